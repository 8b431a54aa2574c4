// picnix.hpp -- C++17 host-side mirror of pic-nix's PicChunk operator surface
// over the C ABI of picnix.h.
//
// The reference drives every chunk through PicChunk's virtual operators
// (pic/pic_chunk.hpp:90-142) from PicApplication::push_openmp
// (pic/pic_application.cpp:219-292).  picnix::Simulation keeps those names;
// each call acts on all chunks of the run at once, on the GPU.  Errors are
// thrown as picnix::Error (the reference logs and calls MPI_Abort,
// pic/pic_chunk.cpp:14-16).
#pragma once

#include <cstdint>
#include <stdexcept>
#include <string>
#include <vector>

#include "picnix.h"

namespace picnix {

class Error : public std::runtime_error {
public:
  Error(int code, const std::string& what) : std::runtime_error(what), code_(code) {}
  int code() const { return code_; }

private:
  int code_;
};

// Boundary modes, pic/pic.hpp:40-46.
enum BoundaryMode {
  BoundaryEmf      = PN_BOUNDARY_EMF,
  BoundaryCur      = PN_BOUNDARY_CUR,
  BoundaryMom      = PN_BOUNDARY_MOM,
  BoundaryParticle = PN_BOUNDARY_PARTICLE,
};

class Simulation {
public:
  // config.toml text; load = false leaves every chunk without particles.
  explicit Simulation(const std::string& toml, bool load = true)
  {
    int rc = pn_create(toml.c_str(), load ? 1 : 0, &sim_);
    if (rc != PN_OK) {
      std::string msg = sim_ ? pn_last_error(sim_) : "pn_create failed";
      if (sim_)
        pn_destroy(sim_);
      sim_ = nullptr;
      throw Error(rc, msg);
    }
    check(pn_num_chunks(sim_, &nchunk_));
    check(pn_num_species(sim_, &ns_));
  }
  ~Simulation()
  {
    if (sim_)
      pn_destroy(sim_);
  }
  Simulation(const Simulation&)            = delete;
  Simulation& operator=(const Simulation&) = delete;

  int num_chunks() const { return nchunk_; }
  int num_species() const { return ns_; }

  // PicChunk operators (pic/pic_chunk.cpp:491-550), all chunks
  void push_bfd(double delt) { check(pn_push_bfd(sim_, -1, delt)); }
  void push_efd(double delt) { check(pn_push_efd(sim_, -1, delt)); }
  void push_velocity(double delt) { check(pn_push_velocity(sim_, -1, delt)); }
  void push_position(double delt) { check(pn_push_position(sim_, -1, delt)); }
  void deposit_current(double delt) { check(pn_deposit_current(sim_, -1, delt)); }
  void sort_particle() { check(pn_sort_particle(sim_, -1)); }
  // set_boundary_{pack,begin,end,unpack}(mode) as one call
  void set_boundary(BoundaryMode mode) { check(pn_exchange(sim_, mode)); }
  // PicApplication::push: one full timestep
  void push(double delt) { check(pn_step(sim_, delt)); }
  // replay push() as a captured hipGraph (no reference counterpart)
  void set_step_graph(bool enable) { check(pn_set_step_graph(sim_, enable ? 1 : 0)); }
  void synchronize() { check(pn_synchronize(sim_)); }

  // FieldDiag / ParticleDiag files and the application profile
  void write_field_diag(const std::string& dir, int step, double time, int decimate = 1)
  {
    check(pn_write_field_diag(sim_, dir.c_str(), step, time, decimate));
  }
  void write_particle_diag(const std::string& dir, int step, double time, double fraction = 0.01,
                           int seed = 0)
  {
    check(pn_write_particle_diag(sim_, dir.c_str(), step, time, fraction, seed));
  }
  void write_profile(const std::string& path, int step = 0, double time = 0.0)
  {
    check(pn_write_profile(sim_, path.c_str(), step, time));
  }
  // StateHandler::save / load
  void save_state(const std::string& prefix, int step, double time)
  {
    check(pn_save_state(sim_, prefix.c_str(), step, time));
  }
  void load_state(const std::string& prefix, int& step, double& time)
  {
    check(pn_load_state(sim_, prefix.c_str(), &step, &time));
  }

  // PicChunk::get_energy / get_diverror, summed over chunks (pic/diag/history.hpp:44-61)
  void get_energy(double& efd, double& bfd, std::vector<double>& particle)
  {
    particle.assign(ns_, 0.0);
    check(pn_get_energy(sim_, &efd, &bfd, particle.data()));
  }
  void get_diverror(double& efd, double& bfd) { check(pn_get_diverror(sim_, &efd, &bfd)); }

  // particle records of one chunk/species, AoS [np][7]
  std::vector<double> particles(int chunk, int species)
  {
    int64_t np = 0;
    check(pn_get_np(sim_, chunk, species, &np));
    std::vector<double> xu(static_cast<size_t>(np) * 7);
    check(pn_get_particles(sim_, chunk, species, xu.data(), np, &np));
    return xu;
  }
  void set_particles(int chunk, int species, const std::vector<double>& xu)
  {
    check(pn_set_particles(sim_, chunk, species, xu.data(), static_cast<int64_t>(xu.size() / 7)));
  }

  pn_sim* handle() { return sim_; }

private:
  void check(int rc)
  {
    if (rc != PN_OK)
      throw Error(rc, pn_last_error(sim_));
  }

  pn_sim* sim_    = nullptr;
  int     nchunk_ = 0;
  int     ns_     = 0;
};

} // namespace picnix
