// picnix_run.cpp -- run a pic-nix config.toml on one MI355X.
//
// The driver loop of nix::Application::main (nix/application.cpp:50-87):
// while curtime < tmax + delt { diagnostic; push; curstep++; curtime += delt },
// with the command-line options -c (config), -t (tmax), -e (emax, seconds) of
// nix::ArgParser (nix/argparser.hpp:18-23), plus -l (load) and -s (save): a
// restart from <basedir>/<load> and a checkpoint to <basedir>/<save> at the
// end (Application::setup_chunks / finalize, nix/application.cpp:113-120,
// 311-318).  The 'history' diagnostic writes
// <basedir>/history.txt in the format of HistoryDiag (pic/diag/history.hpp:
// 75-111); 'field' and 'particle' write FieldDiag / ParticleDiag snapshots
// under <basedir>/<prefix>/ (pic/diag/field.hpp, pic/diag/particle.hpp); the
// application profile goes to <basedir>/profile.msgpack before the first step
// (nix/application.cpp:50-60,409-414).  The per-step elapsed time of the push
// is printed like the reference's "push" log record
// (pic/pic_application.cpp:196-216).
#include <chrono>
#include <cstdio>
#include <cstdlib>
#include <filesystem>
#include <fstream>
#include <iostream>
#include <limits>
#include <sstream>
#include <string>
#include <vector>

#include "../../include/picnix.hpp"
#include "../csrc/pn_config.hpp"

namespace {

struct Diag {
  std::string name, prefix;
  int         interval = 1, begin = 0, end = std::numeric_limits<int>::max();
  int         decimate = 1;    // FieldDiag default (pic/diag/field.hpp:72)
  double      fraction = 0.01; // ParticleDiag default (pic/diag/particle.hpp:87)
  bool        due(int step) const
  {
    return step >= begin && step <= end && (step - begin) % interval == 0;
  }
};

std::string fmt(const char* f, double v)
{
  char buf[64];
  std::snprintf(buf, sizeof(buf), f, v);
  return buf;
}

} // namespace

int main(int argc, char** argv)
{
  std::string config = "config.toml";
  double      tmax   = std::numeric_limits<double>::max();
  double      emax   = 3600.0;
  std::string load, save;
  for (int i = 1; i + 1 < argc; i += 2) {
    std::string k = argv[i];
    if (k == "-c" || k == "--config")
      config = argv[i + 1];
    else if (k == "-t" || k == "--tmax")
      tmax = std::atof(argv[i + 1]);
    else if (k == "-e" || k == "--emax")
      emax = std::atof(argv[i + 1]);
    else if (k == "-l" || k == "--load")
      load = argv[i + 1];
    else if (k == "-s" || k == "--save")
      save = argv[i + 1];
    else {
      std::cerr << "usage: picnix_run -c config.toml [-t tmax] [-e emax] [-l load] [-s save]\n";
      return 2;
    }
  }
  std::ifstream in(config);
  if (!in) {
    std::cerr << "cannot open " << config << "\n";
    return 2;
  }
  std::stringstream ss;
  ss << in.rdbuf();
  std::string text = ss.str();

  picnix::TomlMap kv;
  std::string     err;
  if (!picnix::parse_toml(text, kv, err)) {
    std::cerr << err << "\n";
    return 2;
  }
  double delt = std::atof(kv["parameter.delt"].c_str());
  std::string basedir = kv.count("application.basedir") ? kv["application.basedir"] : "";
  if (const char* tmp = std::getenv("PICNIX_TMPDIR")) // pic/pic_application.cpp:159-169
    basedir = (std::filesystem::path(tmp) / basedir).string();
  std::vector<Diag> diags;
  int ndiag = kv.count("diagnostic.#count") ? std::atoi(kv["diagnostic.#count"].c_str()) : 0;
  for (int i = 0; i < ndiag; i++) {
    std::string p = "diagnostic[" + std::to_string(i) + "].";
    Diag        d;
    d.name = kv[p + "name"];
    if (kv.count(p + "interval"))
      d.interval = std::atoi(kv[p + "interval"].c_str());
    if (kv.count(p + "begin"))
      d.begin = std::atoi(kv[p + "begin"].c_str());
    if (kv.count(p + "end"))
      d.end = std::atoi(kv[p + "end"].c_str());
    d.prefix = kv.count(p + "prefix") ? kv[p + "prefix"] : d.name;
    if (kv.count(p + "decimate"))
      d.decimate = std::atoi(kv[p + "decimate"].c_str());
    if (kv.count(p + "fraction"))
      d.fraction = std::atof(kv[p + "fraction"].c_str());
    diags.push_back(d);
  }

  try {
    picnix::Simulation sim(text, load.empty());
    std::filesystem::path hist = std::filesystem::path(basedir) / "history.txt";
    if (!basedir.empty())
      std::filesystem::create_directories(basedir);
    const int ns      = sim.num_species();
    auto      t_start = std::chrono::steady_clock::now();
    int       step    = 0;
    double    time    = 0.0;
    namespace fs      = std::filesystem;
    if (!load.empty())
      sim.load_state((fs::path(basedir) / load).string(), step, time);
    else // Application::save_profile, initial runs only (nix/application.cpp:409-419)
      sim.write_profile((fs::path(basedir) / "profile.msgpack").string(), step, time);
    while (time < tmax + delt) {
      for (const Diag& d : diags) {
        if (!d.due(step))
          continue;
        const std::string dir = (std::filesystem::path(basedir) / d.prefix).string();
        if (d.name == "field")
          sim.write_field_diag(dir, step, time, d.decimate);
        if (d.name == "particle")
          sim.write_particle_diag(dir, step, time, d.fraction, 0);
        if (d.name != "history")
          continue;
        double              de, db, ee, eb;
        std::vector<double> ep;
        sim.get_diverror(de, db);
        sim.get_energy(ee, eb, ep);
        std::string msg;
        if (step == d.begin) {
          msg += "# " + std::string(8 - 4, ' ') + "step" + std::string(13 - 4 + 1, ' ') + "time";
          for (const char* h : {"div(E)", "div(B)", "E^2/2", "B^2/2"})
            msg += " " + std::string(13 - std::string(h).size(), ' ') + h;
          for (int s = 0; s < ns; s++) {
            char lab[32];
            std::snprintf(lab, sizeof(lab), "Particle #%02d", s);
            msg += " " + std::string(13 - std::string(lab).size(), ' ') + lab;
          }
          msg += "\n";
          std::filesystem::remove(hist);
        }
        char head[64];
        std::snprintf(head, sizeof(head), "  %08d %13.6e", step, time);
        msg += head;
        for (double v : {de, db, ee, eb})
          msg += fmt(" %13.6e", v);
        for (double v : ep)
          msg += fmt(" %13.6e", v);
        msg += "\n";
        std::cout << msg << std::flush;
        std::ofstream(hist, std::ios::app) << msg;
      }
      auto t0 = std::chrono::steady_clock::now();
      sim.push(delt);
      sim.synchronize();
      auto t1 = std::chrono::steady_clock::now();
      double el = std::chrono::duration<double>(t1 - t0).count();
      if (std::getenv("PICNIX_VERBOSE"))
        std::fprintf(stderr, "step %08d push elapsed %.6e s\n", step, el);
      step++;
      time += delt;
      if (std::chrono::duration<double>(t1 - t_start).count() > emax)
        break;
    }
    if (!save.empty())
      sim.save_state((fs::path(basedir) / save).string(), step, time);
  } catch (const picnix::Error& e) {
    std::cerr << "picnix error " << e.code() << ": " << e.what() << "\n";
    return 1;
  }
  return 0;
}
