// pn_config.hpp -- config.toml reader for the keys pic-nix's hot path consumes.
#pragma once

#include <cstdint>
#include <map>
#include <string>
#include <vector>

namespace picnix {

struct SpeciesConfig {
  int    np = 0; // particles per cell
  double qm = 0, ro = 0, vt = 0, vx = 0, vy = 0, vz = 0;
};

// example/shock ([application] problem = 'shock'): the [parameter] keys of
// example/shock/config.toml read by MainChunk::setup (example/shock/main.cpp:56-92),
// plus an optional density profile n(x): 1 below ramp_x, a linear ramp over ramp_w
// cells, `density` above (density = 1 is the example's uniform plasma).
struct ShockConfig {
  int    nppc = 0;
  double wp = 0, mime = 0, u0 = 0, theta = 0, phi = 0, sigma = 0, betae = 0, betai = 0,
         taper = 0;
  double density = 1.0, ramp_x = 0.0, ramp_w = 1.0;
};

// Parsed [parameter] and [application.option] tables.  Key names, defaults and
// validation follow nix/cfgparser.hpp:145-217 and PicChunk::setup
// (pic/pic_chunk.cpp:135-262).
struct Config {
  int    N[3] = {1, 1, 1}; // Nz, Ny, Nx
  int    C[3] = {1, 1, 1}; // Cz, Cy, Cx
  int    Ns   = 1;
  double cc = 1, delt = 1, delh = 1;
  double E0[3] = {0, 0, 0}, B0[3] = {0, 0, 0};
  std::vector<SpeciesConfig> species;
  int         problem = 0; // 0: species tables (example/thermal, example/beam); 1: shock
  ShockConfig shock;

  int         order        = 2;
  int         pusher       = 0; // Boris
  int         interp       = 0; // MC
  double      friedman     = 0.0;
  double      buffer_ratio = 0.2;
  std::string seed_type    = "random";
  int         seed_offset  = 0;
  std::string precision    = "fp64";
};

// Flat key -> value map of a TOML document ("a.b.c" keys; arrays of tables as
// "a.b[3].c").  Strings keep their text, numbers and booleans their literal.
using TomlMap = std::map<std::string, std::string>;

bool parse_toml(const std::string& text, TomlMap& out, std::string& err);
bool build_config(const TomlMap& kv, Config& cfg, std::string& err);

} // namespace picnix
