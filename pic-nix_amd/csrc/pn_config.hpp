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
