// pn_config.cpp -- minimal TOML reader + the reference's option normalisation.
#include "pn_config.hpp"

#include <cstdlib>
#include <sstream>

namespace picnix {

namespace {

std::string trim(const std::string& s)
{
  size_t a = s.find_first_not_of(" \t\r\n");
  if (a == std::string::npos)
    return "";
  size_t b = s.find_last_not_of(" \t\r\n");
  return s.substr(a, b - a + 1);
}

std::string strip_comment(const std::string& line)
{
  char        quote = 0;
  std::string out;
  for (char ch : line) {
    if (quote) {
      if (ch == quote)
        quote = 0;
    } else if (ch == '\'' || ch == '"') {
      quote = ch;
    } else if (ch == '#') {
      break;
    }
    out.push_back(ch);
  }
  return trim(out);
}

} // namespace

bool parse_toml(const std::string& text, TomlMap& out, std::string& err)
{
  std::istringstream          in(text);
  std::string                 line;
  std::string                 prefix;
  std::map<std::string, int>  array_count;
  int                         lineno = 0;
  while (std::getline(in, line)) {
    lineno++;
    line = strip_comment(line);
    if (line.empty())
      continue;
    if (line.rfind("[[", 0) == 0) {
      size_t e = line.find("]]");
      if (e == std::string::npos) {
        err = "toml: unterminated [[ at line " + std::to_string(lineno);
        return false;
      }
      std::string name = trim(line.substr(2, e - 2));
      int         k    = array_count[name]++;
      prefix           = name + "[" + std::to_string(k) + "].";
      out[name + ".#count"] = std::to_string(k + 1);
    } else if (line[0] == '[') {
      size_t e = line.find(']');
      if (e == std::string::npos) {
        err = "toml: unterminated [ at line " + std::to_string(lineno);
        return false;
      }
      std::string name = trim(line.substr(1, e - 1));
      // a table nested under the latest element of an array of tables
      for (auto& ac : array_count) {
        const std::string& an = ac.first;
        if (name.rfind(an + ".", 0) == 0) {
          name = an + "[" + std::to_string(ac.second - 1) + "]" + name.substr(an.size());
          break;
        }
      }
      prefix = name + ".";
    } else {
      size_t eq = line.find('=');
      if (eq == std::string::npos) {
        err = "toml: expected key = value at line " + std::to_string(lineno);
        return false;
      }
      std::string key = trim(line.substr(0, eq));
      std::string val = trim(line.substr(eq + 1));
      if (val.size() >= 2 && (val[0] == '\'' || val[0] == '"'))
        val = val.substr(1, val.size() - 2);
      out[prefix + key] = val;
    }
  }
  return true;
}

namespace {

bool get_num(const TomlMap& kv, const std::string& k, double& v, bool required, std::string& err)
{
  auto it = kv.find(k);
  if (it == kv.end()) {
    if (required) {
      err = "config: missing mandatory key '" + k + "'";
      return false;
    }
    return true;
  }
  char* end = nullptr;
  v         = std::strtod(it->second.c_str(), &end);
  if (end == it->second.c_str()) {
    err = "config: key '" + k + "' is not a number";
    return false;
  }
  return true;
}

bool get_int(const TomlMap& kv, const std::string& k, int& v, bool required, std::string& err)
{
  double d = v;
  if (!get_num(kv, k, d, required, err))
    return false;
  v = static_cast<int>(d);
  return true;
}

std::string get_str(const TomlMap& kv, const std::string& k, const std::string& def)
{
  auto it = kv.find(k);
  return it == kv.end() ? def : it->second;
}

} // namespace

bool build_config(const TomlMap& kv, Config& cfg, std::string& err)
{
  // nix/cfgparser.hpp:145-217: mandatory keys and N % C == 0
  const char* ints[6] = {"Nz", "Ny", "Nx", "Cz", "Cy", "Cx"};
  int         v[6];
  for (int i = 0; i < 6; i++) {
    v[i] = 1;
    if (!get_int(kv, std::string("parameter.") + ints[i], v[i], true, err))
      return false;
  }
  for (int d = 0; d < 3; d++) {
    cfg.N[d] = v[d];
    cfg.C[d] = v[3 + d];
    if (cfg.N[d] <= 0 || cfg.C[d] <= 0 || cfg.N[d] % cfg.C[d] != 0) {
      err = "config: N must be a positive multiple of C in every dimension";
      return false;
    }
  }
  const std::string problem = get_str(kv, "application.problem", "");
  if (problem == "shock") {
    cfg.problem = 1;
  } else if (!problem.empty()) {
    err = "config: unknown [application] problem: " + problem;
    return false;
  }
  if (cfg.problem == 1) {
    // MainChunk::setup fixes cc = 1 and Ns = 2 (example/shock/main.cpp:60-61)
    cfg.cc = 1.0;
    cfg.Ns = 2;
    ShockConfig& sh = cfg.shock;
    if (!get_num(kv, "parameter.delt", cfg.delt, true, err) ||
        !get_num(kv, "parameter.delh", cfg.delh, true, err) ||
        !get_int(kv, "parameter.nppc", sh.nppc, true, err) ||
        !get_num(kv, "parameter.wp", sh.wp, true, err) ||
        !get_num(kv, "parameter.mime", sh.mime, true, err) ||
        !get_num(kv, "parameter.u0", sh.u0, true, err) ||
        !get_num(kv, "parameter.theta", sh.theta, true, err) ||
        !get_num(kv, "parameter.phi", sh.phi, true, err) ||
        !get_num(kv, "parameter.sigma", sh.sigma, true, err) ||
        !get_num(kv, "parameter.betae", sh.betae, true, err) ||
        !get_num(kv, "parameter.betai", sh.betai, true, err) ||
        !get_num(kv, "parameter.taper", sh.taper, true, err) ||
        !get_num(kv, "parameter.density", sh.density, false, err) ||
        !get_num(kv, "parameter.ramp_x", sh.ramp_x, false, err) ||
        !get_num(kv, "parameter.ramp_w", sh.ramp_w, false, err))
      return false;
    if (sh.nppc <= 0 || sh.density <= 0 || sh.ramp_w <= 0) {
      err = "config: shock needs nppc > 0, density > 0 and ramp_w > 0";
      return false;
    }
  } else if (!get_num(kv, "parameter.delt", cfg.delt, true, err) ||
             !get_num(kv, "parameter.delh", cfg.delh, true, err) ||
             !get_num(kv, "parameter.cc", cfg.cc, true, err) ||
             !get_int(kv, "parameter.Ns", cfg.Ns, true, err))
    return false;
  const char* eb[6] = {"Ex", "Ey", "Ez", "Bx", "By", "Bz"};
  for (int i = 0; i < 6; i++) {
    double x = 0;
    if (!get_num(kv, std::string("parameter.") + eb[i], x, false, err))
      return false;
    if (i < 3)
      cfg.E0[i] = x;
    else
      cfg.B0[i - 3] = x;
  }
  int npart = 0;
  {
    auto it = kv.find("parameter.particle.#count");
    if (it != kv.end())
      npart = std::atoi(it->second.c_str());
  }
  if (cfg.problem == 0 && npart < cfg.Ns) {
    err = "config: fewer [[parameter.particle]] tables than Ns";
    return false;
  }
  cfg.species.resize(cfg.Ns);
  for (int s = 0; s < cfg.Ns && cfg.problem == 0; s++) {
    std::string    p  = "parameter.particle[" + std::to_string(s) + "].";
    SpeciesConfig& sp = cfg.species[s];
    if (!get_int(kv, p + "np", sp.np, true, err) || !get_num(kv, p + "qm", sp.qm, true, err) ||
        !get_num(kv, p + "ro", sp.ro, true, err) || !get_num(kv, p + "vt", sp.vt, false, err) ||
        !get_num(kv, p + "vx", sp.vx, false, err) || !get_num(kv, p + "vy", sp.vy, false, err) ||
        !get_num(kv, p + "vz", sp.vz, false, err))
      return false;
  }

  // [application.option], pic/pic_chunk.cpp:135-262
  const std::string o = "application.option.";
  if (!get_int(kv, o + "order", cfg.order, false, err))
    return false;
  if (cfg.order < 1 || cfg.order > 4) {
    err = "config: order must be 1, 2, 3 or 4";
    return false;
  }
  std::string pusher = get_str(kv, o + "pusher", "Boris");
  if (pusher == "Boris")
    cfg.pusher = 0;
  else if (pusher == "Vay")
    cfg.pusher = 1;
  else if (pusher == "HigueraCary")
    cfg.pusher = 2;
  else {
    err = "config: invalid pusher: " + pusher;
    return false;
  }
  std::string interp = get_str(kv, o + "interpolation", "MC");
  if (interp == "MC")
    cfg.interp = 0;
  else if (interp == "WT")
    cfg.interp = 1;
  else {
    err = "config: invalid interpolation: " + interp;
    return false;
  }
  cfg.seed_type = get_str(kv, o + "seed_type", "random");
  if (cfg.seed_type != "random" && cfg.seed_type != "fixed") {
    err = "config: invalid seed type: " + cfg.seed_type;
    return false;
  }
  if (!get_int(kv, o + "seed_offset", cfg.seed_offset, false, err) ||
      !get_num(kv, o + "friedman", cfg.friedman, false, err) ||
      !get_num(kv, o + "buffer_ratio", cfg.buffer_ratio, false, err))
    return false;
  cfg.precision = get_str(kv, o + "precision", "fp64");
  if (cfg.precision != "fp64" && cfg.precision != "fp32") {
    err = "config: precision must be 'fp64' or 'fp32'";
    return false;
  }
  // dimension check (nix/application.cpp:263-285): 1D, 2D (no z) or 3D
  bool has[3];
  for (int d = 0; d < 3; d++)
    has[d] = !(cfg.N[d] == 1 && cfg.C[d] == 1);
  bool is_1d = !has[0] && !has[1] && has[2];
  bool is_2d = !has[0] && has[1] && has[2];
  bool is_3d = has[0] && has[1] && has[2];
  if (!is_1d && !is_2d && !is_3d) {
    err = "config: the domain must be 1D (Nz = Ny = 1), 2D (Nz = 1) or 3D";
    return false;
  }
  // minimum chunk width (pic/pic_chunk.cpp:192-210)
  int nb = (cfg.order + 3) / 2;
  for (int d = 0; d < 3; d++) {
    if (has[d] && cfg.N[d] / cfg.C[d] < nb) {
      err = "config: number of grid points per chunk smaller than the ghost margin";
      return false;
    }
  }
  return true;
}

} // namespace picnix
