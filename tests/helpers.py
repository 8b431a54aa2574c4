"""Shared helpers for the test-suite: config construction and comparisons."""
import copy
import os

import numpy as np

import tomlmini

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
CONFIG_DIR = os.path.join(ROOT, "pic-nix_amd", "configs")


def load_config(name):
    return tomlmini.load(os.path.join(CONFIG_DIR, name + ".toml"))


def dumps(cfg):
    """Render the config dict back to TOML text (subset used by pic-nix)."""
    def val(v):
        if isinstance(v, bool):
            return "true" if v else "false"
        if isinstance(v, str):
            return f"'{v}'"
        if isinstance(v, float):
            return repr(v)
        return str(v)

    out = []
    app = cfg.get("application", {})
    out.append("[application]")
    for k, v in app.items():
        if not isinstance(v, dict):
            out.append(f"  {k} = {val(v)}")
    for k, v in app.items():
        if isinstance(v, dict):
            out.append(f"[application.{k}]")
            for k2, v2 in v.items():
                out.append(f"  {k2} = {val(v2)}")
    par = cfg["parameter"]
    out.append("[parameter]")
    for k, v in par.items():
        if k != "particle":
            out.append(f"  {k} = {val(v)}")
    for sp in par.get("particle", []):
        out.append("[[parameter.particle]]")
        for k, v in sp.items():
            out.append(f"  {k} = {val(v)}")
    return "\n".join(out) + "\n"


def small_thermal(N=(16, 16, 16), C=(2, 2, 2), np_=8, order=2, **opt):
    cfg = copy.deepcopy(load_config("thermal64"))
    p = cfg["parameter"]
    p.update(Nz=N[0], Ny=N[1], Nx=N[2], Cz=C[0], Cy=C[1], Cx=C[2])
    for s in p["particle"]:
        s["np"] = np_
    cfg["application"]["option"]["order"] = order
    cfg["application"]["option"].update(opt)
    return cfg


def small_twostream(N=(1, 16, 64), C=(1, 1, 4), np_=16, order=1, **opt):
    cfg = copy.deepcopy(load_config("twostream2d"))
    p = cfg["parameter"]
    p.update(Nz=N[0], Ny=N[1], Nx=N[2], Cz=C[0], Cy=C[1], Cx=C[2])
    for s in p["particle"]:
        s["np"] = np_
    cfg["application"]["option"]["order"] = order
    cfg["application"]["option"].update(opt)
    return cfg


def small_shock(N=(1, 1, 64), C=(1, 1, 4), nppc=8, order=2, **opt):
    """example/shock on a small domain (reflecting wall at x = 0, injection at x = Lx)."""
    cfg = copy.deepcopy(load_config("shock1d"))
    p = cfg["parameter"]
    p.update(Nz=N[0], Ny=N[1], Nx=N[2], Cz=C[0], Cy=C[1], Cx=C[2], nppc=nppc)
    for k in list(opt):
        if k in ("density", "ramp_x", "ramp_w", "u0", "taper", "mime", "delt", "delh", "theta",
                 "phi", "sigma", "betae", "betai"):
            p[k] = opt.pop(k)
    cfg["application"]["option"]["order"] = order
    cfg["application"]["option"].update(opt)
    return cfg


def rel_err(a, b):
    """max |a - b| / max |b| (0 when both vanish)."""
    a = np.asarray(a, dtype=np.float64)
    b = np.asarray(b, dtype=np.float64)
    den = np.max(np.abs(b)) if b.size else 0.0
    num = np.max(np.abs(a - b)) if a.size else 0.0
    if den == 0.0:
        return num
    return num / den


def interior(arr, info, dim):
    """Interior block of a ghost-inclusive chunk array."""
    nb = int(info[9])
    nz, ny, nx = (int(v) for v in info[6:9])
    if dim == 3:
        return arr[nb:nz - nb, nb:ny - nb, nb:nx - nb]
    return arr[nb:nb + 1, nb:ny - nb, nb:nx - nb]
