"""ctypes binding of the FP64 CPU oracle (oracle/liboracle.so).

Test infrastructure: the oracle is the checker for the HIP path and is only
loaded by tests, ``__graft_entry__.smoke()`` and ``bench.py``'s CPU baseline.
"""
import ctypes
import os

import numpy as np

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
LIB_PATH = os.path.join(ROOT, "oracle", "liboracle.so")

MAXS = 8


class Params(ctypes.Structure):
    _fields_ = [
        ("Nx", ctypes.c_int), ("Ny", ctypes.c_int), ("Nz", ctypes.c_int),
        ("Cx", ctypes.c_int), ("Cy", ctypes.c_int), ("Cz", ctypes.c_int),
        ("Ns", ctypes.c_int), ("order", ctypes.c_int), ("pusher", ctypes.c_int),
        ("load", ctypes.c_int), ("loader", ctypes.c_int), ("seed_offset", ctypes.c_int),
        ("cc", ctypes.c_double), ("delt", ctypes.c_double), ("delh", ctypes.c_double),
        ("friedman", ctypes.c_double), ("buffer_ratio", ctypes.c_double),
        ("Ex", ctypes.c_double), ("Ey", ctypes.c_double), ("Ez", ctypes.c_double),
        ("Bx", ctypes.c_double), ("By", ctypes.c_double), ("Bz", ctypes.c_double),
        ("np", ctypes.c_int * MAXS),
        ("qm", ctypes.c_double * MAXS), ("ro", ctypes.c_double * MAXS),
        ("vt", ctypes.c_double * MAXS), ("vx", ctypes.c_double * MAXS),
        ("vy", ctypes.c_double * MAXS), ("vz", ctypes.c_double * MAXS),
        ("interp", ctypes.c_int),
        ("problem", ctypes.c_int), ("emf_order", ctypes.c_int), ("nppc", ctypes.c_int),
        ("wp", ctypes.c_double), ("mime", ctypes.c_double), ("u0", ctypes.c_double),
        ("theta", ctypes.c_double), ("phi", ctypes.c_double), ("sigma", ctypes.c_double),
        ("betae", ctypes.c_double), ("betai", ctypes.c_double), ("taper", ctypes.c_double),
        ("density", ctypes.c_double), ("ramp_x", ctypes.c_double), ("ramp_w", ctypes.c_double),
    ]


_lib = None


def lib():
    global _lib
    if _lib is None:
        if not os.path.exists(LIB_PATH):
            raise FileNotFoundError(f"{LIB_PATH} is not built (run __graft_entry__.build())")
        L = ctypes.CDLL(LIB_PATH)
        P = ctypes.c_void_p
        d = ctypes.c_double
        i = ctypes.c_int
        i64 = ctypes.c_int64
        dp = ctypes.POINTER(ctypes.c_double)
        ip = ctypes.POINTER(ctypes.c_int)
        L.pno_create.restype = P
        L.pno_create.argtypes = [ctypes.POINTER(Params)]
        L.pno_destroy.argtypes = [P]
        L.pno_num_chunks.argtypes = [P]
        L.pno_chunk_info.argtypes = [P, i, ip]
        L.pno_get_np.restype = i64
        L.pno_get_np.argtypes = [P, i, i]
        for f in ("pno_get_particles", "pno_get_particles_old"):
            getattr(L, f).argtypes = [P, i, i, dp]
        L.pno_set_particles.argtypes = [P, i, i, dp, i64]
        L.pno_get_gindex.argtypes = [P, i, i, ctypes.POINTER(ctypes.c_int32)]
        L.pno_get_pindex.argtypes = [P, i, i, ctypes.POINTER(ctypes.c_int32)]
        L.pno_get_field.argtypes = [P, i, i, dp]
        L.pno_set_field.argtypes = [P, i, i, dp]
        for f in ("pno_push_bfd", "pno_push_efd", "pno_push_velocity", "pno_push_position",
                  "pno_deposit_current"):
            getattr(L, f).argtypes = [P, d]
        L.pno_sort_particle.argtypes = [P]
        L.pno_exchange.argtypes = [P, i]
        L.pno_step.argtypes = [P]
        L.pno_init_friedman.argtypes = [P]
        L.pno_deposit_moment.argtypes = [P]
        L.pno_history.argtypes = [P, dp]
        L.pno_set_threads.argtypes = [i]
        L.pno_shape_mc.argtypes = [i, d, d, d, dp]
        L.pno_digitize.argtypes = [d, d, d]
        L.pno_lorentz_factor.restype = d
        L.pno_lorentz_factor.argtypes = [d, d, d, d]
        L.pno_push_boris.argtypes = [dp, dp, d]
        L.pno_push_vay.argtypes = [dp, dp, d]
        L.pno_push_higuera_cary.argtypes = [dp, dp, d]
        L.pno_esirkepov3d.argtypes = [i, dp, dp, d, d, d, d, d, dp, ip]
        L.pno_esirkepov2d.argtypes = [i, dp, dp, d, d, d, d, dp, ip]
        L.pno_esirkepov1d.argtypes = [i, dp, dp, d, d, d, dp, ip]
        L.pno_mt19937_64.restype = ctypes.c_uint64
        L.pno_mt19937_64.argtypes = [ctypes.c_uint64, i64]
        L.pno_uniform.argtypes = [ctypes.c_uint64, i64, dp]
        L.pno_normal.argtypes = [ctypes.c_uint64, i64, dp]
        L.pno_chunk_map.argtypes = [i, i, i, ip]
        L.pno_gamma.argtypes = [d, d, ctypes.c_uint64, i64, dp]
        L.pno_poisson.restype = i
        L.pno_poisson.argtypes = [d, ctypes.c_uint64, i64, ip]
        L.pno_maxwell_juttner.argtypes = [d, d, ctypes.c_uint64, i64, dp]
        L.pno_shock_state.argtypes = [P, i, ctypes.POINTER(ctypes.c_int64)]
        L.pno_adj_tangent.argtypes = [i, i, d, d, d, d, d, i64, dp, dp, i, dp, dp]
        _lib = L
    return _lib


def _dp(a):
    return a.ctypes.data_as(ctypes.POINTER(ctypes.c_double))


def _ip(a):
    return a.ctypes.data_as(ctypes.POINTER(ctypes.c_int))


def make_params(cfg, load=True):
    """Build the oracle parameter block from a parsed config dict (see config.py)."""
    p = Params()
    par = cfg["parameter"]
    app = cfg.get("application", {})
    opt = app.get("option", {})
    shock = app.get("problem", "") == "shock"
    for k in ("Nx", "Ny", "Nz", "Cx", "Cy", "Cz"):
        setattr(p, k, int(par[k]))
    p.Ns = 2 if shock else int(par["Ns"])
    p.density = 1.0
    if shock:
        p.problem = 1
        p.emf_order = int(opt.get("emf_order", 0))
        p.nppc = int(par["nppc"])
        for k in ("wp", "mime", "u0", "theta", "phi", "sigma", "betae", "betai", "taper"):
            setattr(p, k, float(par[k]))
        p.density = float(par.get("density", 1.0))
        p.ramp_x = float(par.get("ramp_x", 0.0))
        p.ramp_w = float(par.get("ramp_w", 1.0))
    p.order = int(opt.get("order", 2))
    p.pusher = {"Boris": 0, "Vay": 1, "HigueraCary": 2}[opt.get("pusher", "Boris")]
    p.load = 1 if load else 0
    p.loader = 0
    p.seed_offset = int(opt.get("seed_offset", 0))
    p.cc = 1.0 if shock else float(par["cc"])
    p.delt = float(par["delt"])
    p.delh = float(par["delh"])
    p.friedman = float(opt.get("friedman", 0.0))
    p.interp = {"MC": 0, "WT": 1}[opt.get("interpolation", "MC")]
    p.buffer_ratio = float(opt.get("buffer_ratio", 0.2))
    for k in ("Ex", "Ey", "Ez", "Bx", "By", "Bz"):
        setattr(p, k, float(par.get(k, 0.0)))
    for s, sp in enumerate([] if shock else par["particle"]):
        p.np[s] = int(sp["np"])
        p.qm[s] = float(sp["qm"])
        p.ro[s] = float(sp["ro"])
        p.vt[s] = float(sp["vt"])
        p.vx[s] = float(sp.get("vx", 0.0))
        p.vy[s] = float(sp.get("vy", 0.0))
        p.vz[s] = float(sp.get("vz", 0.0))
    return p


class Oracle:
    """Thin object wrapper with the PicChunk-like operator names."""

    def __init__(self, cfg, load=True):
        self.L = lib()
        self.params = make_params(cfg, load)
        self.h = self.L.pno_create(ctypes.byref(self.params))
        self.Ns = self.params.Ns
        self.nchunks = self.L.pno_num_chunks(self.h)

    def close(self):
        if self.h:
            self.L.pno_destroy(self.h)
            self.h = None

    def __del__(self):
        self.close()

    def chunk_info(self, c):
        info = np.zeros(11, dtype=np.int32)
        self.L.pno_chunk_info(self.h, c, _ip(info))
        return info

    def np(self, c, s):
        return int(self.L.pno_get_np(self.h, c, s))

    def particles(self, c, s):
        n = self.np(c, s)
        out = np.zeros((n, 7), dtype=np.float64)
        if n:
            self.L.pno_get_particles(self.h, c, s, _dp(out))
        return out

    def particles_old(self, c, s):
        n = self.np(c, s)
        out = np.zeros((n, 7), dtype=np.float64)
        if n:
            self.L.pno_get_particles_old(self.h, c, s, _dp(out))
        return out

    def set_particles(self, c, s, xu):
        xu = np.ascontiguousarray(xu, dtype=np.float64)
        self.L.pno_set_particles(self.h, c, s, _dp(xu), xu.shape[0])

    def gindex(self, c, s):
        out = np.zeros(self.np(c, s), dtype=np.int32)
        if out.size:
            self.L.pno_get_gindex(self.h, c, s, out.ctypes.data_as(ctypes.POINTER(ctypes.c_int32)))
        return out

    def pindex(self, c, s):
        ng = int(self.chunk_info(c)[10])
        out = np.zeros(ng + 1, dtype=np.int32)
        self.L.pno_get_pindex(self.h, c, s, out.ctypes.data_as(ctypes.POINTER(ctypes.c_int32)))
        return out

    def shock_state(self, c):
        out = (ctypes.c_int64 * 6)()
        self.L.pno_shock_state(self.h, c, out)
        return list(out)

    def field(self, c, which):
        info = self.chunk_info(c)
        nz, ny, nx = (int(v) for v in info[6:9])
        ncomp = {0: 6, 1: 4, 2: 9, 3: 14 * self.Ns}[which]
        out = np.zeros((nz, ny, nx, ncomp), dtype=np.float64)
        self.L.pno_get_field(self.h, c, which, _dp(out))
        return out

    def set_field(self, c, which, arr):
        arr = np.ascontiguousarray(arr, dtype=np.float64)
        self.L.pno_set_field(self.h, c, which, _dp(arr))

    def push_bfd(self, dt):
        self.L.pno_push_bfd(self.h, dt)

    def push_efd(self, dt):
        self.L.pno_push_efd(self.h, dt)

    def push_velocity(self, dt):
        self.L.pno_push_velocity(self.h, dt)

    def push_position(self, dt):
        self.L.pno_push_position(self.h, dt)

    def deposit_current(self, dt):
        self.L.pno_deposit_current(self.h, dt)

    def init_friedman(self):
        self.L.pno_init_friedman(self.h)

    def deposit_moment(self):
        """Moments + Mom exchange (PicChunk::deposit_moment, set_boundary(Mom))."""
        self.L.pno_deposit_moment(self.h)

    def sort_particle(self):
        self.L.pno_sort_particle(self.h)

    def exchange(self, mode):
        self.L.pno_exchange(self.h, mode)

    def step(self, n=1):
        for _ in range(n):
            self.L.pno_step(self.h)

    def history(self):
        out = np.zeros(self.Ns + 4, dtype=np.float64)
        self.L.pno_history(self.h, _dp(out))
        return out


def set_threads(n):
    lib().pno_set_threads(int(n))


def adj_tangent(nx, ny, delh, delt, cc, q, m, xu6, dxu6, nsteps):
    """L and dL of the adjoint form's forward along dxu6 (tangent-linear, CPU)."""
    L = lib()
    xu6 = np.ascontiguousarray(xu6, dtype=np.float64)
    dxu6 = np.ascontiguousarray(dxu6, dtype=np.float64)
    loss = ctypes.c_double()
    dloss = ctypes.c_double()
    L.pno_adj_tangent(nx, ny, delh, delt, cc, q, m, xu6.shape[0], _dp(xu6), _dp(dxu6), nsteps,
                      ctypes.byref(loss), ctypes.byref(dloss))
    return loss.value, dloss.value
