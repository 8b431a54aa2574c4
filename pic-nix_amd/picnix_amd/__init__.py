"""Python binding of libpicnix_amd.so (the MI355X pic-nix hot path) over its C ABI.

The library is required: importing works without it, but creating a
:class:`Simulation` raises if ``libpicnix_amd.so`` is missing, so no code path
can silently fall back to a CPU implementation.

Method names mirror the reference's PicChunk operator surface
(pic/pic_chunk.hpp:90-142); each call acts on every chunk, as one
``omp for`` phase of PicApplication::push_openmp does, or on the one chunk
its ``chunk`` argument names, as a PicChunk method does.
"""
import ctypes
import os

import numpy as np

PKG_DIR = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
LIB_PATH = os.path.join(PKG_DIR, "libpicnix_amd.so")
CONFIG_DIR = os.path.join(PKG_DIR, "configs")

BOUNDARY_EMF, BOUNDARY_CUR, BOUNDARY_MOM, BOUNDARY_PARTICLE = 0, 1, 2, 3
FIELD_UF, FIELD_UJ, FIELD_UM = 0, 1, 3

ERRORS = {-1: "PN_ERR_ARG", -2: "PN_ERR_CONFIG", -3: "PN_ERR_HIP", -4: "PN_ERR_CAPACITY",
          -5: "PN_ERR_UNSUPPORTED", -6: "PN_ERR_STATE"}

# every symbol include/picnix.h declares
EXPORTS = [
    "pn_create", "pn_destroy", "pn_last_error", "pn_num_chunks", "pn_num_species",
    "pn_chunk_info", "pn_get_np", "pn_get_particles", "pn_set_particles",
    "pn_get_particles_old", "pn_get_gindex", "pn_get_pindex", "pn_get_fields",
    "pn_set_fields", "pn_push_bfd", "pn_push_efd", "pn_push_velocity", "pn_push_position",
    "pn_deposit_current", "pn_deposit_moment", "pn_init_friedman", "pn_set_step_graph",
    "pn_sort_particle", "pn_exchange", "pn_step",
    "pn_get_energy", "pn_get_energy_static",
    "pn_get_diverror", "pn_synchronize", "pn_timing", "pn_timing_report",
    "pn_write_field_diag", "pn_write_particle_diag", "pn_write_profile",
    "pn_save_state", "pn_load_state", "pn_get_boundary_state", "pn_autotune",
    "pn_get_tuning",
    # include/picnix_adjoint.h (signatures in picnix_amd.adjoint)
    "pn_adj_create", "pn_adj_destroy", "pn_adj_last_error", "pn_adj_modes", "pn_adj_forward",
    "pn_adj_backward", "pn_adj_set_state", "pn_adj_get_state", "pn_adj_step",
    "pn_adj_num_particles",
]

PHASES = ["bfd", "velocity", "position", "deposit", "fold", "efd", "sort", "rho"]


class PicnixError(RuntimeError):
    def __init__(self, code, msg):
        super().__init__(f"{ERRORS.get(code, code)}: {msg}")
        self.code = code


_lib = None


def load_library(path=LIB_PATH):
    """Load libpicnix_amd.so and declare the C signatures (no GPU work)."""
    global _lib
    if _lib is not None:
        return _lib
    if not os.path.exists(path):
        raise FileNotFoundError(
            f"{path} is missing: build it with `python pic-nix_amd/build.py` "
            "(there is no CPU fallback)")
    L = ctypes.CDLL(path)
    P = ctypes.c_void_p
    i = ctypes.c_int
    d = ctypes.c_double
    i64 = ctypes.c_int64
    dp = ctypes.POINTER(ctypes.c_double)
    ip = ctypes.POINTER(ctypes.c_int)
    i32p = ctypes.POINTER(ctypes.c_int32)
    sig = {
        "pn_create": [ctypes.c_char_p, i, ctypes.POINTER(P)],
        "pn_destroy": [P],
        "pn_num_chunks": [P, ip],
        "pn_num_species": [P, ip],
        "pn_chunk_info": [P, i, ip],
        "pn_get_np": [P, i, i, ctypes.POINTER(i64)],
        "pn_get_particles": [P, i, i, dp, i64, ctypes.POINTER(i64)],
        "pn_set_particles": [P, i, i, dp, i64],
        "pn_get_particles_old": [P, i, i, dp, i64],
        "pn_get_gindex": [P, i, i, i32p, i64],
        "pn_get_pindex": [P, i, i, i32p],
        "pn_get_fields": [P, i, i, dp],
        "pn_set_fields": [P, i, i, dp],
        "pn_push_bfd": [P, i, d],
        "pn_push_efd": [P, i, d],
        "pn_push_velocity": [P, i, d],
        "pn_push_position": [P, i, d],
        "pn_deposit_current": [P, i, d],
        "pn_sort_particle": [P, i],
        "pn_deposit_moment": [P, i],
        "pn_init_friedman": [P],
        "pn_set_step_graph": [P, i],
        "pn_exchange": [P, i],
        "pn_step": [P, d],
        "pn_get_energy": [P, dp, dp, dp],
        "pn_get_energy_static": [P, dp, dp, dp],
        "pn_get_diverror": [P, dp, dp],
        "pn_synchronize": [P],
        "pn_timing": [P, i],
        "pn_timing_report": [P, dp, ip],
        "pn_write_field_diag": [P, ctypes.c_char_p, i, d, i],
        "pn_write_particle_diag": [P, ctypes.c_char_p, i, d, d, i],
        "pn_write_profile": [P, ctypes.c_char_p, i, d],
        "pn_save_state": [P, ctypes.c_char_p, i, d],
        "pn_load_state": [P, ctypes.c_char_p, ip, dp],
        "pn_get_boundary_state": [P, i, ctypes.POINTER(i64)],
        "pn_autotune": [P, i],
        "pn_get_tuning": [P, ip],
    }
    for name, args in sig.items():
        f = getattr(L, name)
        f.argtypes = args
        f.restype = ctypes.c_int
    L.pn_last_error.argtypes = [P]
    L.pn_last_error.restype = ctypes.c_char_p
    _lib = L
    return L


def _dp(a):
    return a.ctypes.data_as(ctypes.POINTER(ctypes.c_double))


class Simulation:
    """All chunks of one pic-nix run, resident on the current HIP device."""

    def __init__(self, toml_text, load=True):
        self.L = load_library()
        h = ctypes.c_void_p()
        rc = self.L.pn_create(toml_text.encode(), 1 if load else 0, ctypes.byref(h))
        self.h = h
        if rc != 0:
            msg = self.L.pn_last_error(h).decode() if h else ""
            if h:
                self.L.pn_destroy(h)
            self.h = None
            raise PicnixError(rc, msg)
        n = ctypes.c_int()
        self._chk(self.L.pn_num_chunks(self.h, ctypes.byref(n)))
        self.nchunks = n.value
        self._chk(self.L.pn_num_species(self.h, ctypes.byref(n)))
        self.Ns = n.value

    @classmethod
    def from_file(cls, path, load=True):
        with open(path) as f:
            return cls(f.read(), load)

    def _chk(self, rc):
        if rc != 0:
            raise PicnixError(rc, self.L.pn_last_error(self.h).decode())

    def close(self):
        if getattr(self, "h", None):
            self.L.pn_destroy(self.h)
            self.h = None

    def __del__(self):
        self.close()

    # ---- geometry ----
    def chunk_info(self, c):
        info = np.zeros(11, dtype=np.int32)
        self._chk(self.L.pn_chunk_info(self.h, c, info.ctypes.data_as(ctypes.POINTER(ctypes.c_int))))
        return info

    def np(self, c, s):
        n = ctypes.c_int64()
        self._chk(self.L.pn_get_np(self.h, c, s, ctypes.byref(n)))
        return n.value

    # ---- particles ----
    def particles(self, c, s):
        n = self.np(c, s)
        out = np.zeros((n, 7), dtype=np.float64)
        got = ctypes.c_int64()
        self._chk(self.L.pn_get_particles(self.h, c, s, _dp(out), n, ctypes.byref(got)))
        return out

    def set_particles(self, c, s, xu):
        xu = np.ascontiguousarray(xu, dtype=np.float64)
        self._chk(self.L.pn_set_particles(self.h, c, s, _dp(xu), xu.shape[0]))

    def particles_old(self, c, s):
        n = self.np(c, s)
        out = np.zeros((n, 3), dtype=np.float64)
        self._chk(self.L.pn_get_particles_old(self.h, c, s, _dp(out), n))
        return out

    def gindex(self, c, s):
        n = self.np(c, s)
        out = np.zeros(n, dtype=np.int32)
        self._chk(self.L.pn_get_gindex(self.h, c, s,
                                       out.ctypes.data_as(ctypes.POINTER(ctypes.c_int32)), n))
        return out

    def pindex(self, c, s):
        ng = int(self.chunk_info(c)[10])
        out = np.zeros(ng + 1, dtype=np.int32)
        self._chk(self.L.pn_get_pindex(self.h, c, s,
                                       out.ctypes.data_as(ctypes.POINTER(ctypes.c_int32))))
        return out

    # ---- fields ----
    def field(self, c, which):
        info = self.chunk_info(c)
        nz, ny, nx = (int(v) for v in info[6:9])
        ncomp = {FIELD_UF: 6, FIELD_UJ: 4, FIELD_UM: 14 * self.Ns}[which]
        out = np.zeros((nz, ny, nx, ncomp), dtype=np.float64)
        self._chk(self.L.pn_get_fields(self.h, c, which, _dp(out)))
        return out

    def set_field(self, c, which, arr):
        arr = np.ascontiguousarray(arr, dtype=np.float64)
        self._chk(self.L.pn_set_fields(self.h, c, which, _dp(arr)))

    # ---- operators (PicChunk names) ----
    def push_bfd(self, dt, chunk=-1):
        self._chk(self.L.pn_push_bfd(self.h, chunk, dt))

    def push_efd(self, dt, chunk=-1):
        self._chk(self.L.pn_push_efd(self.h, chunk, dt))

    def push_velocity(self, dt, chunk=-1):
        self._chk(self.L.pn_push_velocity(self.h, chunk, dt))

    def push_position(self, dt, chunk=-1):
        self._chk(self.L.pn_push_position(self.h, chunk, dt))

    def deposit_current(self, dt, chunk=-1):
        self._chk(self.L.pn_deposit_current(self.h, chunk, dt))

    def deposit_moment(self):
        """PicChunk::deposit_moment; follow with exchange(BOUNDARY_MOM)."""
        self._chk(self.L.pn_deposit_moment(self.h, -1))

    def step_graph(self, enable=True):
        """Replay pn_step as a captured hipGraph (pn_set_step_graph)."""
        self._chk(self.L.pn_set_step_graph(self.h, 1 if enable else 0))

    def init_friedman(self):
        self._chk(self.L.pn_init_friedman(self.h))

    def sort_particle(self, chunk=-1):
        self._chk(self.L.pn_sort_particle(self.h, chunk))

    def exchange(self, mode):
        self._chk(self.L.pn_exchange(self.h, mode))

    def step(self, dt, n=1):
        for _ in range(n):
            self._chk(self.L.pn_step(self.h, dt))

    def synchronize(self):
        self._chk(self.L.pn_synchronize(self.h))

    def timing(self, enable=True):
        self._chk(self.L.pn_timing(self.h, 1 if enable else 0))

    def timing_report(self):
        """{phase: (total_ms, launches)} since the last timing(True)."""
        ms = np.zeros(len(PHASES), dtype=np.float64)
        n = np.zeros(len(PHASES), dtype=np.int32)
        self._chk(self.L.pn_timing_report(self.h, _dp(ms),
                                          n.ctypes.data_as(ctypes.POINTER(ctypes.c_int))))
        return {p: (float(ms[i]), int(n[i])) for i, p in enumerate(PHASES)}

    def energy(self):
        e, b = ctypes.c_double(), ctypes.c_double()
        p = np.zeros(self.Ns, dtype=np.float64)
        self._chk(self.L.pn_get_energy(self.h, ctypes.byref(e), ctypes.byref(b), _dp(p)))
        return e.value, b.value, p

    def energy_static(self):
        """pn_get_energy_static: the same sums with no device allocation."""
        e, b = ctypes.c_double(), ctypes.c_double()
        p = np.zeros(self.Ns, dtype=np.float64)
        self._chk(self.L.pn_get_energy_static(self.h, ctypes.byref(e), ctypes.byref(b), _dp(p)))
        return e.value, b.value, p

    def diverror(self):
        e, b = ctypes.c_double(), ctypes.c_double()
        self._chk(self.L.pn_get_diverror(self.h, ctypes.byref(e), ctypes.byref(b)))
        return e.value, b.value

    def write_field_diag(self, directory, step, time, decimate=1):
        """FieldDiag files <directory>/<step>.data/.json (pn_write_field_diag)."""
        self._chk(self.L.pn_write_field_diag(self.h, str(directory).encode(), int(step),
                                             float(time), int(decimate)))

    def write_particle_diag(self, directory, step, time, fraction=0.01, seed=0):
        """ParticleDiag files <directory>/<step>.data/.json (pn_write_particle_diag)."""
        self._chk(self.L.pn_write_particle_diag(self.h, str(directory).encode(), int(step),
                                                float(time), float(fraction), int(seed)))

    def write_profile(self, path, step=0, time=0.0):
        """The application profile as msgpack (pn_write_profile)."""
        self._chk(self.L.pn_write_profile(self.h, str(path).encode(), int(step), float(time)))

    def save_state(self, prefix, step=0, time=0.0):
        """Checkpoint <prefix>.msgpack + <prefix>.data (pn_save_state)."""
        self._chk(self.L.pn_save_state(self.h, str(prefix).encode(), int(step), float(time)))

    def load_state(self, prefix):
        """Restart from a checkpoint; returns (step, time) (pn_load_state)."""
        st = ctypes.c_int(0)
        tm = ctypes.c_double(0.0)
        self._chk(self.L.pn_load_state(self.h, str(prefix).encode(), ctypes.byref(st),
                                       ctypes.byref(tm)))
        return st.value, tm.value

    def autotune(self, enable=True):
        """Start (or cancel) an autotuning session over the next steps."""
        self._chk(self.L.pn_autotune(self.h, 1 if enable else 0))

    def tuning(self):
        """{state, steps_left, deposit_rows, rho_rows, gather_rows (-1: per particle),
        gather_per_species, deposit_queue, gather_queue, fused_push}."""
        out = (ctypes.c_int * 9)()
        self._chk(self.L.pn_get_tuning(self.h, out))
        keys = ("state", "steps_left", "deposit_rows", "rho_rows", "gather_rows",
                "gather_per_species", "deposit_queue", "gather_queue", "fused_push")
        return dict(zip(keys, list(out)))

    def boundary_state(self, c):
        """{random_seed, lastid, influx[2], efflux[2]} of chunk c (shock problem)."""
        out = (ctypes.c_int64 * 6)()
        self._chk(self.L.pn_get_boundary_state(self.h, c, out))
        return list(out)

    def history(self):
        """[div(E)-rho, div(B), E^2/2, B^2/2, particle energy per species] (history.txt columns)."""
        de, db = self.diverror()
        e, b, p = self.energy()
        return np.concatenate([[de, db, e, b], p])


def config_path(name):
    return os.path.join(CONFIG_DIR, name if name.endswith(".toml") else name + ".toml")
