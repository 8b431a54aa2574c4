"""Adjoint form of the PIC step (include/picnix_adjoint.h) from Python.

    python -m picnix_amd.adjoint --check [--steps 200 --every 20 --nx 256 --nppc 32]

`AdjointForm` drives the C ABI; `weibel_loss` goes through the libtorch
autograd function (torch::autograd::Function in pic-nix_amd/torch_ext/pn_torch.cpp)
whose forward runs the steps (checkpointing every K) and whose backward runs the
hand-written adjoint kernels, so that

    theta = torch.tensor([6.0] + [1e-3] * 8, dtype=torch.float64, requires_grad=True)
    loss = weibel_loss(theta, form, 200, 20)
    loss.backward()          # theta.grad = dL/dtheta from the adjoint

`WeibelLoss` is the same loss as a Python torch.autograd.Function.

The default set-up is the 2D Weibel case: 256 x 256 periodic cells, 32
electrons per cell with vth_y = vth_z = 0.05c and vth_x = theta[0] * 0.05c over
immobile ions, eight seeded density modes of amplitude theta[1:], order-1
shapes, FP64, and L = the magnetic energy after 200 steps.  `--check` compares
the adjoint gradient with central finite differences of the same forward.
"""
import argparse
import ctypes
import json
import os
import sys
import time

import numpy as np

from . import LIB_PATH, load_library

MAX_MODES = 16


class Params(ctypes.Structure):
    _fields_ = [("Nx", ctypes.c_int), ("Ny", ctypes.c_int),
                ("delh", ctypes.c_double), ("delt", ctypes.c_double), ("cc", ctypes.c_double),
                ("nppc", ctypes.c_int), ("qm", ctypes.c_double), ("ro", ctypes.c_double),
                ("vth", ctypes.c_double), ("nmodes", ctypes.c_int), ("kmax", ctypes.c_int),
                ("seed", ctypes.c_uint64), ("precision", ctypes.c_int)]


def _lib():
    L = load_library(LIB_PATH)
    if getattr(L, "_adj_ready", False):
        return L
    P = ctypes.c_void_p
    dp = ctypes.POINTER(ctypes.c_double)
    sig = {
        "pn_adj_create": [ctypes.POINTER(Params), ctypes.POINTER(P)],
        "pn_adj_destroy": [P],
        "pn_adj_modes": [P, dp],
        "pn_adj_forward": [P, dp, ctypes.c_int, ctypes.c_int, dp],
        "pn_adj_backward": [P, dp],
        "pn_adj_set_state": [P, dp, ctypes.c_int64, dp, dp],
        "pn_adj_get_state": [P, dp, dp, dp],
        "pn_adj_step": [P, ctypes.c_int],
        "pn_adj_num_particles": [P, ctypes.POINTER(ctypes.c_int64)],
    }
    for name, args in sig.items():
        f = getattr(L, name)
        f.argtypes = args
        f.restype = ctypes.c_int
    L.pn_adj_last_error.argtypes = [P]
    L.pn_adj_last_error.restype = ctypes.c_char_p
    L._adj_ready = True
    return L


def _dp(a):
    return a.ctypes.data_as(ctypes.POINTER(ctypes.c_double))


class AdjointForm:
    """One adjoint-form set-up (seeded base particles and modes) on the GPU.
    precision: "fp64" or "fp32" (the particles, fields and co-states; theta, the
    loss and the gradient stay FP64)."""

    def __init__(self, nx=256, ny=256, delh=1.0, delt=0.5, cc=1.0, nppc=32, qm=-1.0, ro=1.0,
                 vth=0.05, nmodes=8, kmax=4, seed=12345, precision="fp64"):
        if precision not in ("fp64", "fp32"):
            raise ValueError("precision must be 'fp64' or 'fp32'")
        self.L = _lib()
        self.p = Params(nx, ny, delh, delt, cc, nppc, qm, ro, vth, nmodes, kmax, seed,
                        1 if precision == "fp32" else 0)
        self.precision = precision
        self.h = ctypes.c_void_p()
        rc = self.L.pn_adj_create(ctypes.byref(self.p), ctypes.byref(self.h))
        self._chk(rc)
        self.nmodes = nmodes
        self.nx, self.ny = nx, ny
        self.tape = 0

    def _chk(self, rc):
        if rc != 0:
            msg = self.L.pn_adj_last_error(self.h).decode() if self.h else ""
            raise RuntimeError(f"adjoint form error {rc}: {msg}")

    def close(self):
        if self.h:
            self.L.pn_adj_destroy(self.h)
            self.h = None

    def __del__(self):
        self.close()

    def modes(self):
        out = np.zeros(3 * self.nmodes)
        self._chk(self.L.pn_adj_modes(self.h, _dp(out)))
        return out.reshape(-1, 3)

    def forward(self, theta, nsteps, every):
        th = np.ascontiguousarray(theta, dtype=np.float64)
        loss = ctypes.c_double()
        self._chk(self.L.pn_adj_forward(self.h, _dp(th), int(nsteps), int(every),
                                        ctypes.byref(loss)))
        self.tape += 1  # the checkpoints pn_adj_backward differentiates
        return loss.value

    def backward(self):
        g = np.zeros(1 + self.nmodes)
        self._chk(self.L.pn_adj_backward(self.h, _dp(g)))
        return g

    def num_particles(self):
        n = ctypes.c_int64()
        self._chk(self.L.pn_adj_num_particles(self.h, ctypes.byref(n)))
        return n.value

    def set_state(self, xu6, E=None, B=None):
        xu6 = np.ascontiguousarray(xu6, dtype=np.float64)
        e = None if E is None else _dp(np.ascontiguousarray(E, dtype=np.float64))
        b = None if B is None else _dp(np.ascontiguousarray(B, dtype=np.float64))
        self._chk(self.L.pn_adj_set_state(self.h, _dp(xu6), xu6.shape[0], e, b))

    def get_state(self):
        n = self.num_particles()
        xu = np.zeros((n, 6))
        E = np.zeros((3, self.ny, self.nx))
        B = np.zeros((3, self.ny, self.nx))
        self._chk(self.L.pn_adj_get_state(self.h, _dp(xu), _dp(E), _dp(B)))
        return xu, E, B

    def step(self, n=1):
        self._chk(self.L.pn_adj_step(self.h, int(n)))


def weibel_loss(theta, form, nsteps, every):
    """L(theta) through the libtorch autograd function of pic-nix_amd/torch_ext
    (picnix_torch*.so, a torch::autograd::Function in C++): forward runs the
    steps with a checkpoint every `every`, backward the adjoint kernels; the
    gradient comes back on theta's device in theta's dtype."""
    import picnix_torch
    return picnix_torch.weibel_loss(theta, int(form.h.value), int(nsteps), int(every))


try:
    import torch

    class WeibelLoss(torch.autograd.Function):
        """The same loss as a Python torch.autograd.Function over the C ABI."""

        @staticmethod
        def forward(ctx, theta, form, nsteps, every):
            ctx.form = form
            loss = form.forward(theta.detach().cpu().numpy(), nsteps, every)
            ctx.tape = form.tape
            ctx.device, ctx.dtype = theta.device, theta.dtype
            return torch.tensor(loss, dtype=theta.dtype, device=theta.device)

        @staticmethod
        def backward(ctx, grad_out):
            if ctx.form.tape != ctx.tape:
                raise RuntimeError("another forward ran on this adjoint form after the one "
                                   "being differentiated; backward needs its own checkpoints")
            g = torch.from_numpy(ctx.form.backward()).to(ctx.device, ctx.dtype)
            return g * grad_out, None, None, None
except ImportError:  # torch is optional for the C ABI itself
    WeibelLoss = None


def check(args):
    form = AdjointForm(nx=args.nx, ny=args.nx, nppc=args.nppc, nmodes=args.modes, seed=args.seed)
    theta = np.array([args.ratio] + [args.eps] * args.modes)
    t0 = time.perf_counter()
    loss = form.forward(theta, args.steps, args.every)
    t1 = time.perf_counter()
    grad = form.backward()
    t2 = time.perf_counter()
    print(f"[adjoint] {args.nx}^2 cells, {form.num_particles()} particles, {args.steps} steps: "
          f"L = {loss:.10e}; forward {t1 - t0:.2f} s, backward {t2 - t1:.2f} s", flush=True)
    rows = []
    # order-1 shapes make L piecewise smooth: the difference quotient is taken at
    # several steps h (relative to max(|theta_j|, 1)); particles crossing grid
    # lines inside [theta - h, theta + h] bend it at the larger ones
    for j in range(len(theta)):
        fds = []
        for hrel in args.h:
            h = hrel * max(abs(theta[j]), 1.0)
            tp, tm = theta.copy(), theta.copy()
            tp[j] += h
            tm[j] -= h
            fd = (form.forward(tp, args.steps, args.every) -
                  form.forward(tm, args.steps, args.every)) / (2 * h)
            fds.append({"h": h, "fd": fd, "rel_error": abs(grad[j] - fd) / max(abs(fd), 1e-300)})
        best = min(fds, key=lambda r: r["rel_error"])
        rows.append({"param": "ratio" if j == 0 else f"eps[{j - 1}]", "adjoint": grad[j],
                     "fd": best["fd"], "h": best["h"], "rel_error": best["rel_error"], "scan": fds})
        print(f"[adjoint] {rows[-1]['param']:>8}: adjoint {grad[j]:+.10e}  FD {best['fd']:+.10e} "
              f"(h {best['h']:.0e})  rel {best['rel_error']:.2e}  [" +
              ", ".join(f"{r['rel_error']:.1e}" for r in fds) + "]", flush=True)
    print(json.dumps({"loss": loss, "steps": args.steps, "every": args.every, "nx": args.nx,
                      "nppc": args.nppc, "modes": form.modes().tolist(), "rows": rows,
                      "forward_s": t1 - t0, "backward_s": t2 - t1}))
    form.close()
    return 0


def main(argv=None):
    ap = argparse.ArgumentParser()
    ap.add_argument("--check", action="store_true")
    ap.add_argument("--nx", type=int, default=256)
    ap.add_argument("--nppc", type=int, default=32)
    ap.add_argument("--steps", type=int, default=200)
    ap.add_argument("--every", type=int, default=20)
    ap.add_argument("--modes", type=int, default=8)
    ap.add_argument("--ratio", type=float, default=6.0)
    ap.add_argument("--eps", type=float, default=1e-3)
    ap.add_argument("--h", type=float, nargs="+", default=[1e-6, 1e-7, 1e-8],
                    help="finite-difference steps, relative to max(|theta_j|, 1)")
    ap.add_argument("--seed", type=int, default=12345)
    args = ap.parse_args(argv)
    if args.check:
        return check(args)
    ap.print_help()
    return 0


if __name__ == "__main__":
    sys.exit(main())
