"""Build the libtorch autograd extension picnix_torch*.so in-tree (pic-nix_amd/).

Usage: python pic-nix_amd/torch_ext/build.py

A C++ translation unit over torch/extension.h, compiled with g++ against the
image's torch headers and linked with libtorch and libpicnix_amd.so (rpath to
both), so that `import picnix_torch` from pic-nix_amd/ finds it.
"""
import os
import subprocess
import sys
import sysconfig

HERE = os.path.dirname(os.path.abspath(__file__))
PKG = os.path.dirname(HERE)


def build(verbose=True):
    import torch
    from torch.utils import cpp_extension as ce
    suffix = sysconfig.get_config_var("EXT_SUFFIX") or ".so"
    out = os.path.join(PKG, "picnix_torch" + suffix)
    abi = int(torch.compiled_with_cxx11_abi())
    import pybind11
    inc = ce.include_paths() + [sysconfig.get_paths()["include"], pybind11.get_include()]
    libdir = ce.library_paths()[0]
    cmd = ["g++", "-O2", "-std=c++17", "-fPIC", "-shared", f"-D_GLIBCXX_USE_CXX11_ABI={abi}",
           "-DTORCH_EXTENSION_NAME=picnix_torch", "-DTORCH_API_INCLUDE_EXTENSION_H",
           os.path.join(HERE, "pn_torch.cpp"), "-o", out]
    cmd += [f"-I{p}" for p in inc]
    cmd += [f"-L{libdir}", f"-L{PKG}", "-lc10", "-ltorch", "-ltorch_cpu", "-ltorch_python",
            "-lpicnix_amd", f"-Wl,-rpath,{libdir}", "-Wl,-rpath,$ORIGIN"]
    r = subprocess.run(cmd, capture_output=True, text=True)
    if r.returncode != 0:
        raise RuntimeError(f"torch extension build failed: {' '.join(cmd)}\n{r.stdout}\n{r.stderr}")
    if verbose:
        print(f"built {out}")
    return out


if __name__ == "__main__":
    try:
        build()
    except RuntimeError as e:
        print(e, file=sys.stderr)
        sys.exit(1)
