"""The C-ABI library loads and exports every entry point include/picnix.h declares.

CPU only: nothing here touches a GPU (no pn_create), so it runs in the no-GPU suite.
"""
import os
import re

import pytest

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))


def declared_symbols():
    names = []
    for f in sorted(os.listdir(os.path.join(ROOT, "include"))):
        if not f.endswith(".h"):
            continue
        text = open(os.path.join(ROOT, "include", f)).read()
        names += re.findall(r"^\s*(?:const\s+)?\w[\w\s\*]*?\b(pn_\w+)\s*\(", text, re.M)
    return sorted(set(names))


def test_header_declares_the_abi():
    names = declared_symbols()
    for n in ("pn_create", "pn_destroy", "pn_step", "pn_push_bfd", "pn_push_efd",
              "pn_push_velocity", "pn_push_position", "pn_deposit_current", "pn_sort_particle",
              "pn_exchange", "pn_get_energy", "pn_get_diverror", "pn_last_error"):
        assert n in names


def test_library_exports_every_declared_symbol():
    import picnix_amd
    if not os.path.exists(picnix_amd.LIB_PATH):
        pytest.fail("libpicnix_amd.so is not built (run __graft_entry__.build())")
    L = picnix_amd.load_library()
    missing = [n for n in declared_symbols() if not hasattr(L, n)]
    assert not missing, missing
    # the Python binding declares the same set
    assert sorted(picnix_amd.EXPORTS) == declared_symbols()


def test_missing_library_fails_loudly(tmp_path):
    import picnix_amd
    with pytest.raises(FileNotFoundError):
        picnix_amd._lib_probe = None
        # load from a path that does not exist: no silent fallback
        import ctypes  # noqa: F401
        saved = picnix_amd._lib
        try:
            picnix_amd._lib = None
            picnix_amd.load_library(str(tmp_path / "nope.so"))
        finally:
            picnix_amd._lib = saved


def test_config_rejects_bad_keys():
    """Config validation happens before any device allocation, so it is testable on CPU
    only through the Python reader; the library's checks mirror nix/cfgparser.hpp."""
    import helpers
    cfg = helpers.load_config("thermal64")
    assert cfg["parameter"]["Nx"] % cfg["parameter"]["Cx"] == 0
