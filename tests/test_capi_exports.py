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


BAD_CONFIGS = [
    # (edit of the thermal64 config, text the library's error must contain)
    (lambda c: c["parameter"].update(Nx=65), "multiple of C"),
    (lambda c: c["parameter"].pop("delt"), "missing mandatory key 'parameter.delt'"),
    (lambda c: c["parameter"].update(Ns=3), "fewer [[parameter.particle]] tables than Ns"),
    (lambda c: c["application"]["option"].update(order=5), "order must be 1, 2, 3 or 4"),
    (lambda c: c["application"]["option"].update(pusher="Leapfrog"), "invalid pusher"),
    (lambda c: c["application"]["option"].update(interpolation="XX"), "invalid interpolation"),
    (lambda c: c["application"]["option"].update(seed_type="sometimes"), "invalid seed type"),
    (lambda c: c["application"]["option"].update(precision="fp16"), "precision must be"),
    (lambda c: c["application"].update(problem="vortex"), "unknown [application] problem"),
    (lambda c: c["parameter"].update(Nx=4, Cx=4), "smaller than the ghost margin"),
]


@pytest.mark.parametrize("case", range(len(BAD_CONFIGS)))
def test_config_rejects_bad_keys(case):
    """pn_create validates the configuration before it touches a device
    (nix/cfgparser.hpp:145-217, pic/pic_chunk.cpp:135-262), so every rejected
    key comes back as PN_ERR_CONFIG with its reason on this CPU-only host."""
    import ctypes

    import helpers
    import picnix_amd
    edit, reason = BAD_CONFIGS[case]
    cfg = helpers.load_config("thermal64")
    edit(cfg)
    L = picnix_amd.load_library()
    h = ctypes.c_void_p()
    rc = L.pn_create(helpers.dumps(cfg).encode(), 0, ctypes.byref(h))
    try:
        assert rc == -2, (rc, L.pn_last_error(h))
        assert reason in L.pn_last_error(h).decode()
    finally:
        if h:
            L.pn_destroy(h)


def test_create_rejects_null_arguments():
    import ctypes

    import picnix_amd
    L = picnix_amd.load_library()
    h = ctypes.c_void_p()
    assert L.pn_create(None, 0, ctypes.byref(h)) == -1
    assert L.pn_step(None, 0.5) == -1
    assert L.pn_destroy(None) == -1
