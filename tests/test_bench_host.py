"""Host logic of bench.py (no GPU): the HBM traffic lookup in the committed
rocprofv3 summaries and the contract's refusal of more than one GPU."""
import importlib.util
import os
import subprocess
import sys

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))


def _bench():
    spec = importlib.util.spec_from_file_location("bench_host", os.path.join(ROOT, "bench.py"))
    mod = importlib.util.module_from_spec(spec)
    spec.loader.exec_module(mod)
    return mod


def test_traffic_takes_the_most_dispatched_form(tmp_path, monkeypatch):
    b = _bench()
    prof = tmp_path / "profiles"
    prof.mkdir()
    (prof / "hbm_traffic_demo.txt").write_text(
        "# header\n"
        "FETCH_SIZE  (copy probe 256.0 MiB, counter 131087 KiB -> factor 2.000)\n"
        "  kernel                                              n   KiB/dispatch  MB/dispatch (cal.)\n"
        "  k_deposit_march<2, false>                          34       995326.5              2038.2\n"
        "  k_deposit_march<2, true>                            4       990000.0              9999.0\n"
        "  k_push_position                                    38       790392.2              1618.5\n"
        "WRITE_SIZE  (copy probe 256.0 MiB, counter 262144 KiB -> factor 1.000)\n"
        "  kernel                                              n   KiB/dispatch  MB/dispatch (cal.)\n"
        "  k_deposit_march<2, true>                            4       200000.0              8888.0\n"
        "  k_deposit_march<2, false>                          34       207496.7               212.5\n")
    monkeypatch.setattr(b, "ROOT", str(tmp_path))
    assert b.profile_traffic("demo", "k_deposit_march") == (2038.2 + 212.5) * 1e6
    # a kernel with only one counter section, and a missing summary: no figure
    assert b.profile_traffic("demo", "k_push_position") is None
    assert b.profile_traffic("absent", "k_deposit_march") is None


def test_committed_summary_has_the_dominant_kernel():
    b = _bench()
    t = b.profile_traffic("thermal64", "k_deposit_march")
    assert t is not None and t > 1e9


def test_more_than_one_gpu_is_refused():
    r = subprocess.run([sys.executable, os.path.join(ROOT, "bench.py"), "--gpus", "2"],
                       capture_output=True, text=True, timeout=120, cwd=ROOT)
    assert r.returncode == 2
    assert "one GPU" in r.stderr


def test_rho_bytes_follow_the_step_form():
    """The periodic workloads carry rho by continuity (a stencil over the grid,
    no particle reads); the shock problem deposits it from the particles."""
    b = _bench()
    assert b.algorithmic_bytes("rho", 10 ** 6, 1000) == 40 * 1000
    assert b.algorithmic_bytes("rho", 10 ** 6, 1000, rho_direct=True) == 24 * 10 ** 6 + 8 * 1000
    # the gather with the position push (124 B) and the deferred record move (+20 B);
    # the deferred sort reads the cell index and writes the source index
    assert b.algorithmic_bytes("velocity", 10 ** 6, 1000, fused=True) == 124 * 10 ** 6 + 48 * 1000
    assert (b.algorithmic_bytes("velocity", 10 ** 6, 1000, fused=True, deferred=True)
            == 144 * 10 ** 6 + 48 * 1000)
    assert b.algorithmic_bytes("sort", 10 ** 6, 1000, deferred=True) == 8 * 10 ** 6 + 12 * 1000
    assert b.PHASE_KERNEL["rho"] == "k_rho_continuity"


def test_host_cpu_record():
    b = _bench()
    h = b.host_cpu()
    assert h["nproc"] >= 1 and h["logical_cpus"] >= h["nproc"]
