"""bench.py's contract pieces that need no GPU: the one-GPU refusal and the
profiled traffic figure it reports beside the roofline."""
import importlib.util
import os
import subprocess
import sys

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))


def _bench():
    spec = importlib.util.spec_from_file_location("bench_mod", os.path.join(ROOT, "bench.py"))
    mod = importlib.util.module_from_spec(spec)
    spec.loader.exec_module(mod)
    return mod


def test_more_than_one_gpu_is_refused():
    r = subprocess.run([sys.executable, os.path.join(ROOT, "bench.py"), "--gpus", "2"], capture_output=True,
                       text=True, timeout=120)
    assert r.returncode == 2
    assert "one GPU only" in r.stderr
    assert r.stdout == ""


def test_profiled_traffic_is_read_from_the_committed_profile():
    b = _bench()
    t = b.profiled_traffic("gm_local_kernel_q24_rmat22.md")
    assert t is not None and 1e10 < t < 1e13
    assert b.profiled_traffic("no_such_profile.md") is None
