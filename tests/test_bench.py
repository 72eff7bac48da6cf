"""CPU checks of the measurement script's contract (no GPU is touched)."""
import importlib.util
import json
import os
import subprocess
import sys

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))


def _bench():
    spec = importlib.util.spec_from_file_location("bench_mod", os.path.join(ROOT, "bench.py"))
    mod = importlib.util.module_from_spec(spec)
    spec.loader.exec_module(mod)
    return mod


def test_yardstick_bytes_per_element_stage():
    """8*4*Np*(1 + 1 + 2/3) + 112 + 12: U read, U written, step-start state in two of three
    stages, the reference's 14 geometry doubles and 3 neighbour ids; the kernels move the
    same state with 5 geometry doubles (72 B less)."""
    b = _bench()
    assert round(b.algorithmic_bytes_per_element_stage(3)) == 380
    assert round(b.algorithmic_bytes_per_element_stage(6)) == 636
    assert round(b.algorithmic_bytes_per_element_stage(10)) == 977
    for np_ in (3, 6, 10):
        assert abs(b.algorithmic_bytes_per_element_stage(np_) - b.moved_bytes_per_element_stage(np_) - 72) < 1e-9


def test_more_than_one_gpu_is_refused_with_a_message():
    r = subprocess.run([sys.executable, os.path.join(ROOT, "bench.py"), "--gpus", "2"], capture_output=True, text=True,
                       timeout=120)
    assert r.returncode == 2
    assert "one GPU" in r.stderr
    assert r.stdout.strip() == ""  # no JSON line for a configuration that was not measured


def test_entry_points_exist():
    spec = importlib.util.spec_from_file_location("graft_entry", os.path.join(ROOT, "__graft_entry__.py"))
    mod = importlib.util.module_from_spec(spec)
    spec.loader.exec_module(mod)
    assert callable(mod.build) and callable(mod.smoke)
