"""bench.py's host-side contract, on the CPU: the algorithmic byte counts the
roofline divides by (the 96^3 moment-form apply is 234 911 280 bytes: u read
and y written, 3 x 97^3 values each, plus 27 moments per element), the
roofline object's fields, and the refusal of more than one GPU."""
import importlib.util
import os
import subprocess
import sys

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))


def _bench():
    spec = importlib.util.spec_from_file_location("bench_module", os.path.join(ROOT, "bench.py"))
    mod = importlib.util.module_from_spec(spec)
    spec.loader.exec_module(mod)
    return mod


def test_algorithmic_bytes():
    b = _bench()
    N, E = 97 ** 3, 96 ** 3
    assert b.k_moments_bytes(3, N, E) == 234_911_280 == 2 * 8 * 3 * N + 8 * 27 * E
    assert b.k_moments_bytes(3, N, E, b=4) == 234_911_280 // 2
    assert b.k_moments_bytes(2, 33 ** 2, 32 ** 2) == 2 * 8 * 2 * 33 ** 2 + 8 * 9 * 32 ** 2
    P = 16 * E
    assert b.k_apply_bytes(3, N, E, P) == 2 * 8 * 3 * N + 32 * P + 4 * (E + 1)
    assert b.k_apply_bytes(3, N, E, P, b=4) == 2 * 4 * 3 * N + 16 * P + 4 * (E + 1)


def test_roofline_object():
    b = _bench()
    r = b.roofline_obj(234_911_280, 0.0855, "k", traffic=303_648_598)
    assert r["bound"] == "hbm" and r["unit"] == "GB/s" and r["peak"] == 8000.0
    assert abs(r["achieved"] - 234_911_280 / 0.0855e-3 / 1e9) < 1e-9
    assert abs(r["frac"] - r["achieved"] / 8000.0) < 1e-12
    assert r["traffic"] == 303_648_598 and r["algorithmic_bytes"] == 234_911_280


def test_more_than_one_gpu_is_refused():
    p = subprocess.run([sys.executable, os.path.join(ROOT, "bench.py"), "--gpus", "2"], capture_output=True,
                       text=True, timeout=300, cwd=ROOT)
    assert p.returncode != 0
    assert "one GPU only" in (p.stdout + p.stderr)
