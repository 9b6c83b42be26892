"""A plain C program calling the C ABI (tests/c_abi/sinker_c1.c, the 2D
sinker of the reference's examples) is compiled with gcc against the
installed headers and the built library.  On the CPU it must see the library
refuse to run without a device; on the GPU it solves and advects, and its
result equals the same workload driven through the Python mirror."""
import os
import re
import subprocess

import numpy as np
import pytest

from conftest import ROOT

SRC = os.path.join(ROOT, "tests", "c_abi", "sinker_c1.c")
LIBDIR = os.path.join(ROOT, "gale_amd")


def _build(tmp_path):
    exe = str(tmp_path / "sinker_c1")
    cmd = ["gcc", "-std=c11", "-Wall", "-Wextra", "-Werror", "-I", os.path.join(ROOT, "include"), SRC,
           "-L", LIBDIR, "-lgale_stokes", f"-Wl,-rpath,{LIBDIR}", "-lm", "-o", exe]
    r = subprocess.run(cmd, capture_output=True, text=True)
    assert r.returncode == 0, r.stderr
    return exe


def test_c_caller_builds_and_refuses_without_device(tmp_path):
    import gale_amd
    if gale_amd.device_count() > 0:
        pytest.skip("a GPU is visible: the no-device path is checked on CPU-only hosts")
    exe = _build(tmp_path)
    r = subprocess.run([exe], capture_output=True, text=True, timeout=120)
    assert r.returncode == 3, (r.stdout, r.stderr)
    assert "refused" in r.stdout


@pytest.mark.gpu
def test_c_caller_solves_like_the_python_mirror(tmp_path, gale):
    from gale_amd import configs
    exe = _build(tmp_path)
    print("running the C caller (2D sinker 32^2, 20 ppc)", flush=True)
    r = subprocess.run([exe], capture_output=True, text=True, timeout=120)
    assert r.returncode == 0, (r.stdout, r.stderr)
    m = re.search(r"uzawa (\d+) its, rel residual ([^,]+), div ([^,]+), \|v\|max ([^,]+), dt ([^,]+),", r.stdout)
    assert m, r.stdout
    its, vmax, dt = int(m.group(1)), float(m.group(4)), float(m.group(5))
    cfg = configs.sinker_2d(res=32, ppc=20)
    pb = gale.Problem(**cfg.problem_kwargs())
    s = gale.Stokes_SLE_UzawaSolver(pb, gale.uzawa_params(tolerance=1e-6))
    u, p, rep = s.Solve()
    assert rep.uzawa_iterations == its
    vm = np.sqrt((u.reshape(-1, 2) ** 2).sum(1)).max()
    assert abs(vm - vmax) <= 1e-6 * vm          # printed with 7 significant digits
    assert abs(s.compute_dt(0.5) - dt) <= 1e-6 * dt
    s.close()
    pb.close()
