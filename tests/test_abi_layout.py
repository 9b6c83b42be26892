"""The ctypes mirrors in gale_amd/__init__.py against the C ABI in include/*.h:
a small C program built with gcc from the headers prints sizeof and offsetof
of every member the mirror names, and each must equal the mirror's size and
field offset.  A member added to a header but not to the mirror (or the
reverse, or in another order) fails here instead of corrupting arguments at
run time.  CPU only."""
import ctypes as C
import os
import shutil
import subprocess

import pytest

import gale_amd as G

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))

PAIRS = [
    ("gs_material", G.Material), ("gs_problem_desc", G.ProblemDesc), ("gs_uzawa_params", G.UzawaParams),
    ("gs_picard_params", G.PicardParams), ("gs_report", G.Report), ("gs_shape", G.Shape),
    ("gs_initial_damage_params", G.InitialDamageParams), ("gs_pcdvc_params", G.PcdvcParams),
    ("gs_lpc_train_params", G.TrainParams), ("gs_batch_params", G.BatchParams),
    ("gs_batch_result", G.BatchResult),
]


@pytest.fixture(scope="module")
def c_layout(tmp_path_factory):
    if shutil.which("gcc") is None:
        pytest.skip("gcc not available")
    lines = ["#include <stddef.h>", "#include <stdio.h>"]
    for h in sorted(os.listdir(os.path.join(ROOT, "include"))):
        if h.endswith(".h"):
            lines.append(f'#include "{h}"')
    lines.append("int main(void) {")
    for cname, cls in PAIRS:
        lines.append(f'  printf("{cname} sizeof %zu\\n", sizeof({cname}));')
        for f in cls._fields_:
            lines.append(f'  printf("{cname} {f[0]} %zu\\n", offsetof({cname}, {f[0]}));')
    lines += ["  return 0;", "}"]
    d = tmp_path_factory.mktemp("abi")
    src, exe = d / "layout.c", d / "layout"
    src.write_text("\n".join(lines) + "\n")
    r = subprocess.run(["gcc", "-std=c11", "-I", os.path.join(ROOT, "include"), str(src), "-o", str(exe)],
                       capture_output=True, text=True)
    assert r.returncode == 0, r.stderr
    out = subprocess.run([str(exe)], capture_output=True, text=True, check=True).stdout
    table = {}
    for ln in out.splitlines():
        struct, field, value = ln.split()
        table[(struct, field)] = int(value)
    return table


@pytest.mark.parametrize("cname,cls", PAIRS, ids=[p[0] for p in PAIRS])
def test_mirror_matches_header(c_layout, cname, cls):
    assert C.sizeof(cls) == c_layout[(cname, "sizeof")], (cname, C.sizeof(cls), c_layout[(cname, "sizeof")])
    for f in cls._fields_:
        assert getattr(cls, f[0]).offset == c_layout[(cname, f[0])], (cname, f[0])
