"""Writes the golden fixtures under tests/golden/ from data files that the
reference's own tests hold (run once in a tree where the reference source is
available; the outputs are committed and the tests never read the reference).

  k_isoviscous_3x3.npz        StgFEM/Assembly/tests/expected/testIsoviscousSiffness.matrix
  k_material_viscosity_10x10.npz
                              Underworld/Rheology/tests/expected/testMaterialViscosity2D.matrix
  constitutive_2d.json        Underworld/Rheology/tests/expected/testConstitutiveMatrix_Cartesian2D.expected
  dvc_weights.json            the ten particle positions of PICellerator/Weights/tests/DVCWeightsSuite.cxx
                              (2D and 3D) with the centroids and weights of
                              PICellerator/Weights/tests/expected/DVCWeightsSuite_testGetCentroids{,2D}.txt
  pcdvc_expected.json         PICellerator/Utils/tests/expected/testPCDVC_*.expected (tolerance and value of
                              the mean and standard deviation of the relative integration error)
  sobol.json                  the 100 (degree, coefficient, initial direction numbers) rows of the published
                              Joe-Kuo table as tabulated in StgDomain/Utils/src/SobolGenerator.cxx, and the
                              sampled outputs (index % 773 == 3 of 200000, printed %.4g) of
                              StgDomain/Utils/tests/expected/testSobolGeneratorOutput.<degree>-<coeff>.expected
  spacefiller.json            StgDomain/Swarm/tests/expected/testSpaceFillerParticleLayoutOutput.expected
                              (cell, in-cell index, global index, coordinates as printed %7.5g)
  rayleigh_taylor_step10/     the HDF5 checkpoint files (Mesh, VelocityField, PressureField, materialSwarm,
                              timeInfo at step 10) of the reference's regression fixtures
  lid_driven_pic_step10/      Underworld/SysTest/RegressionTests/expected/RayleighTaylorBenchmark.xml/,
  stokes_multigrid_step10/    PICellerator/SysTest/RegressionTests/expected/LidDrivenPIC2D.xml/,
  rayleigh_taylor_lowres_step5/  StgFEM/SysTest/RegressionTests/expected/Multigrid.xml/ and
                              Underworld/SysTest/RegressionTests/lowres_expected/RayleighTaylorBenchmark.xml/
                              (the suite's own low-resolution run, step 5), copied unchanged as data

The .matrix files are PETSc binary AIJ matrices (big-endian int32 header
1211216, M, N, nnz; row lengths; column indices; big-endian float64 values);
they are parsed here as data.
"""
import json
import os
import re
import sys

import numpy as np

REF = sys.argv[1] if len(sys.argv) > 1 else "/root/reference"
OUT = os.path.dirname(os.path.abspath(__file__))


def load_petsc_aij(path):
    b = open(path, "rb").read()
    cls, M, N, nnz = np.frombuffer(b[:16], dtype=">i4")
    assert cls == 1211216
    off = 16
    rl = np.frombuffer(b[off:off + 4 * M], dtype=">i4").astype(np.int64)
    off += 4 * M
    cols = np.frombuffer(b[off:off + 4 * nnz], dtype=">i4").astype(np.int32)
    off += 4 * nnz
    vals = np.frombuffer(b[off:off + 8 * nnz], dtype=">f8").astype(np.float64)
    indptr = np.concatenate([[0], np.cumsum(rl)])
    return dict(shape=np.array([M, N]), indptr=indptr, indices=cols, data=vals)


def main():
    pairs = [
        ("StgFEM/Assembly/tests/expected/testIsoviscousSiffness.matrix", "k_isoviscous_3x3.npz"),
        ("Underworld/Rheology/tests/expected/testMaterialViscosity2D.matrix", "k_material_viscosity_10x10.npz"),
    ]
    for src, dst in pairs:
        m = load_petsc_aij(os.path.join(REF, src))
        np.savez_compressed(os.path.join(OUT, dst), **m)
        print(dst, m["shape"], len(m["data"]))
    txt = open(os.path.join(REF, "Underworld/Rheology/tests/expected/testConstitutiveMatrix_Cartesian2D.expected")).read()
    block = txt.split("SetMatrixWithViscosity7")[1].split("SetMatrixWithSecondViscosity2")[0]
    D = [[float(v) for v in re.findall(r"matrixData\[ %d \]\[ \d \] = ([-\d.]+)" % i, block)] for i in range(3)]
    visc = float(re.search(r"Viscosity =\s+([-\d.]+)", block).group(1))
    nums = [float(v) for v in re.findall(r"^\s+([-\d.]+)\s+([-\d.]+)\s*$", block, re.M) for v in v]
    strain = [[nums[0], nums[1]], [nums[2], nums[3]]]
    stress = [[nums[4], nums[5]], [nums[6], nums[7]]]
    json.dump(dict(viscosity=visc, D=D, strain_rate=strain, stress=stress),
              open(os.path.join(OUT, "constitutive_2d.json"), "w"), indent=1)
    print("constitutive_2d.json", visc, D, strain, stress)
    dvc_fixtures()


def dvc_fixtures():
    src = open(os.path.join(REF, "PICellerator/Weights/tests/DVCWeightsSuite.cxx")).read()
    out = {}
    for dim, lst, fn in ((3, "pList", "DVCWeightsSuite_testGetCentroids.txt"),
                         (2, "pList2D", "DVCWeightsSuite_testGetCentroids2D.txt")):
        xs = []
        for ax in "xyz"[:dim]:
            v = [float(t) for t in re.findall(r"data->%s\[i\]\.%s = ([-+\d.]+)" % (lst, ax), src)]
            assert len(v) == 10, (lst, ax, len(v))
            xs.append(v)
        xi = [[xs[d][i] for d in range(dim)] for i in range(10)]
        exp = open(os.path.join(REF, "PICellerator/Weights/tests/expected", fn)).read()
        cent = [[float(t) for t in m] for m in re.findall(r"centroids: \([^)]*\) = \(" + ",".join([r"\s*([-\d.]+)"] * dim) + r"\)", exp)]
        w = [float(t) for t in re.findall(r"weight = ([-\d.]+)", exp)]
        assert len(cent) == 10 and len(w) == 10
        out["%dd" % dim] = dict(res=[10] * dim, xi=xi, centroids=cent, weights=w, tol_sq=1e-6)
    json.dump(out, open(os.path.join(OUT, "dvc_weights.json"), "w"), indent=1)
    print("dvc_weights.json", {k: len(v["xi"]) for k, v in out.items()})
    pc = {}
    for name in ("CircleInterface", "ExponentialInterface", "PolynomialFunction"):
        lines = open(os.path.join(REF, "PICellerator/Utils/tests/expected/testPCDVC_%s.expected" % name)).read().split("\n")
        (mt, mv), (st, sv) = [[float(t) for t in l.split()] for l in lines[:2]]
        pc[name] = dict(mean_tol=mt, mean=mv, std_tol=st, std=sv)
    json.dump(pc, open(os.path.join(OUT, "pcdvc_expected.json"), "w"), indent=1)
    print("pcdvc_expected.json", pc)
    sobol_fixtures()
    regression_fixtures()


def regression_fixtures():
    import shutil
    swarm = ("Mesh", "VelocityField", "PressureField", "materialSwarm", "timeInfo")
    for src, name, files in (
            ("Underworld/SysTest/RegressionTests/expected/RayleighTaylorBenchmark.xml", "rayleigh_taylor_step10", swarm),
            ("PICellerator/SysTest/RegressionTests/expected/LidDrivenPIC2D.xml", "lid_driven_pic_step10", swarm),
            ("StgFEM/SysTest/RegressionTests/expected/Multigrid.xml", "stokes_multigrid_step10",
             ("Mesh.linearMesh", "VelocityField", "PressureField", "timeInfo")),
            ("Underworld/SysTest/RegressionTests/lowres_expected/RayleighTaylorBenchmark.xml",
             "rayleigh_taylor_lowres_step5",
             ("Mesh.linearMesh", "VelocityField", "PressureField", "materialSwarm", "timeInfo"))):
        dst = os.path.join(OUT, name)
        os.makedirs(dst, exist_ok=True)
        step = "00005" if name.endswith("step5") else "00010"
        for fn in files:
            shutil.copyfile(os.path.join(REF, src, fn + "." + step + ".h5"), os.path.join(dst, fn + "." + step + ".h5"))
            os.chmod(os.path.join(dst, fn + "." + step + ".h5"), 0o644)
        print(name, sorted(os.listdir(dst)))


def c_array(src, name):
    body = re.search(name + r"\[\]\s*(?:\[\d+\])?\s*=\s*\{(.*?)\};", src, re.S).group(1)
    return re.sub(r"/\*.*?\*/", "", body, flags=re.S)


def sobol_fixtures():
    src = open(os.path.join(REF, "StgDomain/Utils/src/SobolGenerator.cxx")).read()
    deg = [int(t) for t in re.findall(r"\d+", c_array(src, "SobolGenerator_PolynomialDegree"))]
    coef = [int(t) for t in re.findall(r"\d+", c_array(src, "SobolGenerator_PolynomialCoefficient"))]
    m = [[int(t) for t in re.findall(r"\d+", row)]
         for row in re.findall(r"\{([^{}]*)\}", c_array(src, "SobolGenerator_InitialDirectionNumbers"))]
    assert len(deg) == len(coef) == len(m) == 100 and all(len(m[i]) == deg[i] for i in range(100))
    gens = []
    for i in range(100):
        fn = "StgDomain/Utils/tests/expected/testSobolGeneratorOutput.%03d-%03d.expected" % (deg[i], coef[i])
        lines = [l.strip() for l in open(os.path.join(REF, fn)).read().split("\n")[1:] if l.strip()]
        assert len(lines) == len(range(3, 200000, 773)), (fn, len(lines))
        gens.append(dict(degree=deg[i], coefficient=coef[i], m=m[i], sampled=lines))
    json.dump(dict(stride=773, offset=3, count=200000, generators=gens),
              open(os.path.join(OUT, "sobol.json"), "w"), separators=(",", ":"))
    print("sobol.json", len(gens))
    txt = open(os.path.join(REF, "StgDomain/Swarm/tests/expected/testSpaceFillerParticleLayoutOutput.expected")).read()
    parts = []
    cell = -1
    for line in txt.split("\n"):
        mc = re.match(r"\s*Local Cell\s+(\d+):", line)
        if mc:
            cell = int(mc.group(1))
            continue
        mp = re.match(r"\s*part\. InCell\s+(\d+) \(local index\s+(\d+)\): \((.*)\)", line)
        if mp:
            parts.append(dict(cell=cell, in_cell=int(mp.group(1)), index=int(mp.group(2)),
                              coord=[t.strip() for t in mp.group(3).split(",")]))
    assert len(parts) == 160
    json.dump(dict(dim=3, res=[4, 2, 1], lo=[0.0, 0.0, 0.0], hi=[400.0, 200.0, 100.0], avg_per_cell=20,
                   format="%7.5g", particles=parts),
              open(os.path.join(OUT, "spacefiller.json"), "w"), indent=0)
    print("spacefiller.json", len(parts))


if __name__ == "__main__":
    main()
