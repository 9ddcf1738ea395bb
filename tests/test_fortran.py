"""The Fortran FEMSolver_ binding (plantfem_gpu module): builds, runs, matches the oracle."""
import subprocess
from pathlib import Path

import numpy as np
import pytest

ROOT = Path(__file__).resolve().parent.parent
BIN = ROOT / "plantfem_amd" / "lib" / "bin"


def test_fortran_binaries_built_and_link():
    for name in ("cube_static", "plugin_solver", "modal_bar", "overset_organs", "organ_meshes"):
        assert (BIN / name).exists(), name
    out = subprocess.run([str(BIN / "cube_static"), "--version"], capture_output=True, text=True, timeout=60)
    assert out.returncode == 0 and "plantfem_gpu library version" in out.stdout


def _oracle_cube(n):
    from oracle import oracle as O
    nod, el = O.cube3d(n, n, n)
    ne = el.shape[0]
    E = 1.0e7 * (1.0 + 0.5 * np.sin(np.arange(1, ne + 1, dtype=np.float64)))
    nu = np.full(ne, 0.3)
    rp, ci = O.zero_matrix(el, nod.shape[0], 3)
    val = O.assemble(nod, el, E, nu, rp, ci)
    rhs = O.traction_box(nod, el, [-1e300, 1e300, -1e300, 1e300, 1 - 1e-9, 1 + 1e-9], -1.0e4)
    fixed = np.zeros(rhs.size, np.uint8)
    bottom = np.nonzero(nod[:, 2] <= 0)[0]
    for d in range(3):
        fixed[3 * bottom + d] = 1
    x, st, *_ = O.solve(rp, ci, val, rhs, fixed, np.zeros(rhs.size), relative_er=1e-14, er0=1e-30)
    assert st == 0
    return x


@pytest.mark.gpu
@pytest.mark.parametrize("n,mode,alg", [(8, "crs", "bicgstab"), (5, "elem", "bicgstab"), (7, "crs", "GaussJordan")])
def test_fortran_cube_static_matches_oracle(tmp_path, n, mode, alg):
    out = tmp_path / "x.bin"
    print("fortran cube_static", n, mode, alg, flush=True)
    res = subprocess.run([str(BIN / "cube_static"), str(n), str(out), mode, alg], capture_output=True, text=True,
                         timeout=120)
    assert res.returncode == 0, res.stdout + res.stderr
    x = np.fromfile(out, dtype=np.float64)
    xo = _oracle_cube(n)
    assert x.size == xo.size
    assert np.linalg.norm(x - xo) <= 1e-10 * np.linalg.norm(xo)


@pytest.mark.gpu
def test_fortran_plugin_solver():
    print("fortran plugin_solver", flush=True)
    res = subprocess.run([str(BIN / "plugin_solver")], capture_output=True, text=True, timeout=120)
    assert res.returncode == 0, res.stdout + res.stderr
    x = np.array([float(v) for v in res.stdout.split("x=", 1)[1].split("\n")[0].split()])
    A = np.array([[2, -1, 0, 0], [-1, 2, -1, 0], [0, 2, 3, 1], [0, 0, 1, 2]], float)
    want = np.linalg.solve(A, [1, 2, 3, 4])
    assert np.allclose(x, want, rtol=1e-9, atol=1e-9)
    # LinearSolver_'s bicgstab_CRS argument list (stop r.r / r0.r0 < 1e-22)
    x_crs = np.array([float(v) for v in res.stdout.split("x_crs=")[1].split("\n")[0].split()])
    assert np.allclose(x_crs, want, rtol=1e-9, atol=1e-9)


@pytest.mark.gpu
@pytest.mark.parametrize("dims,nev", [((2, 2, 6), 0), ((3, 3, 8), 6)])
def test_fortran_modal_bar_matches_scipy(dims, nev):
    """Tutorial/fem/ModalAnalysis_FixBoundary.f90's sequence in Fortran: every pair (nev 0) or
    the lowest nev (LOBPCG) against scipy's dense solver on the oracle's reduced pencil."""
    import scipy.linalg as sl
    from oracle import oracle as O
    print("fortran modal_bar", dims, nev, flush=True)
    args = [str(BIN / "modal_bar")] + [str(d) for d in dims] + ([str(nev)] if nev else [])
    res = subprocess.run(args, capture_output=True, text=True, timeout=120)
    assert res.returncode == 0, res.stdout + res.stderr
    lines = res.stdout.split("count ")[1].split()
    lam = np.array([float(v) for v in lines[1:1 + int(lines[0])]])
    nod, el = O.cube3d(*dims)
    ne = el.shape[0]
    rp, ci = O.zero_matrix(el, nod.shape[0], 3)
    A = O.assemble(nod, el, np.full(ne, 70000.0), np.full(ne, 0.3), rp, ci)
    B = O.mass_matrix_crs(nod, el, np.full(ne, 2.7), rp, ci)
    n = rp.size - 1
    z = nod[:, 2]
    nodes = np.nonzero((z == z.max()) | (z == z.min()))[0]
    fixed = np.concatenate([3 * nodes, 3 * nodes + 1, 3 * nodes + 2])
    free = np.setdiff1d(np.arange(n), fixed)

    def dense(v):
        D = np.zeros((n, n))
        for i in range(n):
            for k in range(rp[i] - 1, rp[i + 1] - 1):
                D[i, ci[k] - 1] = v[k]
        return D[np.ix_(free, free)]

    ref = sl.eigh(dense(A), dense(B), eigvals_only=True)
    if nev:
        ref = ref[:nev]
    assert lam.size == ref.size
    np.testing.assert_allclose(lam, ref, rtol=1e-7 if nev else 1e-9, atol=0)


@pytest.mark.gpu
@pytest.mark.parametrize("alg", ["GPP", "P2P"])
def test_fortran_overset_organs_matches_oracle(tmp_path, alg):
    # two overlapping organs through overset / init(NumDomain=2) / setEbOM, against the oracle's
    # restatement of the same sequence
    from oracle import oracle as O
    from test_overset import box, joint
    out = tmp_path / "x.bin"
    print("fortran overset_organs", alg, flush=True)
    res = subprocess.run([str(BIN / "overset_organs"), str(out), alg], capture_output=True, text=True, timeout=120)
    assert res.returncode == 0, res.stdout + res.stderr
    raw = np.fromfile(out, dtype=np.int64, count=2)
    x = np.fromfile(out, dtype=np.float64)[2:]
    A = box(3, 3, 4, 1.0, 1.0, 2.0, (0.0, 0.0, 0.0))
    B = box(3, 3, 3, 1.0, 1.0, 1.2, (0.05, 0.07, 1.43))
    if alg == "GPP":
        me, gp, pe, pos = O.overset_gpp(*A, *B)
    else:
        me, pe, pos = O.overset_p2p(*A, *B)
        gp = np.zeros(me.size, np.int32)
    xyz, en = joint(A, B)
    pair = pe + A[1].shape[0]
    rp, ci = O.multi_crs(xyz.shape[0], en, True, me, pair, gp)
    assert raw[0] == me.size and raw[1] == ci.size
    ne = en.shape[0]
    val = O.assemble(xyz, en, np.full(ne, 1.0e4), np.full(ne, 0.3), rp, ci)
    val = O.ebom(xyz, en, rp, ci, val, me, gp, pair, pos, 1.0e6)
    n = 3 * xyz.shape[0]
    rhs = np.zeros(n)
    top = A[0].shape[0] + np.nonzero(B[0][:, 2] > B[0][:, 2].max() - 1e-9)[0]
    rhs[3 * top] = 0.1
    fixed = np.zeros(n, np.uint8)
    base = np.nonzero(A[0][:, 2] <= A[0][:, 2].min())[0]
    for d in range(3):
        fixed[3 * base + d] = 1
    xo, st, *_ = O.solve(rp, ci, val, rhs, fixed, np.zeros(n), relative_er=1e-12, er0=1e-30)
    assert st == 0 and x.size == n
    assert np.linalg.norm(x - xo) <= 1e-9 * np.linalg.norm(xo)


def test_fortran_organ_meshes_match_python(tmp_path):
    # FEMDomain_%create("Leaf3D" | "Cylinder") of the Fortran module and plantfem_amd.fem's
    # generators give the same meshes, bit for bit (host only, no device needed)
    from plantfem_amd.fem import PF_GLYCINE_MAX, cylinder3d, leaf3d
    out = tmp_path / "m.bin"
    res = subprocess.run([str(BIN / "organ_meshes"), str(out)], capture_output=True, text=True, timeout=60)
    assert res.returncode == 0, res.stdout + res.stderr
    raw = out.read_bytes()
    off = 0
    refs = [leaf3d(8, 2, 20, 0.2, 0.01, 1.0, shaperatio=0.3),
            leaf3d(8, 2, 30, 0.1, 0.01, 0.3, species=PF_GLYCINE_MAX, SoyWidthRatio=0.5),
            cylinder3d(10, 10, 6, 0.08, 0.08, 0.6)]
    for nod, el in refs:
        nn, ne = np.frombuffer(raw, np.int64, 2, off)
        off += 16
        x = np.frombuffer(raw, np.float64, nn * 3, off).reshape(3, nn).T
        off += 24 * nn
        e = np.frombuffer(raw, np.int32, ne * 8, off).reshape(8, ne).T
        off += 32 * ne
        assert np.array_equal(x, nod) and np.array_equal(e, el)
    assert off == len(raw)
