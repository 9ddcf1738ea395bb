"""Mass matrix and the eigen solver (MassMatrixFEMDomain, MassMatrix_as_CRS, eigFEMSolver).

CPU: the oracle's element mass matrix against the closed form of a brick (rho V / 216 per
shared-coordinate factor), row sums against the (separately pinned) mass vector, and the CRS
assembly against an ordered dense scatter.  GPU: element and CRS mass matrices bitwise vs the
oracle; every eigenpair (dense path) and the lowest pairs (LOBPCG) against scipy's dense
generalized symmetric solver on the reduced pencil; the 3x3 CRS eigenproblem of
Tutorial/std/solveEigenProblem_as_CRS.f90 against numpy.
"""
import sys

import numpy as np
import pytest
import scipy.linalg as sl

from oracle import oracle as O
from plantfem_amd.fem import CRS, FEMDomain, FEMSolver, cube3d


def say(*a):
    print(*a, flush=True)
    sys.stdout.flush()


SIGN = np.array([[-1, -1, -1], [1, -1, -1], [1, 1, -1], [-1, 1, -1], [-1, -1, 1], [1, -1, 1], [1, 1, 1], [-1, 1, 1]])


def brick(lx=2.0, ly=3.0, lz=5.0):
    return np.array([[0, 0, 0], [lx, 0, 0], [lx, ly, 0], [0, ly, 0],
                     [0, 0, lz], [lx, 0, lz], [lx, ly, lz], [0, ly, lz]], float) + np.array([1.0, -2.0, 0.5])


def test_mass_matrix_brick_closed_form():
    X = brick()
    rho, V = 2.7, 30.0
    M = O.hex8_mass_matrix(X, rho)
    ref = np.zeros((24, 24))
    for a in range(8):
        for b in range(8):
            f = np.prod([2.0 if SIGN[a, d] == SIGN[b, d] else 1.0 for d in range(3)])
            for k in range(3):
                ref[3 * a + k, 3 * b + k] = rho * V / 216.0 * f
    np.testing.assert_allclose(M, ref, rtol=1e-14, atol=0)
    assert np.array_equal(M, M.T)


def test_mass_matrix_rows_sum_to_mass_vector():
    # sum_b N_a N_b = N_a (partition of unity): the row sums of the consistent mass matrix are
    # the mass vector of a unit acceleration, on distorted elements too
    rng = np.random.default_rng(4)
    X = brick() + 0.3 * rng.standard_normal((8, 3))
    M = O.hex8_mass_matrix(X, 1.3)
    assert np.all(np.linalg.eigvalsh(M) > 0)
    for k in range(3):
        acc = np.zeros(3)
        acc[k] = 1.0
        mv = O.hex8_mass_vector(X.ravel(), 1.3, acc)
        rows = [3 * a + k for a in range(8)]
        np.testing.assert_allclose(M[rows].sum(axis=1), mv[rows], rtol=1e-13, atol=0)


def test_mass_crs_is_ordered_scatter():
    nod, el = cube3d(3, 2, 2)
    rng = np.random.default_rng(6)
    nod = np.asfortranarray(nod + 0.05 * rng.standard_normal(nod.shape))
    rho = 1.0 + rng.random(el.shape[0])
    rp, ci = O.zero_matrix(el, nod.shape[0], 3)
    val = O.mass_matrix_crs(nod, el, rho, rp, ci)
    n = rp.size - 1
    dense = np.zeros((n, n))
    for e in range(el.shape[0]):
        M = O.hex8_mass_matrix(nod[el[e] - 1], rho[e])
        dofs = (3 * (el[e] - 1)[:, None] + np.arange(3)).ravel()
        for p in range(24):
            for q in range(24):
                dense[dofs[p], dofs[q]] = dense[dofs[p], dofs[q]] + M[p, q]
    for i in range(n):
        for k in range(rp[i] - 1, rp[i + 1] - 1):
            assert val[k] == dense[i, ci[k] - 1]


# ------------------------------------------------------------------------------ GPU
def modal_system(ctx, dims, E=70000.0, nu=0.3, rho=2.7, fix_top_bottom=True, jitter=0.0, seed=0):
    nod, el = cube3d(*dims)
    if jitter:
        rng = np.random.default_rng(seed)
        nod = np.asfortranarray(nod + jitter * rng.standard_normal(nod.shape) / max(dims))
    ne = el.shape[0]
    s = FEMSolver(ctx)
    s.setDomain(FEMDomain(nod, el, ctx))
    s.setMassMatrix(np.full(ne, rho))
    s.keepThisMatrixAs("B")
    s.zeros()
    s.setStiffness(np.full(ne, E), np.full(ne, nu))
    s.keepThisMatrixAs("A")
    rp, ci = O.zero_matrix(el, nod.shape[0], 3)
    A = O.assemble(nod, el, np.full(ne, E), np.full(ne, nu), rp, ci)
    B = O.mass_matrix_crs(nod, el, np.full(ne, rho), rp, ci)
    fixed = []
    if fix_top_bottom:
        z = nod[:, 2]
        nodes = np.nonzero((z == z.max()) | (z == z.min()))[0] + 1
        fixed = np.sort(np.concatenate([3 * nodes - 2, 3 * nodes - 1, 3 * nodes])).astype(np.int32)
        s.fix_eig(fixed)
    return s, rp, ci, A, B, np.asarray(fixed, np.int32)


def dense(rp, ci, v):
    n = rp.size - 1
    D = np.zeros((n, n))
    for i in range(n):
        for k in range(rp[i] - 1, rp[i + 1] - 1):
            D[i, ci[k] - 1] = v[k]
    return D


@pytest.mark.gpu
def test_gpu_mass_matrices_bitwise(ctx):
    nod, el = cube3d(3, 4, 2)
    rng = np.random.default_rng(8)
    nod = np.asfortranarray(nod + 0.04 * rng.standard_normal(nod.shape))
    ne = el.shape[0]
    rho = 1.0 + rng.random(ne)
    d = FEMDomain(nod, el, ctx)
    ids = np.arange(1, ne + 1)
    Mg = d.MassMatrix(ids, Density=rho)
    for e in range(ne):
        assert np.array_equal(Mg[e], O.hex8_mass_matrix(nod[el[e] - 1], rho[e]))
    s = FEMSolver(ctx)
    s.setDomain(d)
    s.setMassMatrix(rho)
    crs = s.getCRS()
    rp, ci = O.zero_matrix(el, nod.shape[0], 3)
    assert np.array_equal(crs.val, O.mass_matrix_crs(nod, el, rho, rp, ci))


@pytest.mark.gpu
def test_gpu_eig_all_pairs_match_scipy(ctx):
    # Tutorial/fem/ModalAnalysis_FixBoundary.f90's set-up on a small bar: mass "B", stiffness
    # "A", top and bottom faces fixed, every eigenpair
    s, rp, ci, A, B, fixed = modal_system(ctx, (2, 2, 6))
    lam, V = s.eig()
    free = np.setdiff1d(np.arange(rp.size - 1), fixed - 1)
    Ad, Bd = dense(rp, ci, A)[np.ix_(free, free)], dense(rp, ci, B)[np.ix_(free, free)]
    ref = sl.eigh(Ad, Bd, eigvals_only=True)
    say("eig dense", lam.size, lam[:4], ref[:4])
    assert lam.size == free.size
    np.testing.assert_allclose(lam, ref, rtol=1e-9, atol=0)
    assert np.all(V[fixed - 1] == 0.0)
    Bf = dense(rp, ci, B)
    np.testing.assert_allclose(V.T @ Bf @ V, np.eye(lam.size), atol=1e-9)
    Af = dense(rp, ci, A)
    R = (Af @ V - (Bf @ V) * lam)[free]  # the reduced problem: free rows only
    assert np.abs(R).max() <= 1e-8 * np.abs((Af @ V)[free]).max()


@pytest.mark.gpu
@pytest.mark.parametrize("dims,fix", [((3, 3, 8), True), ((2, 3, 4), False)])
def test_gpu_eig_lobpcg_lowest(ctx, dims, fix):
    s, rp, ci, A, B, fixed = modal_system(ctx, dims, fix_top_bottom=fix, jitter=0.1)
    nev = 6
    if not fix:  # a free body: six rigid-body modes with eigenvalue 0 first
        nev = 9
    lam, V = s.eig(num_eigen=nev, tol=1e-8)
    st = s.last_stats
    n = rp.size - 1
    free = np.setdiff1d(np.arange(n), fixed - 1) if fix else np.arange(n)
    Ad, Bd = dense(rp, ci, A)[np.ix_(free, free)], dense(rp, ci, B)[np.ix_(free, free)]
    ref = sl.eigh(Ad, Bd, eigvals_only=True)[:nev]
    say("lobpcg", dims, "iterations", st["iterations"], "res", st["true_relres"], lam, ref)
    assert st["status"] == 0
    scale = np.abs(ref).max()
    np.testing.assert_allclose(lam, ref, rtol=1e-7, atol=1e-9 * scale)
    Bf = dense(rp, ci, B)
    np.testing.assert_allclose(V.T @ Bf @ V, np.eye(nev), atol=1e-7)
    if fix:
        assert np.all(V[fixed - 1] == 0.0)


@pytest.mark.gpu
def test_gpu_eig_reference_crs_3x3(ctx):
    # Tutorial/std/solveEigenProblem_as_CRS.f90: the symmetric 3x3 [1 2 3; 2 1 1; 3 1 8]
    # (B = I on the same pattern)
    Am = np.array([[1.0, 2.0, 3.0], [2.0, 1.0, 1.0], [3.0, 1.0, 8.0]])
    rp = np.array([1, 4, 7, 10], np.int64)
    ci = np.array([1, 2, 3, 1, 2, 3, 1, 2, 3], np.int32)
    s = FEMSolver(ctx)
    s.setCRS(CRS=CRS(rp, ci, Am.ravel()))
    s.keepThisMatrixAs("A")
    s.setCRS(CRS=CRS(rp, ci, np.eye(3).ravel()))
    s.keepThisMatrixAs("B")
    lam, V = s.eig()
    np.testing.assert_allclose(lam, np.linalg.eigvalsh(Am), rtol=1e-12, atol=1e-13)
    np.testing.assert_allclose(np.abs(V.T @ V), np.eye(3), atol=1e-12)
