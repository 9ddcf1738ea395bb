"""Direct solve: solve(algorithm="GaussJordan") (FEMSolverClass.f90:1795-1798).

The reference's gauss_jordan_crs (FEMSolverClass.f90:2413-2460) stops "Under construction", so
there is no reference output to match: the answer is fixed by the mathematics, x = A^-1 b of the
eliminated system.  The oracle restates LAPACK's band Cholesky (dpbtf2 / dpbtrs, lower); the CPU
tests pin it against SciPy's LAPACK (cholesky_banded, solveh_banded), dense numpy solves and
SciPy's sparse LU, and check its planted-failure paths.  The GPU tests compare the device factor
entry by entry with the oracle's factor of the same renumbered band matrix (the GPU's
permutation and bandwidth, read back), the solutions with the oracle's and with SciPy's sparse
LU, every launch of the batch kernel on organ meshes, and the error paths (matrix not positive
definite, pattern change, factor reuse).

Tolerances: the GPU factor sums the Schur updates in blocks of 4 columns with fused
multiply-adds (MFMA), the oracle column by column without contraction; both are backward
stable, so the factors agree to a few ulps of the largest entry on these well-conditioned
meshes (1e-12 relative is asserted).  A backward-stable solve leaves a residual of order
eps |A| |x|, which for a slab clamped along one edge is ~1e-12 |b| (SciPy's sparse LU reaches
the same), so residuals are bounded by 10x the oracle's (or SciPy's) on the same system, or by
100x the rounding floor eps |(|A| |x|)| / |b| of evaluating the residual itself.  Solutions of two backward-stable solvers
differ by up to about cond(A) * eps; the slender meshes have cond(A) ~ 1e6-1e9, so solution
differences are bounded by 1e-8 (1e-10 on the compact and small meshes).
"""
import sys

import numpy as np
import pytest
import scipy.linalg as sl
import scipy.sparse as sp
import scipy.sparse.linalg as spl

from oracle import oracle as O
from plantfem_amd import batch as B
from plantfem_amd import workloads as W


def say(*a):
    print(*a, flush=True)
    sys.stdout.flush()


def random_band_spd(n, kd, rng, shift=None):
    A = np.zeros((n, n))
    for i in range(n):
        for j in range(max(0, i - kd), i + 1):
            A[i, j] = A[j, i] = rng.standard_normal()
    A += np.eye(n) * (2 * kd + 3 if shift is None else shift)
    ab = np.zeros((n, kd + 1))
    for j in range(n):
        for d in range(min(kd, n - 1 - j) + 1):
            ab[j, d] = A[j + d, j]
    return A, ab


def organ_system(o: B.Organ):
    """The oracle's eliminated system of one organ mesh (ZeroMatrix, assembly, traction, fix,
    elimination)."""
    rp, ci = O.zero_matrix(o.elemnod, o.nn, 3)
    val = O.assemble(o.nodcoord, o.elemnod, o.E, o.nu, rp, ci)
    rhs = O.traction_box(o.nodcoord, o.elemnod, o.box, o.pressure)
    fixed = np.zeros(o.ndof, np.uint8)
    for d in range(3):
        fixed[3 * o.fixed_nodes - 3 + d] = 1
    v2, r2 = O.apply_dirichlet(rp, ci, val, rhs, fixed, np.zeros(o.ndof))
    return rp, ci, v2, r2


def to_scipy(rp, ci, val):
    n = rp.size - 1
    return sp.csr_matrix((val, ci.astype(np.int64) - 1, rp - 1), shape=(n, n))


def relres(A, x, b):
    return float(np.linalg.norm(b - A @ x) / np.linalg.norm(b))


def rounding_floor(A, x, b):
    """eps |(|A| |x|)| / |b|: the size of the rounding error of computing b - A x in FP64, below
    which no FP64 solve's residual can be measured."""
    return float(np.finfo(float).eps * np.linalg.norm(abs(A) @ np.abs(x)) / np.linalg.norm(b))


def natural_bandwidth(rp, ci):
    n = rp.size - 1
    rows = np.repeat(np.arange(n), np.diff(rp))
    return int(np.abs(rows - (ci.astype(np.int64) - 1)).max())


# ------------------------------------------------------------------------------ CPU


@pytest.mark.parametrize("n,kd", [(1, 0), (7, 0), (40, 3), (200, 17), (64, 80)])
def test_band_cholesky_matches_lapack(n, kd):
    rng = np.random.default_rng(n + kd)
    A, ab = random_band_spd(n, min(kd, max(n - 1, 0)), rng)
    info, L = O.band_cholesky(ab)
    assert info == 0
    ref = sl.cholesky_banded(ab.T, lower=True)
    assert np.abs(L.T - ref).max() <= 1e-14 * np.abs(ref).max()
    b = rng.standard_normal(n)
    x = O.band_solve(L, b)
    assert np.linalg.norm(x - np.linalg.solve(A, b)) <= 1e-13 * np.linalg.norm(x)
    assert np.linalg.norm(x - sl.solveh_banded(ab.T, b, lower=True)) <= 1e-13 * np.linalg.norm(x)


def test_band_cholesky_reports_the_first_bad_pivot():
    rng = np.random.default_rng(3)
    A, ab = random_band_spd(60, 5, rng)
    ab[37, 0] = -1.0e3  # A[37][37] negative: the pivot of column 37 fails
    info, _ = O.band_cholesky(ab)
    assert info == 38
    ab2 = np.zeros((4, 2))
    ab2[:, 0] = [1.0, 0.0, 2.0, 3.0]  # a zero pivot in column 1 ("A is singular")
    assert O.band_cholesky(ab2)[0] == 2


def test_band_from_crs_rejects_entries_outside_the_band():
    rng = np.random.default_rng(1)
    o = B.stem(6, rng)
    rp, ci, val, _ = organ_system(o)
    kd = natural_bandwidth(rp, ci)
    perm = np.arange(rp.size - 1, dtype=np.int32)
    O.band_from_crs(rp, ci, val, perm, kd)
    with pytest.raises(ValueError):
        O.band_from_crs(rp, ci, val, perm, kd - 1)


@pytest.mark.parametrize("kind", ["stem", "leaf"])
def test_oracle_direct_solve_matches_sparse_lu(kind):
    rng = np.random.default_rng(5)
    o = B.stem(10, rng) if kind == "stem" else B.leaf(9, 7, rng)
    rp, ci, val, rhs = organ_system(o)
    n = rp.size - 1
    kd = natural_bandwidth(rp, ci)
    # a random renumbering within the band of the natural order's blocks is still a band
    perm = np.arange(n, dtype=np.int32)
    x = O.direct_solve(rp, ci, val, rhs, perm, kd)
    A = to_scipy(rp, ci, val)
    xs = spl.spsolve(A.tocsc(), rhs)
    assert np.linalg.norm(x - xs) <= 1e-9 * np.linalg.norm(xs)
    # backward stable: the residual is of the order the sparse LU leaves
    assert relres(A, x, rhs) <= max(10.0 * relres(A, xs, rhs), 1e-14)
    # reversing the order keeps the bandwidth and the solution
    rev = perm[::-1].copy()
    xr = O.direct_solve(rp, ci, val, rhs, rev, kd)
    assert np.linalg.norm(xr - x) <= 1e-9 * np.linalg.norm(x)


# ------------------------------------------------------------------------------ GPU


def solver_for_organ(o: B.Organ, ctx):
    from plantfem_amd.fem import FEMDomain, FEMSolver
    s = FEMSolver(ctx)
    s.setDomain(FEMDomain(o.nodcoord, o.elemnod, ctx))
    s.setStiffness(o.E, o.nu)
    s.addTraction(o.box, o.pressure)
    ids = np.sort(np.concatenate([3 * o.fixed_nodes - 2, 3 * o.fixed_nodes - 1, 3 * o.fixed_nodes]))
    s.fix(1, ids, FixValue=0.0)
    return s


def small_organs(seed=0):
    rng = np.random.default_rng(seed)
    return [B.leaf(8, 6, rng), B.stem(12, rng), B.leaf(10, 12, rng), B.stem(20, rng), B.leaf(5, 5, rng),
            B.stem(7, rng), B.block(4, rng)]


@pytest.mark.gpu
@pytest.mark.parametrize("mfma", [True, False])
def test_gpu_direct_factor_matches_oracle(ctx, mfma):
    for k, o in enumerate(small_organs()):
        say("organ", k, o.kind, o.dims, "mfma", mfma)
        s = solver_for_organ(o, ctx)
        s.setDirectKernel(mfma)
        x = s.solve(algorithm="GaussJordan", refine=0)
        st = s.last_stats
        info = s.directInfo(factor=True)
        rp, ci, val, rhs = organ_system(o)
        n = rp.size - 1
        assert info["n"] == n
        perm = info["perm"]
        assert np.array_equal(np.sort(perm), np.arange(n))
        b = info["bandwidth"]
        ab = O.band_from_crs(rp, ci, val, perm, b)  # raises if an entry lies outside the band
        oinfo, lo = O.band_cholesky(ab)
        assert oinfo == 0
        lg = info["factor"].T  # (n, b+1)
        err = np.abs(lg - lo).max() / np.abs(lo).max()
        xo = O.direct_solve(rp, ci, val, rhs, perm, b)
        xerr = np.linalg.norm(x - xo) / np.linalg.norm(xo)
        A = to_scipy(rp, ci, val)
        ro = relres(A, xo, rhs)
        say("  n", n, "b", b, info["ordering"], "factor err", err, "x err", xerr, "relres", st["true_relres"],
            "oracle relres", ro)
        assert st["status"] == 0
        assert err <= 1e-12
        assert xerr <= 1e-10
        # at this level the residual is itself computed with a rounding error of its size, so
        # the device's and the host's evaluations are both bounded rather than compared
        assert relres(A, x, rhs) <= max(10.0 * ro, 1e-14)
        assert st["true_relres"] <= max(10.0 * ro, 1e-14)
        s.close()


@pytest.mark.gpu
def test_gpu_direct_orderings(ctx):
    """The leaf slab is renumbered along its longer in-plane axis, the stem keeps the natural
    order: the chosen half bandwidth is 3 (3 min(n, m) + 7) + 2 for an n x m x 2 leaf and
    3 * 31 + 2 for a 4 x 4 x L stem."""
    rng = np.random.default_rng(2)
    for o, want in ((B.leaf(30, 12, rng), 3 * (3 * 12 + 7) + 2), (B.leaf(12, 30, rng), 3 * (3 * 12 + 7) + 2),
                    (B.stem(40, rng), 3 * 31 + 2)):
        s = solver_for_organ(o, ctx)
        x = s.solve(algorithm="GaussJordan")
        info = s.directInfo()
        rp, ci, val, rhs = organ_system(o)
        A = to_scipy(rp, ci, val)
        fl = rounding_floor(A, x, rhs)
        say(o.kind, o.dims, info["ordering"], info["bandwidth"], "want", want, "relres", s.last_stats["true_relres"],
            "rounding floor", fl)
        assert info["bandwidth"] <= want
        assert relres(A, x, rhs) <= 100.0 * fl and s.last_stats["true_relres"] <= 100.0 * fl
        s.close()


@pytest.mark.gpu
def test_gpu_direct_config1_matches_sparse_lu(ctx):
    w = W.config(1)
    from plantfem_amd.fem import FEMDomain, FEMSolver
    s = FEMSolver(ctx)
    s.setDomain(FEMDomain(w.nodcoord, w.elemnod, ctx))
    s.setStiffness(w.E, w.nu)
    s.addTraction(w.traction_box, w.pressure)
    s.fix(1, w.fixed_dofs, FixValue=0.0)
    x = s.solve(algorithm="GaussJordan", refine=1)
    st = s.last_stats
    info = s.directInfo()
    rp, ci, val, rhs, fixed, ub = O.build_system(w)
    v2, r2 = O.apply_dirichlet(rp, ci, val, rhs, fixed, ub)
    A = to_scipy(rp, ci, v2)
    xs = spl.spsolve(A.tocsc(), r2)
    err = np.linalg.norm(x - xs) / np.linalg.norm(xs)
    res = np.linalg.norm(r2 - A @ x) / np.linalg.norm(r2)
    rs = relres(A, xs, r2)
    say("C1 direct: b", info["bandwidth"], info["ordering"], "err vs sparse LU", err, "relres", res,
        "gpu relres", st["true_relres"], "sparse LU relres", rs, "ms", st["solve_ms"])
    assert st["status"] == 0
    assert info["bandwidth"] <= 3 * (21 * 21 + 21 + 1) + 2
    assert res <= max(10.0 * rs, 1e-14) and res <= 100.0 * rounding_floor(A, x, r2)
    assert err <= 1e-9
    assert np.all(x[w.fixed_dofs - 1] == 0.0)
    s.close()


@pytest.mark.gpu
def test_gpu_direct_factor_reuse_and_refactor(ctx):
    rng = np.random.default_rng(4)
    o = B.leaf(14, 9, rng)
    s = solver_for_organ(o, ctx)
    s.solve(algorithm="GaussJordan")
    assert s.last_stats["setup_ms"] > 0.0  # band assembled and factored
    rp, ci, val, rhs = organ_system(o)
    A = to_scipy(rp, ci, val)
    free = np.ones(o.ndof, bool)
    for d in range(3):
        free[3 * o.fixed_nodes - 3 + d] = False
    # a new right-hand side with the same matrix: substitution only, no band assembly
    rhs2 = 0.5 * rhs
    rhs2[free] += 1.0e-3 * np.abs(rhs).max() * rng.standard_normal(int(free.sum()))
    s.setRHS(rhs2)
    x2 = s.solve(algorithm="GaussJordan")
    st = s.last_stats
    say("reuse: setup ms", st["setup_ms"], "relres", st["true_relres"])
    assert st["setup_ms"] < 0.05
    assert relres(A, x2, rhs2) <= 100.0 * rounding_floor(A, x2, rhs2)
    # a new stiffness (E doubled): the band is reassembled and refactored, x halves
    s.setStiffness(o.E * 2.0, o.nu)
    x3 = s.solve(algorithm="GaussJordan")
    say("refactor: setup ms", s.last_stats["setup_ms"], "x3 vs x2/2", np.linalg.norm(x3 - 0.5 * x2) / np.linalg.norm(x2))
    assert s.last_stats["setup_ms"] > 0.0
    assert np.linalg.norm(x3 - 0.5 * x2) <= 1e-10 * np.linalg.norm(x2)
    s.close()


def tridiag(n, diag):
    rows, cols, vals = [], [], []
    for i in range(n):
        for j in (i - 1, i, i + 1):
            if 0 <= j < n:
                rows.append(i)
                cols.append(j + 1)
                vals.append(diag[i] if i == j else -1.0)
    rp = np.concatenate([[1], 1 + np.cumsum(np.bincount(rows, minlength=n))]).astype(np.int64)
    return rp, np.array(cols, np.int32), np.array(vals)


@pytest.mark.gpu
def test_gpu_direct_not_positive_definite(ctx):
    from plantfem_amd import _lib as L
    from plantfem_amd.fem import CRS, FEMSolver
    n = 12  # a DOF-level ordering (no node blocks)
    d = np.full(n, 4.0)
    d[5] = -4.0  # the pivot of row 5 is negative
    s = FEMSolver(ctx)
    s.setCRS(CRS=CRS(*tridiag(n, d)))
    s.setRHS(np.ones(n))
    with pytest.raises(L.PFError) as e:
        s.solve(algorithm="GaussJordan")
    assert e.value.code == L.PF_NOT_SPD
    # the definite matrix on the same pattern is solved
    rp, ci, va = tridiag(n, np.full(n, 4.0))
    s.setCRS(CRS=CRS(rp, ci, va))
    s.setRHS(np.ones(n))
    x = s.solve(algorithm="GaussJordan")
    A = to_scipy(rp, ci, va)
    assert np.linalg.norm(np.ones(n) - A @ x) <= 1e-13 * np.sqrt(n)
    s.close()


@pytest.mark.gpu
def test_gpu_direct_batch_matches_oracle(ctx):
    rng = np.random.default_rng(7)
    organs = small_organs(1) + [B.leaf(30, 21, rng), B.stem(60, rng), B.leaf(20, 44, rng)]
    u = B.union(organs)
    s = B.build_solver(u, ctx)
    x, st = s.solveBatch(algorithm="GaussJordan", refine=1)
    for m, o in enumerate(organs):
        rp, ci, val, rhs = organ_system(o)
        xm = x[u.dof_range(m)]
        info = s.directInfo(m)
        xo = O.direct_solve(rp, ci, val, rhs, info["perm"], info["bandwidth"])
        A = to_scipy(rp, ci, val)
        res = relres(A, xm, rhs)
        ro = relres(A, xo, rhs)
        err = np.linalg.norm(xm - xo) / np.linalg.norm(xo)
        say("mesh", m, o.kind, o.dims, "b", info["bandwidth"], info["ordering"], "relres", res, "gpu",
            st["true_relres"][m], "oracle", ro, "err vs oracle", err, "status", st["status"][m])
        assert st["status"][m] == 0
        assert res <= max(10.0 * ro, 1e-14)
        assert st["true_relres"][m] <= max(10.0 * ro, 1e-14)
        assert err <= (1e-10 if o.kind == "block" else 1e-8)
    s.close()
