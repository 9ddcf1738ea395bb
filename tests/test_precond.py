"""The preconditioned forms of solveFEMSolver (FEMSolverClass.f90:1768-1790) in the oracle:
JacobiPreconditionerCRS, ILUCRS(0), ILU_matvecCRS and bicgstab_CRS_ILU.

Pins: the reference's own known answers (Tutorial/std/ILU_test.f90 factors of a 4x4 and a
3x3 matrix, Tutorial/std/PBiCGSTAB_ILU.f90 solution of a 3x3 system), a hand-worked factor of
a row stored out of column order, library LU and triangular solves for fully populated
patterns (where ILU(0) is exact LU), exact power-of-two scalings, and the solution of a cube
system against the unpreconditioned solver.
"""
import numpy as np
import pytest
import scipy.linalg as sl

from oracle import oracle as O
from plantfem_amd import workloads as W


def crs_dense(A, order=None):
    """1-based CRS of every entry of the dense A (zeros kept, as A%update(i,j,0) stores them);
    order(i) gives the stored column order of row i (default ascending)."""
    n = A.shape[0]
    rp = [1]
    ci, va = [], []
    for i in range(n):
        cols = list(range(n)) if order is None else list(order(i))
        ci += [c + 1 for c in cols]
        va += [A[i, c] for c in cols]
        rp.append(len(ci) + 1)
    return np.array(rp, np.int64), np.array(ci, np.int32), np.array(va, np.float64)


def to_dense(rp, ci, va):
    n = rp.size - 1
    A = np.zeros((n, n))
    for i in range(n):
        for k in range(rp[i] - 1, rp[i + 1] - 1):
            A[i, ci[k] - 1] = va[k]
    return A


def sort_rows(rp, ci, va):
    ci2, va2 = ci.copy(), va.copy()
    for i in range(rp.size - 1):
        a, b = rp[i] - 1, rp[i + 1] - 1
        o = np.argsort(ci[a:b], kind="stable")
        ci2[a:b] = ci[a:b][o]
        va2[a:b] = va[a:b][o]
    return ci2, va2


def test_ilu_reference_4x4():
    # Tutorial/std/ILU_test.f90:10-39: the matrix and its factor as the test lists it
    A = np.array([[1.0, 1.0, 0.0, 3.0], [2.0, 1.0, -1.0, 1.0], [3.0, -1.0, -1.0, 2.0], [-1.0, 2.0, 3.0, -1.0]])
    F = np.array([[1.0, 1.0, 0.0, 3.0], [2.0, -1.0, -1.0, -5.0], [3.0, 4.0, 3.0, 13.0], [-1.0, -3.0, 0.0, -13.0]])
    rp, ci, va = crs_dense(A)
    f = O.ilu0(rp, ci, va)
    assert np.array_equal(to_dense(rp, ci, f), F)


def test_ilu_reference_3x3():
    # Tutorial/std/ILU_test.f90:44-175: [5 6 7; 10 20 23; 15 50 67] -> [5 6 7; 2 8 9; 3 4 10]
    A = np.array([[5.0, 6.0, 7.0], [10.0, 20.0, 23.0], [15.0, 50.0, 67.0]])
    rp, ci, va = crs_dense(A)
    f = O.ilu0(rp, ci, va)
    assert np.array_equal(to_dense(rp, ci, f), [[5, 6, 7], [2, 8, 9], [3, 4, 10]])


def test_ilu_stored_order_hand_worked():
    # row 3 stored as columns (2, 1, 3): k = 2 is eliminated first, then k = 1 adds
    # -A_31 * u_12 to the already scaled multiplier of column 2:
    #   row 2: A_21 = 4/2 = 2, u_22 = 5 - 2 = 3, u_23 = 2 - 2 = 0
    #   row 3: A_32 = 3/3 = 1; a_33 = 9 - 1*0 = 9; A_31 = 6/2 = 3; L_32 = 1 - 3*1 = -2;
    #          a_33 = 9 - 3*1 = 6
    A = np.array([[2.0, 1.0, 1.0], [4.0, 5.0, 2.0], [6.0, 3.0, 9.0]])
    rp, ci, va = crs_dense(A, order=lambda i: [1, 0, 2] if i == 2 else [0, 1, 2])
    f = O.ilu0(rp, ci, va)
    assert np.array_equal(to_dense(rp, ci, f)[2], [3.0, -2.0, 6.0])
    # in ascending order the same matrix gives the textbook factor (L_32 = 0)
    rp, ci, va = crs_dense(A)
    assert np.array_equal(to_dense(rp, ci, O.ilu0(rp, ci, va))[2], [3.0, 0.0, 6.0])


@pytest.mark.parametrize("n", [5, 12])
def test_ilu_full_pattern_is_lu(n):
    rng = np.random.default_rng(n)
    A = rng.standard_normal((n, n))
    A += np.diag(np.abs(A).sum(axis=0) + 1.0)  # column dominant: partial pivoting keeps P = I
    P, Lf, Uf = sl.lu(A)
    assert np.array_equal(P, np.eye(n))
    rp, ci, va = crs_dense(A)
    F = to_dense(rp, ci, O.ilu0(rp, ci, va))
    np.testing.assert_allclose(np.tril(F, -1), Lf - np.eye(n), rtol=1e-12, atol=1e-12)
    np.testing.assert_allclose(np.triu(F), Uf, rtol=1e-12, atol=1e-12)


def test_ilu_apply_is_lower_solve_and_diagonal():
    # ILU_matvecCRS's backward loop leaves the vector unchanged: y = D^-1 L^-1 v, not U^-1 L^-1 v
    n = 9
    rng = np.random.default_rng(3)
    A = rng.standard_normal((n, n)) + n * np.eye(n)
    rp, ci, va = crs_dense(A)
    f = O.ilu0(rp, ci, va)
    F = to_dense(rp, ci, f)
    v = rng.standard_normal(n)
    y = O.ilu_apply(rp, ci, f, v)
    z = sl.solve_triangular(np.tril(F, -1) + np.eye(n), v, lower=True, unit_diagonal=True)
    np.testing.assert_allclose(y, z / np.diag(F), rtol=1e-12, atol=1e-14)
    assert not np.allclose(y, np.linalg.solve(np.triu(F), z))  # the full LU solve differs


def test_ilu_apply_zero_diagonal_divides_by_one():
    rp = np.array([1, 2, 4], np.int64)
    ci = np.array([2, 1, 2], np.int32)     # row 1 stores no diagonal
    f = np.array([7.0, 0.5, 4.0])
    y = O.ilu_apply(rp, ci, f, np.array([3.0, 2.0]))
    assert np.array_equal(y, [3.0, (2.0 - 0.5 * 3.0) / 4.0])


def test_jacobi_scaling_exact():
    # power-of-two diagonals make every scaling exact; the LAST stored diagonal entry is used
    # (getCRSval, FEMSolverClass.f90:2488-2503) and zero-diagonal rows are left alone
    rp = np.array([1, 3, 6, 8], np.int64)
    ci = np.array([1, 2, 2, 1, 2, 3, 1], np.int32)
    va = np.array([4.0, 3.0, 1.0, 5.0, 8.0, 0.0, 6.0])
    rhs = np.array([2.0, 16.0, 9.0])
    v2, b2 = O.jacobi_scale(rp, ci, va, rhs)
    assert np.array_equal(v2, [1.0, 0.75, 0.125, 0.625, 1.0, 0.0, 6.0])
    assert np.array_equal(b2, [0.5, 2.0, 9.0])


def test_pbicgstab_ilu_reference_3x3():
    # Tutorial/std/PBiCGSTAB_ILU.f90:10-31: x should be (0.54, -0.35, 0.20)
    A = np.array([[5.0, 6.0, 7.0], [10.0, 20.0, 23.0], [15.0, 50.0, 67.0]])
    rp, ci, va = crs_dense(A)
    x, st, it, rr, rr1 = O.bicgstab_ilu(rp, ci, va, np.array([2.0, 3.0, 4.0]), itrmax=10000, er0=1e-13,
                                        relative_er=1e-13)
    assert st == 0
    np.testing.assert_allclose(x, [0.54, -0.35, 0.20], rtol=1e-10, atol=1e-12)


def cube_system(n, sigma=1.0, sorted_cols=True):
    w = W.box(n, sigma=sigma, seed=0, nu_range=(0.25, 0.35))
    rp, ci, val, rhs, fixed, ubar = O.build_system(w)
    v2, b2 = O.apply_dirichlet(rp, ci, val, rhs, fixed, ubar)
    if sorted_cols:
        ci, v2 = sort_rows(rp, ci, v2)
    return rp, ci, v2, b2


def test_bicgstab_ilu_sorted_cube_matches_plain():
    rp, ci, va, b = cube_system(8)
    x0, st0, it0, _, _ = O.bicgstab(rp, ci, va, b)
    x1, st1, it1, _, _ = O.bicgstab_ilu(rp, ci, va, b)
    assert st0 == 0 and st1 == 0
    assert it1 < it0 / 2  # the preconditioner pays on a sorted pattern
    res = np.linalg.norm(b - O.spmv(rp, ci, va, x1)) / np.linalg.norm(b)
    assert res < 1e-9
    np.testing.assert_allclose(x1, x0, rtol=0, atol=1e-8 * np.abs(x0).max())


def test_bicgstab_ilu_early_return():
    # "if (c1 < er0) return": r.r below er0 before any iteration returns x0 untouched
    rp, ci, va = crs_dense(np.array([[2.0, 0.0], [0.0, 2.0]]))
    x, st, it, rr, _ = O.bicgstab_ilu(rp, ci, va, np.array([1e-6, 0.0]), er0=1e-10)
    assert st == 0 and it == 0 and np.array_equal(x, [0.0, 0.0]) and rr == 1e-12


def test_ilu_first_appearance_cube_breaks_down():
    # On ZeroMatrix's first-appearance column order the stored-order elimination mixes
    # scaled multipliers with unscaled entries; its M^-1 grows without bound (the reference
    # behaves the same way: its ILU is meant for sorted rows, SparseClass.f90:1661-1662)
    rp, ci, va, b = cube_system(4, sigma=0.5, sorted_cols=False)
    f = O.ilu0(rp, ci, va)
    y = O.ilu_apply(rp, ci, f, b)
    assert not np.all(np.isfinite(y)) or np.abs(y).max() > 1e20
    ci_s, va_s = sort_rows(rp, ci, va)
    ys = O.ilu_apply(rp, ci_s, O.ilu0(rp, ci_s, va_s), b)
    assert np.abs(ys).max() < 1.0
