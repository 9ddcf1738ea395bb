"""Element-matrix kernels of the stiffness assembly: the reference's operation order (kind 0,
bitwise) against the isotropic split with the contraction on the vector units (kind 1) or on
the FP64 matrix cores (kind 2).

The split rests on an identity of B^T D B for the isotropic D of DMatrix
(FEMDomainClass.f90:11607-11636): per node pair (a, b),
    Ke(3a+d, 3b+e) = lam M(d,e) + mu M(e,d) + mu [d==e] tr M,  M(p,q) = sum_gp detJ G(p,a) G(q,b).
The CPU test pins it against the oracle's element matrix (StiffnessMatrixFEMDomain,
FEMDomainClass.f90:11457-11515) for both J modes; the GPU tests compare the kernels' CRS values
with the oracle's assembly within 1e-13 of each row's largest entry (the split sums in another
order than the reference's matmul chain, so agreement is to rounding, not bitwise).
"""
import sys

import numpy as np
import pytest

from oracle import oracle as O
from plantfem_amd import _lib as L
from plantfem_amd import workloads as W
from plantfem_amd.fem import FEMDomain, FEMSolver, cube3d

GPV = 0.57735026918962576
SIGNS = np.array([[-1, -1, -1], [1, -1, -1], [1, 1, -1], [-1, 1, -1],
                  [-1, -1, 1], [1, -1, 1], [1, 1, 1], [-1, 1, 1]], dtype=float)


def say(*a):
    print(*a, flush=True)
    sys.stdout.flush()


def gauss_G(X, jmode):
    """G (3 x 8) and det J at the 8 Gauss points in the table order of ShapeFunctionClass.f90:503-529."""
    out = []
    for gp in range(8):
        xi = GPV * SIGNS[gp]
        t = 1.0 + SIGNS * xi  # (8, 3): 1 + s_a xi_k
        dN = np.stack([SIGNS[:, 0] * 0.125 * t[:, 1] * t[:, 2],
                       SIGNS[:, 1] * 0.125 * t[:, 0] * t[:, 2],
                       SIGNS[:, 2] * 0.125 * t[:, 0] * t[:, 1]])
        J = dN @ X
        Ji = np.linalg.inv(J)
        G = Ji.T @ dN if jmode == 0 else Ji @ dN
        out.append((G, np.linalg.det(J)))
    return out


def ke_split(X, E, nu, jmode, plant=None):
    """Ke by the isotropic split; plant names a mistake to make instead (for the tests that
    show the check can fail)."""
    lam = nu * E / (1 + nu) / (1 - 2 * nu)
    mu = E / 2 / (1 + nu)
    if plant == "lam_mu":
        lam, mu = mu, lam
    Ke = np.zeros((24, 24))
    for G, dj in gauss_G(X, jmode):
        for a in range(8):
            for b in range(8):
                M = dj * np.outer(G[:, a], G[:, b])
                if plant == "transpose":
                    blk = lam * M.T + mu * M
                else:
                    blk = lam * M + mu * M.T
                if plant != "trace":
                    blk = blk + mu * np.trace(M) * np.eye(3)
                Ke[3 * a:3 * a + 3, 3 * b:3 * b + 3] += blk
    return Ke


@pytest.mark.parametrize("jmode", [0, 1])
def test_isotropic_split_matches_oracle_element(jmode):
    """The split reproduces the reference's Ke on distorted elements, and each planted mistake
    (M and M^T swapped, the trace term dropped, lam and mu swapped) fails the same check."""
    rng = np.random.default_rng(4 + jmode)
    for _ in range(4):
        X = 0.5 * SIGNS + 0.5 + 0.12 * rng.standard_normal((8, 3))
        E, nu = 1e7 * np.exp(rng.standard_normal()), rng.uniform(0.1, 0.45)
        want = O.hex8_stiffness(np.asfortranarray(X), E, nu, jmode)
        scale = np.abs(want).max()
        assert np.abs(ke_split(X, E, nu, jmode) - want).max() <= 1e-13 * scale
        for plant in ("transpose", "trace", "lam_mu"):
            assert np.abs(ke_split(X, E, nu, jmode, plant) - want).max() > 1e-3 * scale, plant


def permuted_mesh(dims, seed, jitter=0.0):
    """A box mesh with permuted node and element numbering and jittered coordinates."""
    rng = np.random.default_rng(seed)
    nod, el = cube3d(*dims)
    perm = rng.permutation(nod.shape[0])
    inv = np.empty_like(perm)
    inv[perm] = np.arange(perm.size)
    nod2 = np.asfortranarray(nod[perm] + jitter * rng.standard_normal(nod.shape) / max(dims))
    el2 = np.asfortranarray((inv[el - 1] + 1)[rng.permutation(el.shape[0])].astype(np.int32))
    return nod2, el2


def _assemble(ctx, nod, el, E, nu, jmode, kind):
    s = FEMSolver(ctx)
    s.setDomain(FEMDomain(nod, el, ctx))
    s.setCRS(DOF=3)
    s.setKeKernel(kind)
    s.setStiffness(E, nu, jmode=jmode)
    return s, s.getCRS()


def _row_scaled_err(rp, a, b):
    n = rp.size - 1
    rows = np.repeat(np.arange(n), np.diff(rp))
    scale = np.zeros(n)
    np.maximum.at(scale, rows, np.abs(b))
    return (np.abs(a - b) / scale[rows]).max()


@pytest.mark.gpu
@pytest.mark.parametrize("jmode", [0, 1])
def test_ke_kernels_assembly(ctx, jmode):
    nod, el = permuted_mesh((6, 5, 4), 3, jitter=0.15)
    rng = np.random.default_rng(2)
    ne = el.shape[0]
    E = 1e7 * np.exp(0.8 * rng.standard_normal(ne))
    nu = rng.uniform(0.25, 0.35, ne)
    rp, ci = O.zero_matrix(el, nod.shape[0], 3)
    want = O.assemble(nod, el, E, nu, rp, ci, jmode)
    vals = {}
    for kind in (0, 1, 2):
        say("ke kernel", kind, "jmode", jmode)
        _, crs = _assemble(ctx, nod, el, E, nu, jmode, kind)
        assert np.array_equal(crs.row_ptr, rp) and np.array_equal(crs.col_idx, ci)
        vals[kind] = crs.val
    assert np.array_equal(vals[0], want)
    for kind in (1, 2):
        err = _row_scaled_err(rp, vals[kind], want)
        say("kind", kind, "row-scaled error", err)
        assert err <= 1e-13


@pytest.mark.gpu
def test_ke_kernels_partial_wave_and_solve(ctx):
    """Element counts that leave the last wavefront's 8-element group partly empty (ne = 7 x 5 x 3
    = 105), and a full solve with the MFMA split against the reference-order solve."""
    w = W.box(7, sigma=0.6, seed=5)
    nod, el = cube3d(7, 5, 3)
    rng = np.random.default_rng(9)
    ne = el.shape[0]
    assert ne % 8 != 0
    E = 1e7 * np.exp(0.5 * rng.standard_normal(ne))
    nu = rng.uniform(0.2, 0.4, ne)
    rp, ci = O.zero_matrix(el, nod.shape[0], 3)
    want = O.assemble(nod, el, E, nu, rp, ci, 0)
    for kind in (1, 2):
        _, crs = _assemble(ctx, nod, el, E, nu, 0, kind)
        assert _row_scaled_err(rp, crs.val, want) <= 1e-13
    xs = {}
    for kind in (0, 2):
        say("solve with ke kernel", kind)
        s = FEMSolver(ctx)
        s.setDomain(FEMDomain(w.nodcoord, w.elemnod, ctx))
        s.setKeKernel(kind)
        s.setStiffness(w.E, w.nu)
        s.addTraction(w.traction_box, w.pressure)
        s.fix(1, w.fixed_dofs, FixValue=0.0)
        s.relative_er = 1e-14
        s.er0 = 1e-30
        xs[kind] = s.solve()
    assert np.linalg.norm(xs[2] - xs[0]) <= 1e-10 * np.linalg.norm(xs[0])


@pytest.mark.gpu
def test_ke_kernel_rejects_unknown_kind(ctx):
    nod, el = cube3d(2, 2, 2)
    s = FEMSolver(ctx)
    s.setDomain(FEMDomain(nod, el, ctx))
    with pytest.raises(L.PFError):
        s.setKeKernel(3)
