"""Multi-domain assembly with overset penalties in the oracle: FEMDomain%overset (GPP),
Mesh%nearestElementID / InsideOfElement, getLocalCoordinate, the multi-domain setCRS pattern
and setEbOM (FEMDomainClass.f90:14606-14700, :13014-13100, :10046-10100; MeshClass.f90:8611-8866;
FEMSolverClass.f90:466-676, :1891-1985).

Pins: affine hexahedra whose inside set and local coordinates are known in closed form;
structured boxes where the containing element of a point follows from its coordinates; a
set-based restatement of the COO merge; an explicit dense assembly of (Bc^T Bc) detJ; and the
coupling's null space (rigid translations and, between axis-aligned meshes, linear fields)."""
import itertools

import numpy as np
import pytest

from oracle import oracle as O

CORNERS = np.array([[-1, -1, -1], [1, -1, -1], [1, 1, -1], [-1, 1, -1],
                    [-1, -1, 1], [1, -1, 1], [1, 1, 1], [-1, 1, 1]], dtype=np.float64)
G = 1.0 / np.sqrt(3.0)
GP = np.array([[-G, -G, -G], [G, -G, -G], [G, G, -G], [-G, G, -G],
               [-G, -G, G], [G, -G, G], [G, G, G], [-G, G, G]])


def N(xi):
    return 0.125 * (1 + CORNERS[:, 0] * xi[0]) * (1 + CORNERS[:, 1] * xi[1]) * (1 + CORNERS[:, 2] * xi[2])


def box(xn, yn, zn, lx, ly, lz, origin):
    xyz, en = O.cube3d(xn, yn, zn, lx, ly, lz)
    return np.asfortranarray(xyz + np.asarray(origin)), en


def containing_elements(xyz, en, p):
    """Elements of an axis-aligned box mesh whose closed node box holds p."""
    X = xyz[en - 1]
    lo, hi = X.min(axis=1), X.max(axis=1)
    return np.nonzero(np.all((lo <= p) & (p <= hi), axis=1))[0] + 1


def test_inside_hex8_affine():
    rng = np.random.default_rng(0)
    for _ in range(40):
        M = np.eye(3) + 0.4 * rng.standard_normal((3, 3))
        if np.linalg.det(M) < 0.2:
            continue
        c = rng.standard_normal(3)
        X = c + CORNERS @ M.T                      # x = c + M xi
        for _ in range(50):
            xi = rng.uniform(-1.6, 1.6, 3)
            if np.any(np.abs(np.abs(xi) - 1.0) < 1e-6):
                continue
            assert O.inside_hex8(X, c + M @ xi) == bool(np.all(np.abs(xi) <= 1.0))


def test_gauss_point_position():
    X = np.array([0.5, -1.0, 2.0]) + 0.5 * (CORNERS + 1) * np.array([2.0, 0.5, 4.0])
    for g in range(8):
        np.testing.assert_allclose(O.gauss_point_position(X, g + 1), N(GP[g]) @ X, rtol=0, atol=1e-14)


def test_local_coordinate_box_and_transpose_quirk():
    rng = np.random.default_rng(1)
    h = np.array([0.8, 1.3, 0.5])
    c = np.array([1.0, -2.0, 0.25])
    X = c + 0.5 * CORNERS * h
    for _ in range(20):
        xi = rng.uniform(-1, 1, 3)
        p = c + 0.5 * h * xi
        got = O.local_coordinate(X, p)
        np.testing.assert_allclose(got, xi, rtol=0, atol=1e-14)
        np.testing.assert_allclose(N(got) @ X, p, rtol=0, atol=1e-14)
    # sheared element: the reference applies inverse(J) with J(i,j) = dx_j/dxi_i = M^T, so it
    # returns M^-T (x - c) where the true local coordinate is M^-1 (x - c)
    M = np.array([[1.0, 0.3, 0.0], [0.0, 1.0, 0.2], [0.1, 0.0, 1.0]])
    X = c + CORNERS @ M.T
    p = c + M @ np.array([0.2, -0.4, 0.5])
    np.testing.assert_allclose(O.local_coordinate(X, p), np.linalg.solve(M.T, p - c), rtol=1e-13, atol=1e-14)
    assert not np.allclose(O.local_coordinate(X, p), [0.2, -0.4, 0.5])


def test_nearest_element_structured():
    xyz, en = box(5, 4, 3, 4.0, 3.2, 2.4, (0.37, -0.11, 0.53))
    rng = np.random.default_rng(2)
    lo, hi = xyz.min(axis=0), xyz.max(axis=0)
    for _ in range(200):
        p = rng.uniform(lo - 0.3, hi + 0.3)
        got = O.nearest_element(xyz, en, p)
        if np.any(p < lo) or np.any(p > hi):
            assert got == -1
        else:
            assert got in containing_elements(xyz, en, p)


def overlapping_pair():
    A = box(4, 3, 3, 4.0, 3.0, 3.0, (0.0, 0.0, 0.0))
    B = box(5, 5, 4, 4.0, 4.0, 3.2, (2.37, 0.91, 1.13))
    return A, B


def test_overset_gpp_matches_brute_force():
    (xa, ea), (xb, eb) = overlapping_pair()
    me, gp, pe, pos = O.overset_gpp(xa, ea, xb, eb)
    lo, hi = xb.min(axis=0), xb.max(axis=0)
    exp = []
    for e in range(ea.shape[0]):
        X = xa[ea[e] - 1]
        for g in range(8):
            p = N(GP[g]) @ X
            if np.all((lo <= p) & (p <= hi)):
                cand = containing_elements(xb, eb, p)
                assert cand.size == 1  # the offsets keep Gauss points off the other mesh's faces
                exp.append((e + 1, g + 1, cand[0], p))
    assert len(exp) > 20 and len(exp) < 8 * ea.shape[0]
    assert np.array_equal(me, [t[0] for t in exp]) and np.array_equal(gp, [t[1] for t in exp])
    assert np.array_equal(pe, [t[2] for t in exp])
    np.testing.assert_allclose(pos, np.array([t[3] for t in exp]), rtol=0, atol=1e-14)


def joint(A, B):
    (xa, ea), (xb, eb) = A, B
    xyz = np.asfortranarray(np.concatenate([xa, xb]))
    en = np.asfortranarray(np.concatenate([ea, eb + xa.shape[0]]).astype(np.int32))
    return xyz, en


def test_multi_crs_pattern_is_sorted_union():
    A, B = overlapping_pair()
    me, gp, pe, pos = O.overset_gpp(*A, *B)
    xyz, en = joint(A, B)
    ne_a = A[1].shape[0]
    own, pair = me, pe + ne_a
    rp, ci = O.multi_crs(xyz.shape[0], en, True, own, pair)
    nbr = [set() for _ in range(xyz.shape[0])]
    for e in range(en.shape[0]):
        for a, b in itertools.product(en[e], en[e]):
            nbr[a - 1].add(b - 1)
    for c in range(len(own)):
        nodes = np.concatenate([en[own[c] - 1], en[pair[c] - 1]]) - 1
        for a, b in itertools.product(nodes, nodes):
            nbr[a].add(b)
    exp_rp, exp_ci = [1], []
    for i in range(xyz.shape[0]):
        cols = sorted(nbr[i])
        for d in range(3):
            exp_ci += [3 * j + dd + 1 for j in cols for dd in range(3)]
            exp_rp.append(len(exp_ci) + 1)
    assert np.array_equal(rp, exp_rp) and np.array_equal(ci, exp_ci)
    # coupled rows are wider than the block-diagonal ones
    rp0, ci0 = O.multi_crs(xyz.shape[0], en, True, own[:0], pair[:0])
    assert ci0.size < ci.size
    # no domain overset: the reference's COO has no rows at all
    rpe, cie = O.multi_crs(xyz.shape[0], en, False, own, pair)
    assert cie.size == 0 and np.all(rpe == 1)


def test_ebom_dense_assembly():
    A, B = overlapping_pair()
    me, gp, pe, pos = O.overset_gpp(*A, *B)
    xyz, en = joint(A, B)
    own, pair = me, pe + A[1].shape[0]
    rp, ci = O.multi_crs(xyz.shape[0], en, True, own, pair)
    pen = 1.0e7
    val = O.ebom(xyz, en, rp, ci, np.zeros(ci.size), own, gp, pair, pos, pen)
    n = 3 * xyz.shape[0]
    dense = np.zeros((n, n))
    for c in range(len(own)):
        X, Y = xyz[en[own[c] - 1] - 1], xyz[en[pair[c] - 1] - 1]
        dN = np.stack([CORNERS[:, k] * 0.125 * np.prod(1 + np.delete(CORNERS, k, 1) * np.delete(GP[gp[c] - 1], k),
                                                        axis=1) for k in range(3)])
        detJ = np.linalg.det(dN @ X)
        h = Y.max(axis=0) - Y.min(axis=0)
        xi = 2.0 * (pos[c] - Y.mean(axis=0)) / h
        cvec = np.concatenate([N(GP[gp[c] - 1]), -N(xi)])
        Bc = np.kron(cvec[None, :], np.eye(3))       # 3 x 48
        Ae = pen * (Bc.T @ Bc) * detJ
        dofs = (3 * (np.concatenate([en[own[c] - 1], en[pair[c] - 1]]) - 1)[:, None] + np.arange(3)).ravel()
        dense[np.ix_(dofs, dofs)] += Ae
    got = np.zeros((n, n))
    for i in range(n):
        got[i, ci[rp[i] - 1:rp[i + 1] - 1] - 1] = val[rp[i] - 1:rp[i + 1] - 1]
    np.testing.assert_allclose(got, dense, rtol=1e-12, atol=1e-12 * np.abs(dense).max())
    assert np.allclose(got, got.T, rtol=0, atol=1e-9 * np.abs(dense).max())
    # null space: rigid translations, and linear fields between these axis-aligned meshes
    K = got
    for d in range(3):
        u = np.zeros(n)
        u[d::3] = 1.0
        assert np.abs(K @ u).max() < 1e-9 * pen
    Gm = np.array([[0.1, -0.2, 0.3], [0.05, 0.2, -0.1], [-0.3, 0.0, 0.15]])
    u = (xyz @ Gm.T).ravel()
    assert np.abs(K @ u).max() < 1e-8 * pen


def test_overset_system_solve():
    # box A clamped at its base, box B overlapping its top and loaded: with a stiff penalty the
    # two displacement fields agree at the coupled Gauss points
    A = box(3, 3, 4, 1.0, 1.0, 2.0, (0.0, 0.0, 0.0))
    B = box(3, 3, 3, 1.0, 1.0, 1.2, (0.05, 0.07, 1.43))
    me, gp, pe, pos = O.overset_gpp(*A, *B)
    xyz, en = joint(A, B)
    own, pair = me, pe + A[1].shape[0]
    rp, ci = O.multi_crs(xyz.shape[0], en, True, own, pair)
    E = np.full(en.shape[0], 1.0e4)
    nu = np.full(en.shape[0], 0.3)
    val = O.assemble(xyz, en, E, nu, rp, ci)
    val = O.ebom(xyz, en, rp, ci, val, own, gp, pair, pos, 1.0e6)
    n = 3 * xyz.shape[0]
    rhs = np.zeros(n)
    top = np.nonzero(xyz[:, 2] > xyz[:, 2].max() - 1e-9)[0]
    rhs[3 * top + 0] = 0.1
    fixed = np.zeros(n, np.uint8)
    base = np.nonzero(xyz[:A[0].shape[0], 2] < 1e-9)[0]
    for d in range(3):
        fixed[3 * base + d] = 1
    x, st, it, _, _ = O.solve(rp, ci, val, rhs, fixed, np.zeros(n), relative_er=1e-12, er0=1e-30)
    assert st == 0
    ux = x.reshape(-1, 3)
    assert ux[top, 0].min() > 0
    for c in range(len(own)):
        ua = N(GP[gp[c] - 1]) @ ux[en[own[c] - 1] - 1]
        Y = xyz[en[pair[c] - 1] - 1]
        ub = N(O.local_coordinate(Y, pos[c])) @ ux[en[pair[c] - 1] - 1]
        assert np.abs(ua - ub).max() < 2e-2 * np.abs(ux).max()


def test_overset_p2p_matches_brute_force():
    (xa, ea), (xb, eb) = overlapping_pair()
    nd, pe, pos = O.overset_p2p(xa, ea, xb, eb)
    lo, hi = xb.min(axis=0), xb.max(axis=0)
    exp = []
    for i in range(xa.shape[0]):
        p = xa[i]
        if np.all((lo <= p) & (p <= hi)):
            cand = containing_elements(xb, eb, p)
            assert cand.size == 1
            exp.append((i + 1, cand[0]))
    assert len(exp) > 5
    assert np.array_equal(nd, [t[0] for t in exp]) and np.array_equal(pe, [t[1] for t in exp])
    assert np.array_equal(pos, xa[nd - 1])


def test_ebom_p2p_dense_and_null_space():
    A, B = overlapping_pair()
    nd, pe, pos = O.overset_p2p(*A, *B)
    xyz, en = joint(A, B)
    pair = pe + A[1].shape[0]
    gp = np.zeros(nd.size, np.int32)
    rp, ci = O.multi_crs(xyz.shape[0], en, True, nd, pair, gp)
    pen = 1.0e5
    val = O.ebom(xyz, en, rp, ci, np.zeros(ci.size), nd, gp, pair, pos, pen)
    n = 3 * xyz.shape[0]
    dense = np.zeros((n, n))
    for c in range(nd.size):
        Y = xyz[en[pair[c] - 1] - 1]
        h = Y.max(axis=0) - Y.min(axis=0)
        xi = 2.0 * (pos[c] - Y.mean(axis=0)) / h
        Bc = np.kron(np.concatenate([[1.0], -N(xi)])[None, :], np.eye(3))   # 3 x 27, no det J
        dofs = (3 * (np.concatenate([[nd[c]], en[pair[c] - 1]]) - 1)[:, None] + np.arange(3)).ravel()
        dense[np.ix_(dofs, dofs)] += pen * (Bc.T @ Bc)
    got = np.zeros((n, n))
    for i in range(n):
        got[i, ci[rp[i] - 1:rp[i + 1] - 1] - 1] = val[rp[i] - 1:rp[i + 1] - 1]
    np.testing.assert_allclose(got, dense, rtol=1e-12, atol=1e-12 * pen)
    for d in range(3):
        u = np.zeros(n)
        u[d::3] = 1.0
        assert np.abs(got @ u).max() < 1e-9 * pen
    Gm = np.array([[0.1, -0.2, 0.3], [0.05, 0.2, -0.1], [-0.3, 0.0, 0.15]])
    assert np.abs(got @ (xyz @ Gm.T).ravel()).max() < 1e-8 * pen
    # a P2P connect couples 9 nodes: its rows gain the paired element's node blocks only
    rp0, ci0 = O.multi_crs(xyz.shape[0], en, True, nd[:0], pair[:0], gp[:0])
    row = 3 * (nd[0] - 1)
    added = set(ci[rp[row] - 1:rp[row + 1] - 1]) - set(ci0[rp0[row] - 1:rp0[row + 1] - 1])
    assert added == {3 * (m - 1) + d + 1 for m in en[pair[0] - 1] for d in range(3)}
