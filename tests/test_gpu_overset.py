"""Multi-domain assembly with overset penalties on the GPU against the oracle: the overset
search (bitwise connect lists), the multi-domain setCRS pattern, stiffness and setEbOM values
(bitwise), per-domain setMatrix, and the coupled solve (Soybean%getDisplacement's sequence,
SoybeanClass.f90:8340-8533)."""
import numpy as np
import pytest

from oracle import oracle as O
from plantfem_amd import fem
from test_overset import box, joint, overlapping_pair

pytestmark = pytest.mark.gpu


def domain(mesh, ctx):
    return fem.FEMDomain(mesh[0], mesh[1], ctx=ctx)


def test_gpu_overset_connects_bitwise(ctx):
    A, B = overlapping_pair()
    da, db = domain(A, ctx), domain(B, ctx)
    da.overset(db, DomainID=2, algorithm=fem.FEMDomain_Overset_GPP, MyDomainID=1)
    me, gp, pe, pos = O.overset_gpp(*A, *B)
    oc = da.OversetConnect
    assert oc.n == me.size > 0
    assert np.array_equal(oc.ElementID, me) and np.array_equal(oc.GaussPointID, gp)
    assert np.array_equal(oc.PairElementID, pe) and np.array_equal(oc.position, pos)
    assert np.array_equal(oc.DomainIDs12, np.tile([1, 2], (me.size, 1)))
    # a second call appends (num_oversetconnect grows)
    db.overset(da, DomainID=1, MyDomainID=2)
    me2, gp2, pe2, pos2 = O.overset_gpp(*B, *A)
    assert np.array_equal(db.OversetConnect.ElementID, me2) and np.array_equal(db.OversetConnect.position, pos2)
    # disjoint meshes: no connects, but the domain counts as overset
    C = box(2, 2, 2, 1.0, 1.0, 1.0, (10.0, 10.0, 10.0))
    dc = domain(C, ctx)
    dc.overset(da, DomainID=1, MyDomainID=3)
    assert dc.numOversetElements() == 0 and dc.OversetConnect is not None


def test_gpu_overset_larger_mesh_bitwise(ctx):
    # several thousand Gauss points against a 7^3-element mesh
    A = box(12, 10, 9, 6.0, 5.0, 4.5, (0.0, 0.0, 0.0))
    B = box(7, 7, 7, 4.2, 4.2, 4.2, (2.61, 1.33, 0.77))
    da = domain(A, ctx)
    da.overset(domain(B, ctx), DomainID=2)
    me, gp, pe, pos = O.overset_gpp(*A, *B)
    oc = da.OversetConnect
    assert oc.n == me.size > 1000
    assert np.array_equal(oc.ElementID, me) and np.array_equal(oc.PairElementID, pe)
    assert np.array_equal(oc.GaussPointID, gp) and np.array_equal(oc.position, pos)


def multi_solver(A, B, ctx, overset=True):
    da, db = domain(A, ctx), domain(B, ctx)
    if overset:
        da.overset(db, DomainID=2, MyDomainID=1)
    s = fem.FEMSolver(ctx)
    s.init(NumDomain=2)
    s.setDomain(FEMDomains=[da, db], DomainIDs=[1, 2])
    s.setCRS(DOF=3)
    return s, da, db


def test_gpu_multi_crs_pattern_bitwise(ctx):
    A, B = overlapping_pair()
    s, da, db = multi_solver(A, B, ctx)
    me, gp, pe, pos = O.overset_gpp(*A, *B)
    xyz, en = joint(A, B)
    rp, ci = O.multi_crs(xyz.shape[0], en, True, me, pe + A[1].shape[0])
    crs = s.getCRS()
    assert np.array_equal(crs.row_ptr, rp) and np.array_equal(crs.col_idx, ci)
    assert np.all(crs.val == 0)
    assert np.array_equal(s.CRS_ID_Starts_From, [1, 1 + 3 * A[0].shape[0]])
    # no domain overset: the reference's empty CRS
    s2, _, _ = multi_solver(A, B, ctx, overset=False)
    c2 = s2.getCRS()
    assert c2.nnz == 0 and np.all(c2.row_ptr == 1)


def test_gpu_multi_stiffness_and_ebom_bitwise(ctx):
    A, B = overlapping_pair()
    s, da, db = multi_solver(A, B, ctx)
    xyz, en = joint(A, B)
    rng = np.random.default_rng(3)
    E = rng.uniform(1e3, 2e3, en.shape[0])
    nu = rng.uniform(0.25, 0.35, en.shape[0])
    s.setStiffness(YoungModulus=E, PoissonRatio=nu)
    crs = s.getCRS()
    me, gp, pe, pos = O.overset_gpp(*A, *B)
    own, pair = me, pe + A[1].shape[0]
    val = O.assemble(xyz, en, E, nu, crs.row_ptr, crs.col_idx)
    assert np.array_equal(crs.val, val)
    s.setEbOM(penalty=1.0e7, DOF=3)
    val = O.ebom(xyz, en, crs.row_ptr, crs.col_idx, val, own, gp, pair, pos, 1.0e7)
    assert np.array_equal(s.getCRS().val, val)


def test_gpu_multi_setmatrix_per_domain_bitwise(ctx):
    # the reference's per-domain setMatrix / setVector loop (SoybeanClass.f90:8454-8470)
    A, B = overlapping_pair()
    s, da, db = multi_solver(A, B, ctx)
    xyz, en = joint(A, B)
    val = np.zeros(s.nnz)
    crs0 = s.getCRS()
    for d, dom in enumerate([da, db], start=1):
        Ke = dom.StiffnessMatrices(E=np.full(dom.ne(), 500.0 * d), v=np.full(dom.ne(), 0.3))
        for e in range(dom.ne()):
            s.setMatrix(DomainID=d, ElementID=e + 1, DOF=3, Matrix=Ke[e])
            eid = e + 1 + (0 if d == 1 else A[1].shape[0])
            assert O.set_element_matrix(en, eid, 3, Ke[e], crs0.row_ptr, crs0.col_idx, val) == 0
    assert np.array_equal(s.getCRS().val, val)
    rhs = np.zeros(s.n)
    for d, dom in enumerate([da, db], start=1):
        mv = np.stack([dom.MassVector(ElementID=e + 1, Density=2.0, Accel=[0.0, 0.0, -9.8])
                       for e in range(dom.ne())])
        s.setVector(DomainID=d, ElementID=np.arange(1, dom.ne() + 1), DOF=3, Vector=mv)
        off = 0 if d == 1 else A[0].shape[0]
        for e in range(dom.ne()):
            nodes = dom.elemnod[e] - 1 + off
            for a in range(8):
                rhs[3 * nodes[a]:3 * nodes[a] + 3] += mv[e, 3 * a:3 * a + 3]
    assert np.allclose(s.getRHS(), rhs, rtol=1e-14, atol=1e-14)


def test_gpu_overset_solve_matches_oracle(ctx):
    A = box(3, 3, 4, 1.0, 1.0, 2.0, (0.0, 0.0, 0.0))
    B = box(3, 3, 3, 1.0, 1.0, 1.2, (0.05, 0.07, 1.43))
    s, da, db = multi_solver(A, B, ctx)
    xyz, en = joint(A, B)
    E = np.full(en.shape[0], 1.0e4)
    nu = np.full(en.shape[0], 0.3)
    s.setStiffness(YoungModulus=[E[:A[1].shape[0]], E[A[1].shape[0]:]], PoissonRatio=[0.3, 0.3])
    s.setEbOM(penalty=1.0e6, DOF=3)
    n = s.n
    rhs = np.zeros(n)
    top = np.nonzero(xyz[:, 2] > xyz[:, 2].max() - 1e-9)[0]
    rhs[3 * top] = 0.1
    s.setRHS(rhs)
    base = np.nonzero(xyz[:A[0].shape[0], 2] < 1e-9)[0]
    ids = (3 * base[:, None] + np.arange(3) + 1).ravel()
    s.fix(IDs=ids, FixValue=0.0)
    s.relative_er, s.er0 = 1e-12, 1e-30
    x = s.solve()
    crs = s.getCRS()
    me, gp, pe, pos = O.overset_gpp(*A, *B)
    rp, ci = O.multi_crs(xyz.shape[0], en, True, me, pe + A[1].shape[0])
    val = O.ebom(xyz, en, rp, ci, O.assemble(xyz, en, E, nu, rp, ci), me, gp, pe + A[1].shape[0], pos, 1.0e6)
    fixed = np.zeros(n, np.uint8)
    fixed[ids - 1] = 1
    x_o, st, it, _, _ = O.solve(rp, ci, val, rhs, fixed, np.zeros(n), relative_er=1e-12, er0=1e-30)
    assert st == 0 and s.last_stats["status"] == 0
    assert np.linalg.norm(x - x_o) / np.linalg.norm(x_o) < 1e-9
    assert s.last_stats["true_relres"] < 1e-10


def test_gpu_overset_leaf3d_on_petiole_bitwise(ctx):
    # a soybean Leaf3D blade (curved, tapered, thickened elements) whose base sits in a box
    # petiole: the inside tests and local coordinates run on non-brick elements
    leaf = fem.FEMDomain(ctx=ctx).create("Leaf3D", x_num=8, y_num=2, z_num=24, x_len=0.1, y_len=0.01, z_len=0.3,
                                         species=fem.PF_GLYCINE_MAX)
    pet = box(4, 4, 10, 0.043, 0.061, 0.12, (-0.0213, -0.0297, -0.0911))
    dp = domain(pet, ctx)
    leaf.overset(dp, DomainID=1, MyDomainID=2)
    me, gp, pe, pos = O.overset_gpp(leaf.nodcoord, leaf.elemnod, *pet)
    oc = leaf.OversetConnect
    assert oc.n == me.size > 50
    assert np.array_equal(oc.ElementID, me) and np.array_equal(oc.PairElementID, pe)
    assert np.array_equal(oc.GaussPointID, gp) and np.array_equal(oc.position, pos)
    s = fem.FEMSolver(ctx)
    s.init(NumDomain=2)
    s.setDomain(FEMDomains=[dp, leaf], DomainIDs=[1, 2])
    s.setCRS(DOF=3)
    s.setStiffness([2000.0, 1000.0], [0.3, 0.3])
    s.setEbOM(penalty=1.0e7, DOF=3)
    crs = s.getCRS()
    xyz, en = joint(pet, (leaf.nodcoord, leaf.elemnod))
    own = me + pet[1].shape[0]
    rp, ci = O.multi_crs(xyz.shape[0], en, True, own, pe)
    assert np.array_equal(crs.row_ptr, rp) and np.array_equal(crs.col_idx, ci)
    E = np.concatenate([np.full(pet[1].shape[0], 2000.0), np.full(leaf.ne(), 1000.0)])
    val = O.assemble(xyz, en, E, np.full(en.shape[0], 0.3), rp, ci)
    val = O.ebom(xyz, en, rp, ci, val, own, gp, pe, pos, 1.0e7)
    assert np.array_equal(crs.val, val)


def test_gpu_overset_p2p_and_mixed_bitwise(ctx):
    # P2P connects (nodes of A in B) next to GPP ones (C onto B) in one system
    A, B = overlapping_pair()
    Cm = box(3, 3, 3, 1.5, 1.5, 1.5, (3.11, 1.43, 2.02))
    da, db, dc = domain(A, ctx), domain(B, ctx), domain(Cm, ctx)
    da.overset(db, DomainID=2, algorithm=fem.FEMDomain_Overset_P2P, MyDomainID=1)
    dc.overset(db, DomainID=2, algorithm="GPP", MyDomainID=3)
    nd, pe, pos = O.overset_p2p(*A, *B)
    oc = da.OversetConnect
    assert oc.n == nd.size > 5 and np.all(oc.GaussPointID == 0)
    assert np.array_equal(oc.NodeID, nd) and np.array_equal(oc.PairElementID, pe) and np.array_equal(oc.position, pos)
    me, gp, pe2, pos2 = O.overset_gpp(*Cm, *B)
    assert np.array_equal(dc.OversetConnect.ElementID, me) and me.size > 0
    s = fem.FEMSolver(ctx)
    s.init(NumDomain=3)
    s.setDomain(FEMDomains=[da, db, dc], DomainIDs=[1, 2, 3])
    s.setCRS(DOF=3)
    s.setStiffness([1e3, 2e3, 3e3], [0.3, 0.3, 0.3])
    s.setEbOM(penalty=1.0e6, DOF=3)
    crs = s.getCRS()
    xyz = np.asfortranarray(np.concatenate([A[0], B[0], Cm[0]]))
    na, nb = A[0].shape[0], B[0].shape[0]
    ea, eb = A[1].shape[0], B[1].shape[0]
    en = np.asfortranarray(np.concatenate([A[1], B[1] + na, Cm[1] + na + nb]).astype(np.int32))
    own = np.concatenate([nd, me + ea + eb]).astype(np.int32)
    gps = np.concatenate([np.zeros(nd.size, np.int32), gp]).astype(np.int32)
    prs = np.concatenate([pe + ea, pe2 + ea]).astype(np.int32)
    poss = np.concatenate([pos, pos2])
    rp, ci = O.multi_crs(xyz.shape[0], en, True, own, prs, gps)
    assert np.array_equal(crs.row_ptr, rp) and np.array_equal(crs.col_idx, ci)
    E = np.concatenate([np.full(ea, 1e3), np.full(eb, 2e3), np.full(Cm[1].shape[0], 3e3)])
    val = O.assemble(xyz, en, E, np.full(en.shape[0], 0.3), rp, ci)
    val = O.ebom(xyz, en, rp, ci, val, own, gps, prs, poss, 1.0e6)
    assert np.array_equal(crs.val, val)


def test_gpu_overset_cylinder_stem_bitwise(ctx):
    # a round stem (Cylinder mesh: the rounded square core and the ring layers) with one box
    # petiole overset into it by GPP and another by P2P
    stem = fem.FEMDomain(ctx=ctx).create("Cylinder", x_num=8, y_num=8, z_num=12, x_len=0.08, y_len=0.08, z_len=0.6)
    pet = box(10, 3, 3, 0.2, 0.03, 0.03, (0.0113, -0.0141, 0.3017))
    pet2 = box(10, 3, 3, 0.2, 0.03, 0.03, (-0.2113, -0.0137, 0.1519))
    dp, dq = domain(pet, ctx), domain(pet2, ctx)
    dp.overset(stem, DomainID=1, MyDomainID=2)
    dq.overset(stem, DomainID=1, MyDomainID=3, algorithm=fem.FEMDomain_Overset_P2P)
    with pytest.raises(ValueError):
        dq.overset(stem, DomainID=1, MyDomainID=3)  # the reference stops: P2P & GPP cannot be mixed
    me, gp, pe, pos = O.overset_gpp(*pet, stem.nodcoord, stem.elemnod)
    nd, pe2, pos2 = O.overset_p2p(*pet2, stem.nodcoord, stem.elemnod)
    assert me.size > 20 and nd.size > 2
    assert np.array_equal(dp.OversetConnect.ElementID, me) and np.array_equal(dp.OversetConnect.PairElementID, pe)
    assert np.array_equal(dq.OversetConnect.NodeID, nd) and np.array_equal(dq.OversetConnect.PairElementID, pe2)
    assert np.array_equal(dp.OversetConnect.position, pos) and np.array_equal(dq.OversetConnect.position, pos2)
    s = fem.FEMSolver(ctx)
    s.init(NumDomain=3)
    s.setDomain(FEMDomains=[stem, dp, dq], DomainIDs=[1, 2, 3])
    s.setCRS(DOF=3)
    s.setStiffness([1000.0, 1000.0, 1000.0], [0.3, 0.3, 0.3])
    s.setEbOM(penalty=1.0e7, DOF=3)
    ns, es = stem.nn(), stem.ne()
    npt, ept = pet[0].shape[0], pet[1].shape[0]
    xyz = np.asfortranarray(np.concatenate([stem.nodcoord, pet[0], pet2[0]]))
    en = np.asfortranarray(np.concatenate([stem.elemnod, pet[1] + ns, pet2[1] + ns + npt]).astype(np.int32))
    own = np.concatenate([me + es, nd + ns + npt]).astype(np.int32)
    gps = np.concatenate([gp, np.zeros(nd.size, np.int32)]).astype(np.int32)
    prs = np.concatenate([pe, pe2]).astype(np.int32)
    rp, ci = O.multi_crs(xyz.shape[0], en, True, own, prs, gps)
    crs = s.getCRS()
    assert np.array_equal(crs.row_ptr, rp) and np.array_equal(crs.col_idx, ci)
    val = O.assemble(xyz, en, np.full(en.shape[0], 1000.0), np.full(en.shape[0], 0.3), rp, ci)
    val = O.ebom(xyz, en, rp, ci, val, own, gps, prs, np.concatenate([pos, pos2]), 1.0e7)
    assert np.array_equal(crs.val, val)
