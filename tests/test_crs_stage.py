"""CRS product with x staged in LDS per row block (pf_system_spmv kind 4) against the CRS row
kernel (kind 1, sub_crs_matvec, LinearSolverClass.f90:3728-3776): the same products summed in
the same order per lane, so the results must be bitwise equal; both against the oracle's
product.  Patterns: the hex8 ZeroMatrix (every block staged), and a general pattern with empty
rows and blocks, repeated columns, a block above the staging's entry limit and one above its
distinct-column limit (those blocks gather x through the column indices)."""
import sys

import numpy as np
import pytest

from oracle import oracle as O
from plantfem_amd import _lib as L
from plantfem_amd import workloads as W
from plantfem_amd.fem import CRS, FEMDomain, FEMSolver

pytestmark = pytest.mark.gpu


def say(*a):
    print(*a, flush=True)
    sys.stdout.flush()


def test_staged_product_box(ctx):
    w = W.box(12, sigma=1.0, seed=8)
    rp, ci, val, rhs, fixed, ub = O.build_system(w)
    s = FEMSolver(ctx)
    s.setDomain(FEMDomain(w.nodcoord, w.elemnod, ctx))
    s.setStiffness(w.E, w.nu)
    for seed in (3, 4):
        x = np.random.default_rng(seed).standard_normal(rhs.size)
        say("staged spmv box, seed", seed)
        y1, _ = s.spmv(1, x)
        y4, _ = s.spmv(4, x)
        assert np.array_equal(y1, y4)
        ref = O.spmv(rp, ci, val, x)
        assert np.abs(y4 - ref).max() <= 1e-13 * np.abs(ref).max()
    # new values on the same pattern: the staging is rebuilt and stays exact
    s.setStiffness(2.0 * w.E, w.nu)
    x = np.random.default_rng(5).standard_normal(rhs.size)
    y1, _ = s.spmv(1, x)
    y4, _ = s.spmv(4, x)
    assert np.array_equal(y1, y4)
    s.close()


def general_pattern(n, seed):
    rng = np.random.default_rng(seed)
    lens = rng.integers(0, 120, n)
    lens[rng.random(n) < 0.1] = 0       # empty rows
    lens[256:384] = 0                   # empty row blocks
    lens[500] = 9000                    # one row above the staging's entry limit
    lens[1024:1280] = 40                # ~40 distinct columns per row over the whole range
    cols = []
    for i, m in enumerate(lens):
        if 1024 <= i < 1280:
            c = rng.integers(1, n + 1, m)
        else:
            lo = max(1, i - 300)
            c = rng.integers(lo, min(n, i + 300) + 1, m)
            if m > 4:
                c[: m // 5] = c[0]      # repeated columns within the row
        cols.append(c)
    rp = np.concatenate([[1], 1 + np.cumsum(lens)]).astype(np.int64)
    ci = np.concatenate(cols).astype(np.int32)
    val = rng.standard_normal(ci.size)
    return rp, ci, val


def test_staged_product_general_pattern(ctx):
    n = 20000
    rp, ci, val = general_pattern(n, 11)
    s = FEMSolver(ctx)
    s.setCRS(CRS=CRS(rp, ci, val))
    x = np.random.default_rng(12).standard_normal(n)
    say("staged spmv general pattern: n", n, "nnz", ci.size)
    y1, _ = s.spmv(1, x)
    y4, _ = s.spmv(4, x)
    assert np.array_equal(y1, y4)
    ref = O.spmv(rp, ci, val, x)
    assert np.abs(y4 - ref).max() <= 1e-12 * np.abs(ref).max()
    assert np.all(y4[np.diff(rp) == 0] == 0.0)
    s.close()


def test_block_check_accepts_and_rejects(ctx):
    """The DOF=3 node-block check that gates the blocked layout (every node's three rows share one
    list of aligned column triples, the layout ZeroMatrix(DOF=3) produces): the hex8 pattern
    passes and the blocked product equals the CRS one; each broken variant is refused with an
    error instead of being read as blocks."""
    w = W.box(4, sigma=0.5, seed=2)
    rp, ci, val, rhs, fixed, ub = O.build_system(w)
    n = rp.size - 1
    x = np.random.default_rng(7).standard_normal(n)

    def blocked(rp_, ci_, val_):
        s = FEMSolver(ctx)
        s.setCRS(CRS=CRS(rp_, ci_, val_))
        try:
            y2, _ = s.spmv(2, x)
            y1, _ = s.spmv(1, x)
            return y2, y1
        finally:
            s.close()

    y2, y1 = blocked(rp, ci, val)
    assert np.abs(y2 - y1).max() <= 1e-13 * np.abs(y1).max()
    i = 7  # a node; its rows 3i, 3i+1, 3i+2 (0-based)
    r0, r1, r2 = rp[3 * i] - 1, rp[3 * i + 1] - 1, rp[3 * i + 2] - 1
    variants = {}
    c = ci.copy()                                   # a triple out of order in one row
    c[r1], c[r1 + 1] = ci[r1 + 1], ci[r1]
    variants["swapped"] = (rp, c, val)
    c = ci.copy()                                   # the middle row holds another node's triple
    j = 3 if ci[r1] != ci[r1 + 3] else 6
    c[r1:r1 + 3] = ci[r1 + j:r1 + j + 3]
    variants["foreign"] = (rp, c, val)
    c = ci.copy()                                   # a misaligned triple in all three rows
    for rs in (r0, r1, r2):
        c[rs:rs + 3] = (ci[r0:r0 + 3] % n) + 1 if ci[r0 + 2] < n else ci[r0:r0 + 3] - 1
    variants["misaligned"] = (rp, c, val)
    keep = np.ones(ci.size, bool)                   # the first row of the node one triple short
    keep[r0:r0 + 3] = False
    rp_short = rp.copy()
    rp_short[3 * i + 1:] -= 3
    variants["short"] = (rp_short, ci[keep], val[keep])
    for name, (rp_, ci_, val_) in variants.items():
        say("block check variant", name)
        with pytest.raises(L.PFError):
            blocked(rp_, ci_, val_)


def test_staged_layout_option_errors(ctx):
    """spmv_kind 4 (CRS rows with x staged) is an FP64 layout without ILU; kind 3 is no solver
    layout.  Each refusal is an error before any kernel runs, and the system stays usable."""
    w = W.box(5, sigma=0.5, seed=3)
    s = FEMSolver(ctx)
    s.setDomain(FEMDomain(w.nodcoord, w.elemnod, ctx))
    s.setStiffness(w.E, w.nu)
    s.addTraction(w.traction_box, w.pressure)
    s.fix(1, w.fixed_dofs, FixValue=0.0)
    for kw in ({"spmv_kind": 3}, {"spmv_kind": 4, "precision": 32}, {"spmv_kind": 4, "preconditioning": "ILU"},
               {"spmv_kind": 7}):
        say("refused", kw)
        with pytest.raises(L.PFError):
            s.solve(**kw)
    s.relative_er = 1e-12
    x4 = s.solve(spmv_kind=4)
    assert s.last_stats["spmv_kind_used"] == 4 and s.last_stats["status"] == 0
    x1 = s.solve(spmv_kind=1)
    assert np.linalg.norm(x4 - x1) <= 1e-9 * np.linalg.norm(x1)
    s.close()
