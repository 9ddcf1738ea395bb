"""pf_dev_bicgstab_crs_noalloc (the allocation-free solve on device arrays) against
pf_bicgstab_crs, the entry point it is a variant of, on two calls in a row, the second smaller."""
import ctypes as C
import sys

import numpy as np
import pytest

from oracle import oracle as O
from plantfem_amd import _lib as L
from plantfem_amd import workloads as W

pytestmark = pytest.mark.gpu


def say(*a):
    print(*a, flush=True)
    sys.stdout.flush()


def eliminated(k, seed):
    w = W.box(k, sigma=0.8, seed=seed)
    rp, ci, val, rhs, fixed, ub = O.build_system(w)
    v2, r2 = O.apply_dirichlet(rp, ci, val, rhs, fixed, ub)
    return rp, ci, v2, r2


def run_noalloc(ctx, rp, ci, val, b, opts, x0=None):
    import torch
    dev = torch.device("cuda", ctx.device)
    t = [torch.from_numpy(np.ascontiguousarray(a)).to(dev) for a in (rp, ci, val, b)]
    tx = torch.from_numpy(np.zeros(b.size) if x0 is None else np.array(x0)).to(dev)
    torch.cuda.synchronize()
    st = L.Stats()
    code = L.load().pf_dev_bicgstab_crs_noalloc(ctx.h, b.size, *(a.data_ptr() for a in t), tx.data_ptr(),
                                                C.byref(opts), C.byref(st))
    ctx.synchronize()
    return code, tx.cpu().numpy(), st.as_dict()


def run_host(ctx, rp, ci, val, b, opts, x0=None):
    x = np.zeros(b.size) if x0 is None else np.array(x0)
    st = L.Stats()
    code = L.load().pf_bicgstab_crs(ctx.h, b.size, L.ptr(rp, C.c_int64), L.ptr(ci, C.c_int32),
                                    L.ptr(np.ascontiguousarray(val), C.c_double),
                                    L.ptr(np.ascontiguousarray(b), C.c_double), L.ptr(x, C.c_double),
                                    C.byref(opts), C.byref(st))
    return code, x, st.as_dict()


def test_noalloc_matches_bicgstab_crs(ctx):
    for k, seed, stop_mode in ((12, 1, 0), (5, 2, 1)):  # the second call is smaller
        rp, ci, val, b = eliminated(k, seed)
        opts = L.default_opts(spmv_kind=1, stop_mode=stop_mode, relative_er=1e-12, er0=1e-30)
        say("noalloc n", b.size, "stop_mode", stop_mode)
        c1, x1, s1 = run_noalloc(ctx, rp, ci, val, b, opts)
        c2, x2, s2 = run_host(ctx, rp, ci, val, b, opts)
        say("iterations", s1["iterations"], s2["iterations"], "relres", s1["true_relres"], s2["true_relres"])
        assert c1 == c2 == L.PF_OK
        assert s1["iterations"] == s2["iterations"] and s1["spmv_kind_used"] == 1
        assert np.linalg.norm(x1 - x2) <= 1e-12 * np.linalg.norm(x2)
        assert s1["true_relres"] < 1e-10


def test_noalloc_start_vector_and_stop_rule(ctx):
    """x0 on entry and the bicgstab_CRS stop rule (stop_mode 2), as the host entry point."""
    rp, ci, val, b = eliminated(6, 3)
    x0 = 1e-6 * np.random.default_rng(4).standard_normal(b.size)
    opts = L.default_opts(stop_mode=2, er0=1e-16, spmv_kind=1)  # the same layout on both sides
    c1, x1, s1 = run_noalloc(ctx, rp, ci, val, b, opts, x0)
    c2, x2, s2 = run_host(ctx, rp, ci, val, b, opts, x0)
    assert c1 == c2 == L.PF_OK and s1["iterations"] == s2["iterations"]
    assert np.linalg.norm(x1 - x2) <= 1e-12 * np.linalg.norm(x2)


def test_noalloc_rejects_what_it_cannot_hold(ctx):
    import torch
    rp, ci, val, b = eliminated(3, 5)
    dev = torch.device("cuda", ctx.device)
    t = [torch.from_numpy(np.ascontiguousarray(a)).to(dev) for a in (rp, ci, val, b)]
    tx = torch.zeros(b.size, dtype=torch.float64, device=dev)
    torch.cuda.synchronize()
    lib = L.load()
    opts = L.default_opts()
    # above the scratch array's row limit: rejected on the host before any device work
    code = lib.pf_dev_bicgstab_crs_noalloc(ctx.h, L.PF_NOALLOC_MAX_ROWS + 1, *(a.data_ptr() for a in t),
                                           tx.data_ptr(), C.byref(opts), None)
    assert code == L.PF_INVALID and "scratch" in L.last_error()
    # a column outside 1..n: found by the device check, nothing solved
    bad = ci.copy()
    bad[7] = b.size + 3
    tb = torch.from_numpy(bad).to(dev)
    torch.cuda.synchronize()
    code = lib.pf_dev_bicgstab_crs_noalloc(ctx.h, b.size, t[0].data_ptr(), tb.data_ptr(), t[2].data_ptr(),
                                           t[3].data_ptr(), tx.data_ptr(), C.byref(opts), None)
    assert code == L.PF_INVALID
    assert torch.count_nonzero(tx).item() == 0
    # options it does not run
    o32 = L.default_opts(precision=32)
    code = lib.pf_dev_bicgstab_crs_noalloc(ctx.h, b.size, *(a.data_ptr() for a in t), tx.data_ptr(),
                                           C.byref(o32), None)
    assert code == L.PF_INVALID
