"""Benchmark of the static linear-elastic solve on one MI355X.

Workload: the 100^3 hex box (1,030,301 nodes, 3,090,903 DOF, 245,438,109 nonzeros),
E = 1e7 * lognormal(sigma = 1.0) per element, nu ~ U[0.25, 0.35], bottom face clamped,
uniform downward traction 1e4 Pa on the top face (synthetic, seeded).  One step is the
whole device path with the mesh and materials already in HBM: CRS graph build, element
stiffness assembly, traction RHS, Dirichlet elimination and BiCGSTAB to a relative
residual ||b - Ax|| / ||b|| of 1e-8 (FP64).  value = seconds per step (time to solve).

The same line reports: the iteration count under the reference's own stop rule; the
mixed-precision steps (precision 32: FP32 matrix and vectors inside; precision 48: FP32 matrix,
FP64 vectors) and, on this system and on the config 3 system (sigma 2.0 clipped to a 1e4
contrast), every form against a certified FP64 reference at equal scaling (plain FP64, FP64 on
D^-1 A, both mixed forms swept to the first stop target within an error of 1e-8); the SpMV
kernels' bandwidth; the roofline of the dominant kernel (y = A p fused with r0.y, timed live
with hipEvents around every launch); and the CPU baseline (the oracle on the host cores, per
phase, with the residual certification of the references).

    python bench.py [--steps K] [--warmup W] [--size 100] [--no-cpu]

Two further workloads print their own line (not the default):
    python bench.py --workload batch     512 organ meshes (leaf slabs n x m x 2, stems 4 x 4 x L)
                                         in one persistent-kernel launch: the direct banded
                                         Cholesky (meshes/s with every mesh solved, per-mesh
                                         latency, FP64 MFMA roofline), then the BiCGSTAB forms
                                         (autotuned launch shape, DOF-iterations/s)
    python bench.py --workload adjoint   60^3 inverse problem, 64 load cases, 2000 observed surface
                                         nodes: forward + adjoint gradient time, gradient checks
    python bench.py --workload precond   60^3 box with row-sorted columns: unpreconditioned,
                                         PointJacobi and ILU(0) BiCGSTAB (solve(preconditioning=))
    python bench.py --workload modal     mass matrix + eig: the 10 lowest modes of a 20x20x100 bar
                                         (LOBPCG) and every mode of a 10^3 box (dense)
    python bench.py --workload overset   a plant of 17 organ meshes (stem, 8 petioles, 8 leaf blades)
                                         coupled by overset penalties as Soybean%getDisplacement
                                         assembles it: overset search, multi-domain setCRS,
                                         stiffness, setEbOM, gravity, solve
"""
from __future__ import annotations

import argparse
import json
import os
import sys
import time
from pathlib import Path

import numpy as np

ROOT = Path(__file__).resolve().parent
sys.path.insert(0, str(ROOT))

METRIC = "time-to-solve @relres 1e-8, 3M-DOF hex; SpMV GB/s; mixed vs FP64"
HBM_PEAK_GBS = 8000.0


def log(*a):
    print(*a, file=sys.stderr, flush=True)


def spmv_bytes(kind: int, n: int, nnz: int) -> float:
    """Algorithmic HBM bytes of one y = A x (x read once, y written once)."""
    nn = n // 3
    nblocks = nnz // 9
    nslices = (nn + 63) // 64
    if kind in (1, 4):  # CRS: int64 row_ptr, int32 col_idx, FP64 values (4: the same product, x staged)
        return 12.0 * nnz + 8.0 * (n + 1) + 16.0 * n
    if kind == 2:  # blocked FP64: one int32 node index per 3x3 block
        return 8.0 * nnz + 4.0 * nblocks + 12.0 * nslices + 8 + 16.0 * n
    return 4.0 * nnz + 4.0 * nblocks + 12.0 * nslices + 8 + 8.0 * n  # blocked FP32


def main(argv=None):
    ap = argparse.ArgumentParser()
    ap.add_argument("--gpus", type=int, default=1)
    ap.add_argument("--steps", type=int, default=2)
    ap.add_argument("--warmup", type=int, default=1)
    ap.add_argument("--size", type=int, default=100)
    ap.add_argument("--seed", type=int, default=0)
    ap.add_argument("--no-cpu", action="store_true")
    ap.add_argument("--cpu-iters", type=int, default=20, help="oracle iterations timed for the CPU baseline")
    ap.add_argument("--mixed-repeats", type=int, default=5)
    ap.add_argument("--no-c3", action="store_true", help="skip the config 3 (sigma 2.0, 1e4 contrast) leg")
    ap.add_argument("--no-graph", action="store_true", help="direct kernel launches only (for profilers)")
    ap.add_argument("--workload", choices=("box", "batch", "adjoint", "precond", "modal", "overset"), default="box")
    ap.add_argument("--plant-scale", type=int, default=2, help="refinement of the overset plant (1: 23k DOF)")
    ap.add_argument("--plant-kind", choices=("boxes", "organs"), default="boxes",
                    help="overset plant of Cube3D boxes, or of Cylinder stem / box petioles / Leaf3D blades")
    ap.add_argument("--precond-size", type=int, default=60)
    ap.add_argument("--batch-count", type=int, default=512)
    ap.add_argument("--batch-legs", default="direct,fp64,mixed,fp64_pointjacobi",
                    help="batch workload legs to run (comma list; e.g. 'direct' alone for a profile)")
    ap.add_argument("--batch-itrmax", type=int, default=3000,
                    help="iteration budget per mesh of the batched solve")
    ap.add_argument("--adjoint-size", type=int, default=60)
    ap.add_argument("--fd-size", type=int, default=12, help="mesh of the finite-difference gradient check")
    ap.add_argument("--fd-elems", type=int, default=32, help="elements of the difference check (0: no check)")
    ap.add_argument("--fd-rule", choices=("richardson", "central"), default="richardson")
    ap.add_argument("--fd-h", type=float, default=1e-3)
    ap.add_argument("--fd-reduce", type=float, default=1e-11)
    ap.add_argument("--profile", action="store_true",
                    help="short profiling workload: 40 FP64 iterations (blocked, CRS rows, CRS rows with x "
                         "staged), the element-matrix kernels, the SpMV kernels and a 512 MiB device copy "
                         "(the reference kernel for HBM counter calibration)")
    a = ap.parse_args(argv)
    if a.gpus != 1 or int(os.environ.get("WORLD_SIZE", "1")) > 1:
        print("bench.py: this build measures one GPU; the multi-GPU (z-slab sharded) path is a design "
              "note only, so --gpus must be 1", file=sys.stderr)
        sys.exit(2)

    if a.workload == "batch":
        return bench_batch(a)
    if a.workload == "adjoint":
        return bench_adjoint(a)
    if a.workload == "precond":
        return bench_precond(a)
    if a.workload == "overset":
        return bench_overset(a)
    if a.workload == "modal":
        return bench_modal(a)

    from plantfem_amd import _lib as L
    from plantfem_amd import workloads as W
    from plantfem_amd.fem import FEMDomain, FEMSolver

    t0 = time.perf_counter()
    w = W.box(a.size, sigma=1.0, seed=a.seed, nu_range=(0.25, 0.35))
    n = w.ndof
    log(f"workload {w.name}: n={n} ne={w.ne} (mesh built in {time.perf_counter() - t0:.1f}s)")
    ctx = L.default_context()
    s = FEMSolver(ctx)
    s.setDomain(FEMDomain(w.nodcoord, w.elemnod, ctx))
    s.setCRS(DOF=3)
    s.setStiffness(w.E, w.nu)  # materials resident from here on
    ctx.synchronize()
    nnz = s.nnz
    log(f"nnz={nnz}")

    s.itrmax = 40000

    if a.profile:
        s.setCRS(DOF=3)
        s.setStiffness()
        s.addTraction(w.traction_box, w.pressure)
        s.fix(1, w.fixed_dofs, FixValue=0.0)
        s.itrmax = 40
        s.relative_er = 1e-30
        s.solve(stop_mode=1, download=False, use_graph=0, raise_on_breakdown=False)
        log(f"profile: {s.last_stats['iterations']} FP64 iterations")
        s.solve(precision=48, target_relres=1e-30, stop_mode=1, download=False, use_graph=0,
                raise_on_breakdown=False, **P48)
        log(f"profile: {s.last_stats['iterations']} precision 48 iterations")
        for kind in (1, 4):  # the fused iteration on CRS rows, plain and with x staged in LDS
            s.solve(stop_mode=1, download=False, use_graph=0, raise_on_breakdown=False, spmv_kind=kind)
            log(f"profile: {s.last_stats['iterations']} FP64 iterations, spmv kind {s.last_stats['spmv_kind_used']}")
        # element-matrix kernels: the reference-order one (kind 0, in the assembly above) against
        # the isotropic split on the vector units (1) and on the FP64 matrix cores (2)
        for kind in (1, 2, 0):
            s.setKeKernel(kind)
            for _ in range(3):
                t1 = time.perf_counter()
                s.setStiffness()
                ms = (time.perf_counter() - t1) * 1e3
            log(f"profile: stiffness with element kernel {kind} {ms:.2f} ms (host clock, with the row gather)")
        xr = np.random.default_rng(1).standard_normal(n)
        for kind in (1, 4, 2, 3):
            _, ms = s.spmv(kind, xr, repeat=5, download=False)
            log(f"profile: spmv kind {kind} {ms:.3f} ms")
        # HBM counter calibration: a plain 512 MiB device-to-device copy through the HIP
        # runtime's own copy kernel (reads and writes exactly 512 MiB per launch)
        import ctypes as C
        hip = C.CDLL("libamdhip64.so")
        nb = 512 * 1024 * 1024
        src, dst = C.c_void_p(), C.c_void_p()
        assert hip.hipMalloc(C.byref(src), C.c_size_t(nb)) == 0 and hip.hipMalloc(C.byref(dst), C.c_size_t(nb)) == 0
        assert hip.hipMemset(src, 1, C.c_size_t(nb)) == 0
        for _ in range(3):
            assert hip.hipMemcpy(dst, src, C.c_size_t(nb), 3) == 0  # hipMemcpyDeviceToDevice
        assert hip.hipDeviceSynchronize() == 0
        hip.hipFree(src)
        hip.hipFree(dst)
        log("profile: 3 x 512 MiB device copies done")
        return

    def step(precision=64, timing=0, target=1e-8, stop_mode=1, **kw):
        s.setCRS(DOF=3)                       # ZeroMatrix graph build (device)
        s.setStiffness()                      # element stiffness assembly (device)
        s.addTraction(w.traction_box, w.pressure)
        s.fix(1, w.fixed_dofs, FixValue=0.0)
        s.relative_er = target
        s.solve(precision=precision, stop_mode=stop_mode, download=False, kernel_timing=timing,
                target_relres=target, raise_on_breakdown=False, use_graph=0 if a.no_graph else 1, **kw)
        return dict(s.last_stats)

    # ---- FP64 time to solve (timed steps carry hipEvents around every y = A p) ----------
    for i in range(a.warmup):
        st = step(64, 2)
        log(f"warmup {i}: {st['iterations']} it, {st['solve_ms']:.1f} ms, relres {st['true_relres']:.2e}")
    ctx.synchronize()
    times, stats = [], []
    for i in range(a.steps):
        ctx.synchronize()
        t = time.perf_counter()
        st = step(64, 2)
        ctx.synchronize()
        times.append(time.perf_counter() - t)
        stats.append(st)
        log(f"step {i}: {times[-1]:.3f} s, {st['iterations']} it, solve {st['solve_ms']:.1f} ms, "
            f"relres {st['true_relres']:.2e}, K1 {st['spmv_p_ms']:.3f} ms")
    t_step = float(np.median(times))
    st64 = stats[-1]

    # graph-captured solve without events, for the overhead of the timing mode
    t = time.perf_counter()
    st_graph = step(64, 0)
    ctx.synchronize()
    t_graph = time.perf_counter() - t
    # the same solve stopped by the reference's own rule (bicgstab_CRS_2, er0 = relative_er =
    # dble(1.0e-10)), for its iteration count
    st_rc = step(64, 0, target=L.ER0_DEFAULT, stop_mode=0)
    log(f"reference criterion: {st_rc['iterations']} it, relres {st_rc['true_relres']:.2e}")

    # ---- mixed precision on the same matrix --------------------------------------------
    # precision 32 (the FP32 inner BiCGSTAB on D^-1 A, FP32 matrix and vectors, restarted at
    # every FP64 true-residual replacement) and precision 48 (the same with FP64 vectors over
    # the FP32 matrix); each step is the FP64 step's graph build, assembly, BCs and solve
    x64 = s.getX()

    def mixed_steps(prec, target, reps, **kw):
        times, st = [], None
        for i in range(reps):
            ctx.synchronize()
            t = time.perf_counter()
            st = step(prec, 0, target=target, **kw)
            ctx.synchronize()
            times.append(time.perf_counter() - t)
            log(f"precision {prec} to {target:g} ({i}): {times[-1]:.3f} s, status {st['status']}, outer "
                f"{st['outer_iterations']}, inner {st['iterations']}, relres {st['true_relres']:.2e}")
        return float(np.median(times)), st, s.getX()

    reps = max(1, min(a.steps, a.mixed_repeats))
    t_mixed, st32, x32 = mixed_steps(32, 1e-8, reps)
    t_mixed_tight, st32t, x32t = mixed_steps(32, 1e-10, 1)
    t48, st48, x48 = mixed_steps(48, 1e-8, reps, **P48)
    t48t, st48t, x48t = mixed_steps(48, 1e-10, 1, **P48)
    # reference: FP64 BiCGSTAB on the PointJacobi-scaled system to a scaled relres of 1e-13; its
    # residual on the unscaled system is certified on the CPU (oracle elimination + product)
    s.setCRS(DOF=3)
    s.setStiffness()
    s.addTraction(w.traction_box, w.pressure)
    s.fix(1, w.fixed_dofs, FixValue=0.0)
    crs2_pre, rhs2_pre = (s.getCRS(), s.getRHS()) if not a.no_cpu else (None, None)
    s.relative_er = 1e-13
    s.itrmax = 100000
    s.solve(preconditioning="PointJacobi", stop_mode=1, download=False, raise_on_breakdown=False)
    st_ref = dict(s.last_stats)
    x_ref = s.getX()
    s.itrmax = 40000

    def rel(u, ref=x_ref):
        return float(np.linalg.norm(u - ref) / np.linalg.norm(ref))

    mixed_err = rel(x32)
    log(f"mixed: {t_mixed:.3f} s to 1e-8 (err {mixed_err:.2e}), {t_mixed_tight:.3f} s to 1e-10 "
        f"(err {rel(x32t):.2e}); precision 48 {t48:.3f} s (err {rel(x48):.2e}), {t48t:.3f} s to 1e-10 "
        f"(err {rel(x48t):.2e}); FP64 err {rel(x64):.2e}; reference: {st_ref['iterations']} PointJacobi "
        f"iterations, scaled relres {st_ref['true_relres']:.2e}")
    # solve-only comparison on this system at equal scaling, every form swept to the error bar
    comparison2 = compare_forms(ctx, s, lambda: refresh(s, w), x_ref, plain=(t_step, st64, x64))

    # ---- SpMV kernels on the eliminated matrix ----------------------------------------
    refresh(s, w)  # the unscaled system, eliminated by a one-iteration solve
    s.itrmax = 1
    s.solve(stop_mode=1, download=False, raise_on_breakdown=False)
    s.itrmax = 40000
    spmv = {}
    xr = np.random.default_rng(1).standard_normal(n)
    for kind, name in ((1, "crs_f64"), (4, "crs_f64_staged"), (2, "blocked_f64"), (3, "blocked_f32")):
        s.spmv(kind, xr, repeat=3, download=False)
        _, ms = s.spmv(kind, None, repeat=20, download=False)
        gbs = spmv_bytes(kind, n, nnz) / (ms * 1e-3) / 1e9
        spmv[name] = {"ms": round(ms, 4), "GB/s": round(gbs, 1), "frac_hbm": round(gbs / HBM_PEAK_GBS, 3)}
        log(f"spmv {name}: {ms:.3f} ms {gbs:.0f} GB/s")

    # ---- roofline of the dominant kernel: y = A p (+ r0.y), blocked FP64 ---------------
    k1_bytes = spmv_bytes(2, n, nnz) + 8.0 * n  # + r0 read for the fused dot
    k1_ms = float(np.mean([x["spmv_p_ms"] for x in stats]))
    k1_gbs = k1_bytes / (k1_ms * 1e-3) / 1e9
    # measured HBM bytes per K1 launch (rocprofv3 FETCH_SIZE + WRITE_SIZE, copy-kernel calibrated)
    traffic = None
    tfile = ROOT / "profiles" / f"traffic_box{a.size}.json"
    if tfile.exists():
        tr = json.loads(tfile.read_text())
        if tr.get("n_dof") == n and tr.get("nnz") == nnz:
            traffic = tr["bytes_per_launch"]
    roofline = {"bound": "hbm", "achieved": round(k1_gbs, 1), "peak": HBM_PEAK_GBS, "unit": "GB/s",
                "frac": round(k1_gbs / HBM_PEAK_GBS, 4), "traffic": traffic,
                "kernel": "k_spmv_p (blocked FP64 y=A p fused with r0.y)", "bytes_per_launch": k1_bytes,
                "avg_launch_ms": round(k1_ms, 4)}
    s.close()

    # ---- config 3: mixed precision where FP32 robustness is at stake -------------------
    c3 = None if a.no_c3 else bench_c3(a, ctx)

    # ---- CPU baseline (the oracle on the host cores) and the certification of the references
    cpu = None
    if not a.no_cpu:
        cpu = cpu_leg(a, w, st64["iterations"], crs2_pre, rhs2_pre, x_ref, c3)
        del crs2_pre, rhs2_pre
    if c3 is not None:
        c3.pop("_pre", None)

    iters = st64["iterations"]
    line = {
        "metric": METRIC, "value": round(t_step, 4), "unit": "s", "n_gpus": 1, "steps": a.steps,
        "warmup": a.warmup, "ms_per_step": round(1e3 * t_step, 2), "higher_is_better": False,
        "scaling": "weak", "vs_baseline": None, "dtype": "f64", "data": "synthetic",
        "config": {"workload": f"hex box {a.size}^3 (3.09M DOF at 100^3), E=1e7*lognormal(1.0), nu~U[0.25,0.35], "
                               "clamped bottom, 1e4 Pa top traction; graph+assembly+BC+BiCGSTAB to relres 1e-8",
                   "n_dof": n, "nnz": nnz, "seed": a.seed},
        "iterations": iters, "solve_ms": round(st64["solve_ms"], 2), "true_relres": st64["true_relres"],
        "dof_iterations_per_s": round(n * iters / (st64["solve_ms"] * 1e-3), 1),
        "reference_criterion": {"stop": "bicgstab_CRS_2: sqrt(r.r / r1.r1) < dble(1.0e-10) or sqrt(r.r) < "
                                        "dble(1.0e-10) (FEMSolverClass.f90:2275-2289)",
                                "iterations": st_rc["iterations"], "true_relres": st_rc["true_relres"],
                                "solve_ms": round(st_rc["solve_ms"], 2)},
        "solve_ms_graph_mode": round(st_graph["solve_ms"], 2), "step_s_graph_mode": round(t_graph, 4),
        "mixed": {"scheme": "FP32 BiCGSTAB on D^-1 A (row Jacobi), FP64 true residual and solution, restart at "
                            "every replacement (inner tolerance 0.05, at most 300 inner iterations)",
                  "status": st32["status"], "time_to_solve_s": round(t_mixed, 4), "solve_ms": round(st32["solve_ms"], 2),
                  "outer": st32["outer_iterations"], "inner_iterations": st32["iterations"],
                  "true_relres": st32["true_relres"], "rel_err_vs_fp64_reference": mixed_err,
                  "converged": st32["status"] == 0,
                  "speedup_vs_fp64": round(t_step / t_mixed, 3) if st32["status"] == 0 else None,
                  "to_relres_1e-10": {"status": st32t["status"], "time_to_solve_s": round(t_mixed_tight, 4),
                                      "inner_iterations": st32t["iterations"], "outer": st32t["outer_iterations"],
                                      "true_relres": st32t["true_relres"], "rel_err_vs_fp64_reference": rel(x32t),
                                      "speedup_vs_fp64": round(t_step / t_mixed_tight, 3)
                                      if st32t["status"] == 0 else None},
                  "precision48": {"scheme": "the same refinement with FP64 vectors over the FP32 matrix copy "
                                            f"(inner tolerance {P48['inner_tol']}, at most {P48['inner_max']} inner "
                                            "iterations)",
                                  "status": st48["status"], "time_to_solve_s": round(t48, 4),
                                  "inner_iterations": st48["iterations"], "outer": st48["outer_iterations"],
                                  "true_relres": st48["true_relres"], "rel_err_vs_fp64_reference": rel(x48),
                                  "speedup_vs_fp64": round(t_step / t48, 3) if st48["status"] == 0 else None,
                                  "to_relres_1e-10": {"status": st48t["status"], "time_to_solve_s": round(t48t, 4),
                                                      "true_relres": st48t["true_relres"],
                                                      "rel_err_vs_fp64_reference": rel(x48t),
                                                      "speedup_vs_fp64": round(t_step / t48t, 3)
                                                      if st48t["status"] == 0 else None}},
                  "fp64_rel_err_vs_reference": rel(x64),
                  "reference": f"FP64 BiCGSTAB on the PointJacobi-scaled system, {st_ref['iterations']} iterations "
                               f"to scaled relres {st_ref['true_relres']:.2e}; unscaled residual certified in "
                               f"cpu_baseline.certified",
                  "comparison": comparison2},
        "config3": c3,
        "spmv": spmv, "roofline": roofline, "cpu_baseline": cpu,
    }
    print(json.dumps(line), flush=True)


# precision 48's refinement parameters (replacement after a 100x drop of the recursive
# residual, or 1000 inner iterations), chosen on configs 2 and 3 by a sweep on MI355X
P48 = {"inner_tol": 1e-2, "inner_max": 1000}
TARGETS = (1e-8, 1e-9, 1e-10, 1e-11, 1e-12)
ERROR_BAR = 1e-8


def refresh(s, w):
    """graph, assembly, traction and fixes again (the PointJacobi scaling acts in place)"""
    s.setCRS(DOF=3)
    s.setStiffness()
    s.addTraction(w.traction_box, w.pressure)
    s.fix(1, w.fixed_dofs, FixValue=0.0)


def compare_forms(ctx, s, rebuild, x_ref, plain=None):
    """Solve-only times of one system: plain FP64 to relres 1e-8, FP64 on the PointJacobi-scaled
    system D^-1 A (the scaling both mixed forms use), precision 32 and precision 48, the last
    three swept over their stop targets up to the first whose error against x_ref is within
    ERROR_BAR.  `rebuild` restores the unscaled system.  plain = (seconds, stats, x) to reuse."""

    def timed(**kw):
        ctx.synchronize()
        t = time.perf_counter()
        s.solve(download=False, raise_on_breakdown=False, stop_mode=1, **kw)
        ctx.synchronize()
        return time.perf_counter() - t, dict(s.last_stats)

    def err(x):
        return float(np.linalg.norm(x - x_ref) / np.linalg.norm(x_ref))

    out = {"error_bar": ERROR_BAR, "times": "solve() wall seconds: elimination, layouts, scaling and the solve"}
    rebuild()
    s.itrmax = 100000
    if plain is None:
        s.relative_er = 1e-8
        tp, st = timed()
        x = s.getX()
    else:
        tp, st, x = plain
    out["fp64_plain"] = {"relres_target": 1e-8, "s": round(tp, 4), "iterations": st["iterations"],
                         "status": st["status"], "true_relres": st["true_relres"], "rel_err": err(x),
                         "note": "whole step (graph, assembly, BCs, solve)" if plain is not None else None}
    log(f"fp64 plain: {json.dumps(out['fp64_plain'])}")
    legs = []
    for tgt in TARGETS:
        rebuild()
        s.relative_er = tgt
        t, st = timed(preconditioning="PointJacobi")
        leg = {"scaled_relres_target": tgt, "s": round(t, 4), "iterations": st["iterations"], "status": st["status"],
               "rel_err": err(s.getX())}
        legs.append(leg)
        log(f"fp64 PointJacobi: {json.dumps(leg)}")
        if leg["status"] == 0 and leg["rel_err"] <= ERROR_BAR:
            break
    out["fp64_pointjacobi"] = legs
    rebuild()
    s.relative_er = 1e-8
    for prec, name, kw in ((32, "precision32", {}), (48, "precision48", P48)):
        mlegs = []
        for tgt in TARGETS:
            t, st = timed(precision=prec, target_relres=tgt, **kw)
            leg = {"relres_target": tgt, "s": round(t, 4), "status": st["status"], "inner_iterations": st["iterations"],
                   "outer": st["outer_iterations"], "true_relres": st["true_relres"], "rel_err": err(s.getX())}
            mlegs.append(leg)
            log(f"{name}: {json.dumps(leg)}")
            if leg["status"] == 0 and leg["rel_err"] <= ERROR_BAR:
                break
        out[name] = mlegs
    ok_j = [j for j in legs if j["status"] == 0 and j["rel_err"] <= ERROR_BAR]
    summary = {}
    for name in ("precision32", "precision48"):
        ok = [m for m in out[name] if m["status"] == 0 and m["rel_err"] <= ERROR_BAR]
        if ok:
            summary[name] = {"relres_target": ok[0]["relres_target"], "s": ok[0]["s"], "rel_err": ok[0]["rel_err"],
                             "speedup_vs_fp64_same_scaling": round(ok_j[0]["s"] / ok[0]["s"], 3) if ok_j else None,
                             "speedup_vs_fp64_plain_at_1e-8": round(out["fp64_plain"]["s"] / ok[0]["s"], 3)}
        else:
            summary[name] = None
    if ok_j:
        summary["fp64_pointjacobi"] = {"scaled_relres_target": ok_j[0]["scaled_relres_target"], "s": ok_j[0]["s"],
                                       "rel_err": ok_j[0]["rel_err"]}
    out["at_error_bar"] = summary
    return out


def bench_c3(a, ctx):
    """Config 3: the 100^3 box with E = 1e7 lognormal(2.0) clipped to a 1e4 contrast, all forms
    against a tight FP64 reference (FP64 on the PointJacobi-scaled system to 1e-13)."""
    from plantfem_amd import workloads as W
    from plantfem_amd.fem import FEMDomain, FEMSolver
    w3 = W.config(3, seed=a.seed)
    log(f"config 3: E contrast {w3.E.max() / w3.E.min():.3g}")
    s = FEMSolver(ctx)
    s.setDomain(FEMDomain(w3.nodcoord, w3.elemnod, ctx))
    s.setCRS(DOF=3)
    s.setStiffness(w3.E, w3.nu)
    s.addTraction(w3.traction_box, w3.pressure)
    s.fix(1, w3.fixed_dofs, FixValue=0.0)
    crs3_pre, rhs3_pre = (s.getCRS(), s.getRHS()) if not a.no_cpu else (None, None)
    s.itrmax = 100000
    s.relative_er = 1e-13
    s.solve(preconditioning="PointJacobi", stop_mode=1, download=False, raise_on_breakdown=False)
    st = s.last_stats
    x_ref = s.getX()
    ref = {"iterations": st["iterations"], "scaled_relres": st["true_relres"], "status": st["status"]}
    log(f"c3 reference: {json.dumps(ref)}")
    out = {"workload": "hex box 100^3, E=1e7*lognormal(2.0) clipped to a 1e4 contrast, nu~U[0.25,0.35], clamped "
                       "bottom, 1e4 Pa top traction", "e_contrast": float(w3.E.max() / w3.E.min()), "reference": ref}
    out.update(compare_forms(ctx, s, lambda: refresh(s, w3), x_ref))
    s.close()
    out["_pre"] = (crs3_pre, rhs3_pre, w3, x_ref)
    return out


def cpu_leg(a, w, gpu_iters, crs2_pre, rhs2_pre, x_ref2, c3):
    """The oracle on the host cores: config 1 end to end, config 2 per phase (graph, assembly on
    a 100 x 100 x 10 slab scaled by 10, bicgstab_CRS_2 iterations on the same eliminated system),
    a 1-thread figure; and the residual certification of the FP64 references (oracle elimination
    and sub_crs_matvec on the GPU-assembled values, which the GPU tests pin bitwise)."""
    from oracle import oracle as O
    from plantfem_amd import workloads as W
    threads = int(os.environ.get("OMP_NUM_THREADS", "0")) or os.cpu_count()
    try:
        model = next(l.split(":", 1)[1].strip() for l in open("/proc/cpuinfo") if l.startswith("model name"))
    except (OSError, StopIteration):
        model = "unknown"
    out = {"unit": "s", "cores": threads, "kind": "port",
           "host": {"nproc": os.cpu_count(), "cpu_model": model, "threads_used": threads}}

    def certify(crs, rhs, fixed_dofs, x):
        fixed = np.zeros(rhs.size, np.uint8)
        fixed[fixed_dofs - 1] = 1
        v2, r2 = O.apply_dirichlet(crs.row_ptr, crs.col_idx, crs.val, rhs, fixed, np.zeros(rhs.size))
        return float(np.linalg.norm(r2 - O.spmv(crs.row_ptr, crs.col_idx, v2, x, nthreads=threads))
                     / np.linalg.norm(r2)), v2, r2

    cert = {}
    relres2, v2, r2 = certify(crs2_pre, rhs2_pre, w.fixed_dofs, x_ref2)
    cert["config2_reference_relres"] = relres2
    log(f"cpu: config 2 reference certified, relres {relres2:.2e}")
    # config 2 per phase
    t = time.perf_counter()
    O.zero_matrix(w.elemnod, w.nn, 3)
    t_graph = time.perf_counter() - t
    slab = W.box(a.size, sigma=1.0, seed=a.seed, nu_range=(0.25, 0.35), nz=max(1, a.size // 10))
    srp, sci = O.zero_matrix(slab.elemnod, slab.nn, 3)
    t = time.perf_counter()
    O.assemble(slab.nodcoord, slab.elemnod, slab.E, slab.nu, srp, sci)
    t_asm = (time.perf_counter() - t) * w.ne / slab.ne
    del srp, sci
    O.bicgstab(crs2_pre.row_ptr, crs2_pre.col_idx, v2, r2, itrmax=1, nthreads=threads)  # warm
    k = a.cpu_iters
    t = time.perf_counter()
    O.bicgstab(crs2_pre.row_ptr, crs2_pre.col_idx, v2, r2, itrmax=k, er0=0.0, relative_er=0.0, nthreads=threads)
    t_it = (time.perf_counter() - t) / k
    k1 = max(2, k // 10)
    t = time.perf_counter()
    O.bicgstab(crs2_pre.row_ptr, crs2_pre.col_idx, v2, r2, itrmax=k1, er0=0.0, relative_er=0.0, nthreads=1)
    t_it1 = (time.perf_counter() - t) / k1
    del v2, r2
    out["value"] = round(t_graph + t_asm + t_it * gpu_iters, 2)
    out["sample"] = (f"config 2 ({w.ndof} DOF) per phase on {threads} threads: ZeroMatrix graph {t_graph:.2f} s; "
                     f"stiffness assembly (serial ascending-element scatter) {t_asm:.1f} s, timed on a "
                     f"{a.size}x{a.size}x{slab.ne // (a.size * a.size)} slab and scaled by the element count; "
                     f"{k} bicgstab_CRS_2 iterations (OpenMP row SpMV with per-nonzero atomics) "
                     f"{t_it:.3f} s/iteration, times the GPU's {gpu_iters} iterations")
    out["phases"] = {"graph_s": round(t_graph, 3), "assembly_s": round(t_asm, 2),
                     "solve_s_per_iteration": round(t_it, 4), "solve_iterations_timed": k,
                     "solve_s_extrapolated": round(t_it * gpu_iters, 1)}
    out["one_thread"] = {"solve_s_per_iteration": round(t_it1, 4), "iterations_timed": k1,
                         "solve_s_extrapolated": round(t_it1 * gpu_iters, 1)}
    log(f"cpu: config 2 graph {t_graph:.2f} s, assembly {t_asm:.1f} s, {t_it:.3f} s/iteration "
        f"({threads} threads), {t_it1:.3f} s/iteration (1 thread)")
    # config 1 end to end, not extrapolated: graph, assembly, traction, elimination, solve to relres 1e-8
    w1 = W.config(1, seed=a.seed)
    t = time.perf_counter()
    rp, ci, val, rhs, fixed, ub = O.build_system(w1)
    v1, r1 = O.apply_dirichlet(rp, ci, val, rhs, fixed, ub)
    x1, st1, it1, *_ = O.bicgstab(rp, ci, v1, r1, er0=1e-8 * np.linalg.norm(r1), relative_er=0.0, nthreads=threads)
    t_c1 = time.perf_counter() - t
    out["config1_whole_path"] = {"s": round(t_c1, 3), "iterations": it1, "status": st1, "threads": threads,
                                 "stop": "||b - A x|| < 1e-8 ||b|| (recurrence residual)"}
    log(f"cpu: config 1 whole path {t_c1:.2f} s ({it1} iterations)")
    if c3 is not None and "_pre" in c3:
        crs3, rhs3, w3, x3 = c3["_pre"]
        cert["config3_reference_relres"], _, _ = certify(crs3, rhs3, w3.fixed_dofs, x3)
        log(f"cpu: config 3 reference certified, relres {cert['config3_reference_relres']:.2e}")
    out["certified"] = {**cert, "how": "oracle Dirichlet elimination and sub_crs_matvec on the GPU-assembled "
                                       "CRS (bitwise equal to the oracle's assembly), ||b - A x_ref|| / ||b||"}
    return out


def bench_batch(a):
    """512 independent organ meshes solved in one persistent-kernel launch."""
    from plantfem_amd import batch as B
    t = time.perf_counter()
    orgs = B.organ_batch(a.batch_count, seed=a.seed)
    u = B.union(orgs)
    ndof = np.array([o.ndof for o in orgs])
    s = B.build_solver(u)
    log(f"batch: {a.batch_count} organs, {ndof.sum()} DOF ({ndof.min()}..{ndof.max()} per mesh), nnz {s.nnz}, "
        f"setup {time.perf_counter() - t:.2f} s")
    out = {"metric": "batched organ solves: meshes/s (every mesh to relres 1e-8), per-mesh latency",
           "unit": "meshes/s", "n_gpus": 1,
           "higher_is_better": True, "data": "synthetic", "dtype": "f64",
           "config": {"workload": f"{a.batch_count} organ meshes: leaf slabs n x m x 2 (n, m in [20,120], "
                                  "<= 60k DOF) and stems 4 x 4 x L (L in [50,400]); E = 5e6 lognormal(0.5), "
                                  "nu 0.3; clamped at one end/edge; pressure 1e2..1e4 Pa",
                      "n_dof": int(ndof.sum()), "nnz": int(s.nnz), "itrmax_per_mesh": a.batch_itrmax,
                      "stop": "||b - A x|| / ||b|| < 1e-8 per mesh"}}
    legs = set(a.batch_legs.split(","))
    # direct: solve(algorithm="GaussJordan") for every mesh (banded Cholesky, one workgroup per
    # mesh); the first call renumbers the meshes (host, kept per sparsity pattern), the timed
    # call assembles the bands and factors anew (setDirectKernel drops the held factor)
    if "direct" not in legs:
        raise SystemExit("the batch workload line is the direct leg: include 'direct' in --batch-legs")
    t0 = time.perf_counter()
    s.solveBatch(algorithm="GaussJordan", refine=1)
    first_s = time.perf_counter() - t0
    s.setDirectKernel(True)
    x, st = s.solveBatch(algorithm="GaussJordan", refine=1)
    tot = s.last_stats
    info = [s.directInfo(m) for m in range(a.batch_count)]
    nb = np.array([[i["n"], i["bandwidth"]] for i in info], dtype=np.float64)
    flops = nb[:, 0] * nb[:, 1] * (nb[:, 1] + 2.0)  # band Cholesky (dpbtf2) operation count
    ms = tot["setup_ms"] + tot["solve_ms"]
    conv = st["status"] == 0
    lat = st["end_ms"]
    big = int(np.argmax(flops))
    own = st["end_ms"][big] - st["start_ms"][big]
    kinds = {}
    for i in info:
        kinds[i["ordering"]] = kinds.get(i["ordering"], 0) + 1
    leg = {"solve_ms": round(ms, 2), "band_assembly_ms": round(tot["setup_ms"], 2),
           "first_call_s_with_renumbering": round(first_s, 3), "converged": int(conv.sum()),
           "meshes_per_s_converged": round(conv.sum() / (ms * 1e-3), 2),
           "latency_ms_p50": round(float(np.percentile(lat, 50)), 2),
           "latency_ms_p99": round(float(np.percentile(lat, 99)), 2),
           "latency_ms_max": round(float(lat.max()), 2),
           "max_true_relres": float(st["true_relres"].max()), "refinement_steps": 1,
           "orderings": kinds, "max_half_bandwidth": int(nb[:, 1].max()),
           "factor_gflop": round(float(flops.sum()) * 1e-9, 1),
           "largest_mesh": {"n": int(nb[big, 0]), "b": int(nb[big, 1]), "ms": round(float(own), 2),
                            "gflop_per_s_one_workgroup": round(float(flops[big]) / (own * 1e-3) * 1e-9, 1)},
           "roofline": {"bound": "mfma", "achieved": round(float(flops.sum()) / (ms * 1e-3) * 1e-12, 3),
                        "peak": 78.6, "unit": "TFLOP/s",
                        "frac": round(float(flops.sum()) / (ms * 1e-3) * 1e-12 / 78.6, 4)}}
    log(f"batch direct: {json.dumps(leg)}")
    out["direct"] = leg
    for leg_name, prec, pc in (("fp64", 64, None), ("mixed", 32, None), ("fp64_pointjacobi", 64, "PointJacobi")):
        if leg_name not in legs:
            continue
        if pc is not None:  # the scaling is destructive: a fresh system
            s.close()
            s = B.build_solver(u)
        best, table = s.autotuneBatch(iters=40, precision=prec)
        log(f"autotune {leg_name}: best {best}")
        s.itrmax = a.batch_itrmax
        s.relative_er = 1e-8
        s.solveBatch(precision=prec, stop_mode=1, raise_on_breakdown=False, target_relres=1e-8,
                     preconditioning=pc)  # warm
        x, st = s.solveBatch(precision=prec, stop_mode=1, raise_on_breakdown=False, target_relres=1e-8,
                             preconditioning=pc)
        tot = s.last_stats
        ms = tot["solve_ms"]
        conv = st["status"] == 0
        it = st["iterations"].astype(np.int64)
        lat = st["end_ms"]
        row_it = float(np.sum(ndof * it))
        leg = {"launch_shape": best, "solve_ms": round(ms, 2), "converged": int(conv.sum()),
               "meshes_per_s_converged": round(conv.sum() / (ms * 1e-3), 2),
               "dof_iterations_per_s": round(row_it / (ms * 1e-3), 1),
               "iterations_p50": int(np.median(it)), "iterations_max": int(it.max()),
               "latency_ms_p50": round(float(np.percentile(lat, 50)), 2),
               "latency_ms_p99": round(float(np.percentile(lat, 99)), 2),
               "latency_ms_p50_converged": round(float(np.percentile(lat[conv], 50)), 2) if conv.any() else None,
               "max_true_relres_converged": float(st["true_relres"][conv].max()) if conv.any() else None,
               "autotune_ms_per_40_iterations": {f"{t['layout']}/{t['sub']}/{t['block']}": round(v, 2)
                                                 for t, v in table}}
        log(f"batch {leg_name}: {json.dumps(leg)}")
        out[leg_name] = leg
    out["value"] = out["direct"]["meshes_per_s_converged"]
    out["roofline"] = out["direct"]["roofline"]
    out["ms_per_step"] = out["direct"]["solve_ms"]
    print(json.dumps(out), flush=True)


def bench_precond(a):
    """solve(preconditioning=) on a box whose CRS rows hold sorted columns (the order the
    multi-domain setCRS produces; ILU(0) in the reference's stored-order form needs it)."""
    from plantfem_amd import _lib as L
    from plantfem_amd import workloads as W
    from plantfem_amd.fem import CRS, FEMDomain, FEMSolver
    ctx = L.default_context()
    w = W.box(a.precond_size, sigma=1.0, seed=a.seed, nu_range=(0.25, 0.35))
    t = time.perf_counter()
    s0 = FEMSolver(ctx)
    s0.setDomain(FEMDomain(w.nodcoord, w.elemnod, ctx))
    s0.setCRS(DOF=3)
    s0.setStiffness(w.E, w.nu)
    s0.addTraction(w.traction_box, w.pressure)
    crs, rhs = s0.getCRS(), s0.getRHS()
    s0.close()
    rows = np.repeat(np.arange(crs.n, dtype=np.int64), np.diff(crs.row_ptr))
    o = np.lexsort((crs.col_idx, rows))
    sorted_crs = CRS(crs.row_ptr, np.ascontiguousarray(crs.col_idx[o]), np.ascontiguousarray(crs.val[o]))
    del rows, o
    log(f"precond: {a.precond_size}^3 box, n={crs.n} nnz={crs.nnz}, assembled and row-sorted in "
        f"{time.perf_counter() - t:.1f} s")

    def fresh():
        s = FEMSolver(ctx)
        s.setCRS(CRS=sorted_crs)
        s.setRHS(rhs)
        s.fix(1, w.fixed_dofs.astype(np.int32), FixValue=0.0)
        s.relative_er = 1e-8
        return s

    out = {"metric": "time-to-solve @relres 1e-8 with solve(preconditioning=)", "unit": "s", "n_gpus": 1,
           "higher_is_better": False, "data": "synthetic", "dtype": "f64",
           "config": {"workload": f"hex box {a.precond_size}^3, E=1e7*lognormal(1.0), nu~U[0.25,0.35], clamped "
                                  "bottom, 1e4 Pa top traction; CRS rows sorted by column",
                      "n_dof": int(crs.n), "nnz": int(crs.nnz), "stop": "||b - A x|| < 1e-8 ||b||"}}
    legs = {}
    g = 0 if a.no_graph else 1
    for name, pc in (("none", None), ("PointJacobi", "PointJacobi"), ("ILU", "ILU")):
        s = fresh()
        s.solve(preconditioning=pc, stop_mode=1, raise_on_breakdown=False, use_graph=g)  # warm (ILU: factor kept)
        f_levels = s.iluFactor()[1] if pc == "ILU" else None
        setup = s.last_stats["setup_ms"]
        s.solve(preconditioning=pc, stop_mode=1, raise_on_breakdown=False, use_graph=g)
        st = s.last_stats
        leg = {"status": st["status"], "iterations": st["iterations"], "solve_ms": round(st["solve_ms"], 2),
               "ms_per_iteration": round(st["solve_ms"] / max(1, st["iterations"]), 4),
               "true_relres": st["true_relres"], "setup_ms_first_solve": round(setup, 2)}
        if f_levels is not None:
            leg["levels"] = f_levels
            leg["sweep_launches_per_iteration"] = 2 * f_levels
        legs[name] = leg
        log(f"precond {name}: {json.dumps(leg)}")
        if pc == "ILU":
            fval, _ = s.iluFactor()
            ilu_iters = st["iterations"]
        s.close()
    out["legs"] = legs
    out["value"] = legs["ILU"]["solve_ms"] / 1e3
    if not a.no_cpu:
        # oracle bicgstab_CRS_ILU on the same eliminated system with the (bitwise identical)
        # factor from the device, a bounded number of iterations, extrapolated
        from oracle import oracle as O
        s = fresh()
        s.itrmax = 1
        s.solve(stop_mode=1, raise_on_breakdown=False)  # elimination on the device
        e = s.getCRS()
        b = s.getRHS()
        s.close()
        k = 3
        threads = int(os.environ.get("OMP_NUM_THREADS", "0")) or os.cpu_count()
        t = time.perf_counter()
        O.bicgstab_ilu(e.row_ptr, e.col_idx, e.val, b, fval=fval, itrmax=k, nthreads=threads)
        dt = (time.perf_counter() - t) / k
        out["cpu_baseline"] = {"value": round(dt * ilu_iters, 2), "unit": "s", "cores": threads, "kind": "port",
                               "sample": f"{k} bicgstab_CRS_ILU iterations (oracle: OpenMP row SpMV, sequential "
                                         f"ILU_matvec sweeps) timed, x {ilu_iters} GPU iterations; factor excluded"}
    print(json.dumps(out), flush=True)


def bench_modal(a):
    """Modal analysis as Tutorial/fem/ModalAnalysis_FixBoundary.f90 runs it: mass matrix kept as
    "B", stiffness as "A", top and bottom faces removed, eig."""
    from plantfem_amd import _lib as L
    from plantfem_amd.fem import FEMDomain, FEMSolver, cube3d
    ctx = L.default_context()

    def system(dims):
        nod, el = cube3d(*dims, 1.0 * dims[0], 1.0 * dims[1], 1.0 * dims[2])
        ne = el.shape[0]
        s = FEMSolver(ctx)
        s.setDomain(FEMDomain(nod, el, ctx))
        t = time.perf_counter()
        s.setMassMatrix(np.full(ne, 2.7))
        s.keepThisMatrixAs("B")
        s.zeros()
        s.setStiffness(np.full(ne, 70000.0), np.full(ne, 0.3))
        s.keepThisMatrixAs("A")
        ctx.synchronize()
        t_asm = time.perf_counter() - t
        z = nod[:, 2]
        nodes = np.nonzero((z == z.max()) | (z == z.min()))[0] + 1
        s.fix_eig(np.concatenate([3 * nodes - 2, 3 * nodes - 1, 3 * nodes]).astype(np.int32))
        return s, t_asm

    out = {"metric": "modal analysis: time to the lowest modes (LOBPCG) and to every mode (dense)", "unit": "s",
           "n_gpus": 1, "higher_is_better": False, "data": "synthetic", "dtype": "f64",
           "config": {"workload": "Cube3D bar 20x20x100 (unit elements), E 7e4, nu 0.3, rho 2.7, top and bottom "
                                  "faces fixed, 10 lowest modes; and a 10^3 box, every mode"}}
    s, t_asm = system((20, 20, 100))
    s.eig(num_eigen=10)  # warm
    t = time.perf_counter()
    lam, V = s.eig(num_eigen=10)
    t_lobpcg = time.perf_counter() - t
    st = s.last_stats
    out["lobpcg"] = {"n_dof": int(s.n), "assembly_s": round(t_asm, 4), "eig_s": round(t_lobpcg, 4),
                     "iterations": st["iterations"], "status": st["status"], "worst_rel_residual": st["true_relres"],
                     "eigenvalues": [float(v) for v in lam]}
    log(f"modal lobpcg: {json.dumps(out['lobpcg'])}")
    s.close()
    s, _ = system((10, 10, 10))
    s.eig()
    t = time.perf_counter()
    lam, V = s.eig()
    t_dense = time.perf_counter() - t
    out["dense"] = {"n_free": int(lam.size), "eig_s": round(t_dense, 4), "lowest": [float(v) for v in lam[:5]]}
    log(f"modal dense: {json.dumps(out['dense'])}")
    if not a.no_cpu:  # the reference's path: LAPACK (dsygvd via scipy) on the host, same pencil
        import scipy.linalg as sl
        crs = s.getCRS()  # "A" is the current CRS
        n = crs.n
        Ad = np.zeros((n, n))
        rows = np.repeat(np.arange(n), np.diff(crs.row_ptr))
        Ad[rows, crs.col_idx - 1] = crs.val
        s.setMassMatrix(2.7)
        mb = s.getCRS()
        Bd = np.zeros((n, n))
        Bd[rows, mb.col_idx - 1] = mb.val
        keep = np.abs(V).sum(axis=1) > 0  # the free DOFs (eigenvectors vanish at the removed ones)
        t = time.perf_counter()
        ref = sl.eigh(Ad[np.ix_(keep, keep)], Bd[np.ix_(keep, keep)], eigvals_only=False)
        t_cpu = time.perf_counter() - t
        out["dense"]["max_rel_eigenvalue_diff_vs_lapack"] = float(np.max(np.abs(ref[0] - lam) / np.abs(ref[0])))
        out["cpu_baseline"] = {"value": round(t_cpu, 4), "unit": "s", "cores": os.cpu_count(), "kind": "port",
                               "sample": f"LAPACK dsygvd (scipy.linalg.eigh) on the same {int(keep.sum())}-DOF "
                                         "reduced pencil, all pairs with vectors (the reference's default eig "
                                         "calls LAPACK DSPGVD)"}
    out["value"] = out["lobpcg"]["eig_s"]
    print(json.dumps(out), flush=True)


def bench_overset(a):
    """Soybean%getDisplacement's sequence (SoybeanClass.f90:8340-8533) on a synthetic plant: overset
    (GPP) of every petiole onto the stem and every leaf blade onto its petiole, init(NumDomain),
    setDomain, multi-domain setCRS(DOF=3), per-element stiffness (E 1000, nu 0.3: the Soybean
    defaults), setEbOM(penalty 1e7, the default), gravity (rho 1), base of the stem fixed,
    solve (BiCGSTAB, FP64, relres 1e-8)."""
    from oracle import oracle as O
    from plantfem_amd import _lib as L
    from plantfem_amd import fem
    from plantfem_amd import workloads as W
    ctx = L.default_context()
    P = W.plant_organs(a.plant_scale) if a.plant_kind == "organs" else W.plant_scaled(a.plant_scale)
    E0, nu0, pen, rho = 1000.0, 0.3, 1.0e7, 1.0
    stem_xyz = P.meshes[0][0]
    base = np.nonzero(stem_xyz[:, 2] <= stem_xyz[:, 2].min())[0]
    fix_ids = (3 * base[:, None] + np.arange(1, 4)).ravel().astype(np.int32)

    def run(keep=False):
        doms = [fem.FEMDomain(n, e, ctx) for n, e in P.meshes]
        ctx.synchronize()
        t = [time.perf_counter()]
        for my, yours in P.pairs:
            doms[my - 1].overset(doms[yours - 1], DomainID=yours, MyDomainID=my)
        t.append(time.perf_counter())
        s = fem.FEMSolver(ctx)
        s.init(NumDomain=len(doms))
        s.setDomain(FEMDomainPointers=doms)
        s.setCRS(DOF=3)
        ctx.synchronize()
        t.append(time.perf_counter())
        s.setStiffness([E0] * len(doms), [nu0] * len(doms))
        ctx.synchronize()
        t.append(time.perf_counter())
        s.setEbOM(penalty=pen, DOF=3)
        s.addBodyForce(Density=rho, Accel=(0.0, 0.0, -9.8))
        s.fix(IDs=fix_ids, FixValue=0.0)
        ctx.synchronize()
        t.append(time.perf_counter())
        x = s.solve(use_graph=0 if a.no_graph else 1)
        t.append(time.perf_counter())
        phases = dict(zip(("overset", "setCRS", "stiffness", "ebom_load_fix", "solve"), np.diff(t)))
        return s, doms, x, phases, t[-1] - t[0]

    for _ in range(max(a.warmup, 1)):
        s, doms, x, _, _ = run()
        s.close()
    times, phases = [], []
    for _ in range(a.steps):
        s, doms, x, ph, tt = run()
        times.append(tt)
        phases.append(ph)
        st = dict(s.last_stats)
        if len(times) < a.steps:
            s.close()
    k = int(np.argmin(times))
    out = {"metric": "overset multi-domain static solve: time from organ meshes to displacement", "unit": "s",
           "value": round(float(np.median(times)), 4), "n_gpus": 1, "steps": a.steps, "warmup": a.warmup,
           "higher_is_better": False, "data": "synthetic", "dtype": "f64",
           "config": {"workload": f"plant x{a.plant_scale}: stem + 8 petioles + 8 leaf blades (17 domains: "
                                  + ("Cube3D boxes" if a.plant_kind == "boxes" else P.meta["organs"])
                                  + f", {P.meta['elements']} hex8, {P.meta['dof']} DOF), 16 overset pairs (GPP), "
                                  "E 1000, nu 0.3, penalty 1e7, gravity, stem base fixed, relres 1e-8"},
           "phases_s": {k2: round(float(v), 4) for k2, v in phases[k].items()},
           "solve": {"iterations": st["iterations"], "status": st["status"], "true_relres": st["true_relres"]}}
    ncon = sum(d.numOversetElements() for d in doms)
    crs = s.getCRS()
    out["system"] = {"connects": int(ncon), "nnz": int(crs.nnz), "n": int(crs.n)}
    # roofline of the solve's dominant kernel (y = A p fused with r0.y), per-launch hipEvents
    # over 2000 iterations of the same system
    s.itrmax = 2000
    s.solve(kernel_timing=1, download=False, raise_on_breakdown=False, use_graph=0)
    stt = s.last_stats
    kind = int(stt["spmv_kind_used"])
    k1_bytes = spmv_bytes(kind, int(crs.n), int(crs.nnz)) + 8.0 * crs.n
    k1_gbs = k1_bytes / (stt["spmv_p_ms"] * 1e-3) / 1e9
    out["roofline"] = {"bound": "hbm", "achieved": round(k1_gbs, 1), "peak": HBM_PEAK_GBS, "unit": "GB/s",
                       "frac": round(k1_gbs / HBM_PEAK_GBS, 4), "traffic": None,
                       "kernel": "k_spmv_p (" + {1: "CRS rows", 2: "blocked", 4: "CRS rows, x staged in LDS"}[kind]
                                 + " FP64 y=A p fused with r0.y)",
                       "bytes_per_launch": k1_bytes, "avg_launch_ms": round(stt["spmv_p_ms"], 4)}
    # the CRS products standalone on the coupled matrix: row kernel and x staged in LDS
    xr = np.random.default_rng(1).standard_normal(int(crs.n))
    out["spmv"] = {}
    for kind, name in ((1, "crs_f64"), (4, "crs_f64_staged")):
        s.spmv(kind, xr, repeat=3, download=False)
        _, ms = s.spmv(kind, None, repeat=20, download=False)
        gbs = spmv_bytes(kind, int(crs.n), int(crs.nnz)) / (ms * 1e-3) / 1e9
        out["spmv"][name] = {"ms": round(ms, 4), "GB/s": round(gbs, 1), "frac_hbm": round(gbs / HBM_PEAK_GBS, 3)}
    log(f"overset: {json.dumps(out)}")
    # parity of the assembled system with the oracle (pattern and values bitwise)
    nn = [m[0].shape[0] for m in P.meshes]
    ne = [m[1].shape[0] for m in P.meshes]
    noff, eoff = np.concatenate([[0], np.cumsum(nn)]), np.concatenate([[0], np.cumsum(ne)])
    xyz = np.asfortranarray(np.concatenate([m[0] for m in P.meshes]))
    en = np.asfortranarray(np.concatenate([m[1] + noff[i] for i, m in enumerate(P.meshes)]).astype(np.int32))
    t0 = time.perf_counter()
    own, gps, prs, poss = [], [], [], []
    for my, yours in P.pairs:
        me, gp, pe, pos = O.overset_gpp(*P.meshes[my - 1], *P.meshes[yours - 1])
        own.append(me + eoff[my - 1])
        gps.append(gp)
        prs.append(pe + eoff[yours - 1])
        poss.append(pos)
    t_ov = time.perf_counter() - t0
    own, gps, prs, poss = (np.concatenate(v) for v in (own, gps, prs, poss))
    t0 = time.perf_counter()
    rp, ci = O.multi_crs(xyz.shape[0], en, True, own, prs)
    t_crs = time.perf_counter() - t0
    t0 = time.perf_counter()
    val = O.assemble(xyz, en, np.full(en.shape[0], E0), np.full(en.shape[0], nu0), rp, ci)
    t_asm = time.perf_counter() - t0
    t0 = time.perf_counter()
    val = O.ebom(xyz, en, rp, ci, val, own, gps, prs, poss, pen)
    t_ebom = time.perf_counter() - t0
    s2 = fem.FEMSolver(ctx)
    s2.init(NumDomain=len(doms))
    s2.setDomain(FEMDomainPointers=doms)
    s2.setCRS(DOF=3)
    s2.setStiffness([E0] * len(doms), [nu0] * len(doms))
    s2.setEbOM(penalty=pen, DOF=3)
    c2 = s2.getCRS()
    out["parity"] = {"connects_equal": bool(own.size == ncon),
                     "pattern_bitwise": bool(np.array_equal(c2.row_ptr, rp) and np.array_equal(c2.col_idx, ci)),
                     "values_bitwise": bool(np.array_equal(c2.val, val))}
    s2.close()
    if not a.no_cpu:
        b = O.body_force(xyz, en, rho, [0.0, 0.0, -9.8])
        fixed = np.zeros(3 * xyz.shape[0], np.uint8)
        fixed[fix_ids - 1] = 1
        v2, b2 = O.apply_dirichlet(rp, ci, val, b, fixed, np.zeros(b.size))
        nth = os.cpu_count() or 1
        t0 = time.perf_counter()
        O.bicgstab(rp, ci, v2, b2, itrmax=a.cpu_iters * 10, er0=1e-300, relative_er=1e-300, nthreads=nth)
        per_it = (time.perf_counter() - t0) / (a.cpu_iters * 10)
        cpu_total = t_ov + t_crs + t_asm + t_ebom + per_it * st["iterations"]
        out["cpu_baseline"] = {"value": round(cpu_total, 3), "unit": "s", "cores": nth, "kind": "port",
                               "sample": f"oracle overset ({t_ov:.2f} s), setCRS ({t_crs:.2f} s), stiffness "
                                         f"({t_asm:.2f} s), setEbOM ({t_ebom:.2f} s) in full; BiCGSTAB timed over "
                                         f"{a.cpu_iters * 10} iterations ({per_it * 1e3:.2f} ms each) and scaled "
                                         f"to the GPU's {st['iterations']}"}
    s.close()
    print(json.dumps(out), flush=True)


def bench_adjoint(a):
    """Forward + adjoint gradient of a squared displacement misfit w.r.t. per-element log E."""
    import torch
    from plantfem_amd import workloads as W
    from plantfem_amd.autograd import ElasticSystem, static_solve

    def setup(nsize, ncases, nobs, rel):
        p = W.inverse_problem(n=nsize, ncases=ncases, nobs=nobs, seed=a.seed)
        sysm = ElasticSystem(p.nodcoord, p.elemnod, p.fixed_dofs, p.nu, relative_er=rel)
        F = torch.tensor(p.F, device="cuda")
        sysm.assemble(torch.tensor(p.E_hidden, device="cuda"))
        U = sysm.solve(F)
        obs = torch.tensor(p.obs_dofs, device="cuda")
        g = torch.Generator().manual_seed(a.seed)
        D = U[obs] * (1.0 + p.noise * torch.randn(obs.numel(), F.shape[1], generator=g,
                                                  dtype=torch.float64).to("cuda"))
        return p, sysm, F, obs, D

    def loss_grad(sysm, logE, F, obs, D):
        logE = logE.clone().requires_grad_(True)
        U = static_solve(logE, F, sysm)
        r = U[obs] - D
        Lval = (r * r).sum()
        Lval.backward()
        return float(Lval.detach()), logE.grad

    t = time.perf_counter()
    p, sysm, F, obs, D = setup(a.adjoint_size, 64, 2000, 1e-8)
    n, ne = sysm.n, p.elemnod.shape[0]
    log(f"adjoint: {a.adjoint_size}^3 mesh, n={n}, {ne} elements, 64 load cases, 2000 observed nodes "
        f"(setup + synthetic observations {time.perf_counter() - t:.1f} s)")
    logE0 = torch.full((ne,), float(np.log(1e7)), dtype=torch.float64, device="cuda")
    loss_grad(sysm, logE0, F, obs, D)  # warm
    torch.cuda.synchronize()
    t = time.perf_counter()
    Lval, g = loss_grad(sysm, logE0, F, obs, D)
    torch.cuda.synchronize()
    t_grad = time.perf_counter() - t
    st = sysm.last_stats
    fwd_it = sum(c["iterations"] for c in st["cases"])
    log(f"adjoint: loss {Lval:.6e}, gradient in {t_grad:.3f} s, adjoint-solve iterations (sum over cases) {fwd_it}")
    # gradient checks on a smaller mesh: central differences of the loss and a second solve
    # with the same tolerance; a single element of a 216k-element mesh moves the loss by
    # ~1e-9 relative, below what any iterative solve resolves, so the difference check runs
    # where one element's effect is resolvable
    # gradient check: base gradient from tight solves, then central differences of the loss with
    # every perturbed solve started from the unperturbed solution (its residual reduced by
    # fd_reduce); the loss differences are formed as sum (U+ - U-)(U+ + U- - 2D) so nothing
    # cancels; "richardson" combines steps h and 2h as (4 D(h) - D(2h)) / 3
    fd_check = None  # --fd-elems 0: no difference check (a trace of the 60^3 solves alone)
    if a.fd_elems > 0:
        tight = 1e-13 if a.fd_size <= 20 else 1e-12
        p2, s2, F2, obs2, D2 = setup(a.fd_size, 64, min(2000, 6 * a.fd_size * a.fd_size), tight)
        s2.opts["itrmax"] = 40000
        rng = np.random.default_rng(a.seed)
        ne2 = p2.elemnod.shape[0]
        logE2 = torch.tensor(np.log(1e7) + 0.3 * rng.standard_normal(ne2), device="cuda")
        t_fd = time.perf_counter()
        L2, g2 = loss_grad(s2, logE2, F2, obs2, D2)
        g2 = g2.cpu().numpy()
        s2.assemble(torch.exp(logE2))
        U20 = s2.solve(F2)
        h = a.fd_h

        def solve_at(lp):
            s2.assemble(torch.exp(lp))
            return s2.solve(F2, x0=U20, stop_mode=0, relative_er=a.fd_reduce, er0=1e-300)[obs2]

        def dL(e, hh):  # (L(+hh) - L(-hh)) / (2 hh) for element e
            lp, lm = logE2.clone(), logE2.clone()
            lp[e] += hh
            lm[e] -= hh
            up, um = solve_at(lp), solve_at(lm)
            return float(((up - um) * (up + um - 2.0 * D2)).sum()) / (2 * hh)

        rels = []
        for k, e in enumerate(rng.choice(ne2, size=min(a.fd_elems, ne2), replace=False)):
            fd = (4 * dL(e, h) - dL(e, 2 * h)) / 3 if a.fd_rule == "richardson" else dL(e, h)
            rels.append(abs(fd - g2[e]) / abs(g2[e]))
            if a.fd_size > 20:
                log(f"adjoint: fd element {k} ({int(e)}): adjoint {g2[e]:.9e} fd {fd:.9e} rel {rels[-1]:.2e}")
        t_fd = time.perf_counter() - t_fd
        log(f"adjoint: central differences ({a.fd_rule}, h {h}) on {len(rels)} elements of the {a.fd_size}^3 mesh: "
            f"max rel err {max(rels):.2e} ({t_fd:.0f} s)")
        fd_check = {"mesh": f"{a.fd_size}^3", "elements": len(rels), "h": h, "rule": a.fd_rule,
                    "base_relres": tight, "perturbed_solve_residual_reduction": a.fd_reduce,
                    "max_rel_err": max(rels), "seconds": round(t_fd, 1)}
    out = {"metric": "adjoint gradient of a displacement misfit w.r.t. per-element log E", "value": round(t_grad, 4),
           "unit": "s per loss+gradient", "n_gpus": 1, "higher_is_better": False, "dtype": "f64",
           "data": "synthetic",
           "config": {"workload": f"{a.adjoint_size}^3 hex box, 64 random top-face traction fields, 2000 observed "
                                  "surface nodes from a hidden lognormal E + 1% noise, loss = squared misfit",
                      "n_dof": n, "params": ne, "load_cases": 64, "relres": 1e-8},
           "loss": Lval, "grad_norm": float(torch.linalg.norm(g)),
           "adjoint_solve_iterations_sum": fwd_it,
           "dof_iterations_per_s": round(2 * n * fwd_it / t_grad, 1),
           "fd_check": fd_check}
    print(json.dumps(out), flush=True)


if __name__ == "__main__":
    main()
