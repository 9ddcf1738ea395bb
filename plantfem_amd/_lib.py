"""ctypes binding of include/pfgpu.h (libpfgpu.so, built in-tree by plantfem_amd.build).

There is no CPU fallback: if the library is missing every entry point raises.
"""
from __future__ import annotations

import ctypes as C
import os
import sys
from pathlib import Path

import numpy as np

LIB_PATH = Path(__file__).resolve().parent / "lib" / "libpfgpu.so"

PF_OK = 0
PF_BREAKDOWN_VV = 1
PF_BREAKDOWN_C2C3 = 2
PF_NOT_CONVERGED = 3
PF_INVALID = 4
PF_NOT_SPD = 5
PF_ERR_HIP = -1
PF_ERR_NOMEM = -2
PF_NOALLOC_MAX_ROWS = 524288  # include/pfgpu.h

STATUS_NAMES = {
    PF_OK: "ok",
    PF_BREAKDOWN_VV: "breakdown v.v == 0 (Bicgstab devide by zero)",
    PF_BREAKDOWN_C2C3: "breakdown c2*c3 == 0 (Bicgstab devide by zero)",
    PF_NOT_CONVERGED: "itrmax reached",
    PF_INVALID: "invalid input",
    PF_NOT_SPD: "non-positive pivot: the matrix is not symmetric positive definite (A is singular)",
    PF_ERR_HIP: "HIP error",
    PF_ERR_NOMEM: "device out of memory",
}

ER0_DEFAULT = float(np.float32(1.0e-10))  # dble(1.0e-10) of FEMSolverClass.f90:70-71


class SolveOpts(C.Structure):
    _fields_ = [
        ("itrmax", C.c_int32),
        ("precision", C.c_int32),
        ("er0", C.c_double),
        ("relative_er", C.c_double),
        ("inner_tol", C.c_double),
        ("target_relres", C.c_double),
        ("preconditioning", C.c_int32),
        ("spmv_kind", C.c_int32),
        ("check_every", C.c_int32),
        ("use_graph", C.c_int32),
        ("debug", C.c_int32),
        ("stop_mode", C.c_int32),
        ("kernel_timing", C.c_int32),
        ("jacobi", C.c_int32),
        ("inner_max", C.c_int32),
        ("use_x0", C.c_int32),
    ]


class Stats(C.Structure):
    _fields_ = [
        ("status", C.c_int32),
        ("iterations", C.c_int32),
        ("outer_iterations", C.c_int32),
        ("spmv_kind_used", C.c_int32),
        ("rr", C.c_double),
        ("rr1", C.c_double),
        ("true_relres", C.c_double),
        ("solve_ms", C.c_double),
        ("setup_ms", C.c_double),
        ("spmv_p_ms", C.c_double),
        ("spmv_e_ms", C.c_double),
        ("update_ms", C.c_double),
        ("pupdate_ms", C.c_double),
        ("timed_iterations", C.c_int64),
    ]

    def as_dict(self) -> dict:
        return {k: getattr(self, k) for k, _ in self._fields_}


class SystemView(C.Structure):
    _fields_ = [
        ("n", C.c_int64), ("nnz", C.c_int64), ("nn", C.c_int64), ("ne", C.c_int64),
        ("row_ptr", C.c_void_p), ("col_idx", C.c_void_p), ("val", C.c_void_p),
        ("rhs", C.c_void_p), ("x", C.c_void_p), ("nodcoord", C.c_void_p), ("elemnod", C.c_void_p),
    ]


class BatchTune(C.Structure):
    _fields_ = [("layout", C.c_int32), ("sub", C.c_int32), ("block", C.c_int32), ("blocks_per_cu", C.c_int32)]

    def as_dict(self) -> dict:
        return {k: getattr(self, k) for k, _ in self._fields_}


class BatchMeshStats(C.Structure):
    _fields_ = [
        ("status", C.c_int32), ("iterations", C.c_int32), ("replacements", C.c_int32), ("workgroup", C.c_int32),
        ("rr", C.c_double), ("rr1", C.c_double), ("true_relres", C.c_double),
        ("start_ms", C.c_double), ("end_ms", C.c_double),
    ]


# numpy view of an array of BatchMeshStats
BATCH_STATS_DTYPE = np.dtype([
    ("status", np.int32), ("iterations", np.int32), ("replacements", np.int32), ("workgroup", np.int32),
    ("rr", np.float64), ("rr1", np.float64), ("true_relres", np.float64),
    ("start_ms", np.float64), ("end_ms", np.float64),
])


P = C.c_void_p
I32, I64, F64 = C.c_int32, C.c_int64, C.c_double
PI32, PI64, PF64, PF32 = (C.POINTER(C.c_int32), C.POINTER(C.c_int64), C.POINTER(C.c_double),
                          C.POINTER(C.c_float))

# name -> (restype, argtypes); mirrors include/pfgpu.h
SIGNATURES = {
    "pf_solve_opts_default": (None, [C.POINTER(SolveOpts)]),
    "pf_last_error": (C.c_char_p, []),
    "pf_version": (C.c_int, []),
    "pf_ctx_create": (C.c_int, [C.c_int, C.POINTER(P)]),
    "pf_ctx_destroy": (C.c_int, [P]),
    "pf_ctx_synchronize": (C.c_int, [P]),
    "pf_ctx_stream": (P, [P]),
    "pf_crs_graph": (C.c_int, [P, I64, I64, I32, PI32, I32, PI64, PI64, PI32]),
    "pf_hex8_stiffness": (C.c_int, [P, I64, I64, PF64, PI32, PF64, PF64, I32, I64, I64, PF64]),
    "pf_hex8_stiffness_crs": (C.c_int, [P, I64, I64, PF64, PI32, PF64, PF64, I32, PI64, PI32, PF64]),
    "pf_spmv_crs": (C.c_int, [P, I64, PI64, PI32, PF64, PF64, PF64]),
    "pf_bicgstab_crs": (C.c_int, [P, I64, PI64, PI32, PF64, PF64, PF64, C.POINTER(SolveOpts), C.POINTER(Stats)]),
    "pf_dev_spmv_crs": (C.c_int, [P, I64, P, P, P, P, P]),
    "pf_dev_spmv_crs_f32": (C.c_int, [P, I64, P, P, P, P, P]),
    "pf_crs_spmv_real64": (C.c_int, [P, PF64, PI64, PI32, PF64, PF64, I64, I64]),
    "pf_crs_spmv_real32": (C.c_int, [P, PF32, PI64, PI32, PF32, PF32, I64, I64]),
    "pf_dev_bicgstab_crs_noalloc": (C.c_int, [P, I64, P, P, P, P, P, C.POINTER(SolveOpts), C.POINTER(Stats)]),
    "crs_spmv_real64": (None, [PF64, PI64, PI32, PF64, PF64, C.c_size_t, C.c_size_t]),
    "crs_spmv_real32": (None, [PF32, PI64, PI32, PF32, PF32, C.c_size_t, C.c_size_t]),
    "pf_system_create": (C.c_int, [P, C.POINTER(P)]),
    "pf_system_destroy": (C.c_int, [P]),
    "pf_system_set_mesh": (C.c_int, [P, I64, I64, I32, PF64, PI32]),
    "pf_system_zero_matrix": (C.c_int, [P, I32, PI64]),
    "pf_system_set_crs": (C.c_int, [P, I64, I64, PI64, PI32, PF64]),
    "pf_system_sizes": (C.c_int, [P, PI64, PI64, PI64, PI64]),
    "pf_system_get_crs": (C.c_int, [P, PI64, PI32, PF64]),
    "pf_system_stiffness": (C.c_int, [P, PF64, PF64, I32]),
    "pf_system_add_element_matrices": (C.c_int, [P, I64, PI32, PF64]),
    "pf_system_set_rhs": (C.c_int, [P, PF64]),
    "pf_system_add_element_vectors": (C.c_int, [P, I64, PI32, PF64]),
    "pf_system_get_rhs": (C.c_int, [P, PF64]),
    "pf_system_add_traction": (C.c_int, [P, PF64, F64]),
    "pf_system_fix": (C.c_int, [P, I64, PI32, PF64, F64]),
    "pf_system_solve": (C.c_int, [P, C.POINTER(SolveOpts), PF64, PF64, C.POINTER(Stats)]),
    "pf_system_get_x": (C.c_int, [P, PF64]),
    "pf_system_ilu_factor": (C.c_int, [P, PF64, C.POINTER(C.c_int32)]),
    "pf_system_ilu_apply": (C.c_int, [P, PF64, PF64]),
    "pf_system_ilu_reset": (C.c_int, [P]),
    "pf_jacobi_crs": (C.c_int, [P, I64, PI64, PI32, PF64, PF64]),
    "pf_ilu0_crs": (C.c_int, [P, I64, PI64, PI32, PF64, PF64]),
    "pf_ilu_apply_crs": (C.c_int, [P, I64, PI64, PI32, PF64, PF64, PF64]),
    "pf_system_view_get": (C.c_int, [P, C.POINTER(SystemView)]),
    "pf_system_spmv": (C.c_int, [P, I32, PF64, PF64, I32, PF64]),
    "pf_system_add_tractions": (C.c_int, [P, I32, PI64, PF64, PF64]),
    "pf_system_batch_set": (C.c_int, [P, I32, PI64, PF64]),
    "pf_system_batch_solve": (C.c_int, [P, C.POINTER(SolveOpts), C.POINTER(BatchTune), PF64, P,
                                        C.POINTER(Stats)]),
    "pf_system_batch_autotune": (C.c_int, [P, C.POINTER(SolveOpts), I32, C.POINTER(BatchTune), I32,
                                           C.POINTER(BatchTune), PF64, PI32]),
    "pf_hex8_mass_vector": (C.c_int, [P, I64, I64, PF64, PI32, F64, PF64, I64, I64, PF64]),
    "pf_system_add_body_force": (C.c_int, [P, PF64, F64, PF64]),
    "pf_hex8_mass_matrix": (C.c_int, [P, I64, I64, PF64, PI32, PF64, I64, I64, I64, PF64]),
    "pf_system_mass_matrix": (C.c_int, [P, PF64, F64]),
    "pf_system_zeros": (C.c_int, [P]),
    "pf_system_keep_matrix": (C.c_int, [P, C.c_char_p]),
    "pf_overset_gpp": (C.c_int, [P, I64, I64, PF64, PI32, I64, I64, PF64, PI32, I64, PI64, PI32, PI32, PI32, PF64]),
    "pf_overset_p2p": (C.c_int, [P, I64, I64, PF64, PI32, I64, I64, PF64, PI32, I64, PI64, PI32, PI32, PF64]),
    "pf_system_multi_crs": (C.c_int, [P, I32, I32, I64, PI32, PI32, PI32, PI64]),
    "pf_system_ebom": (C.c_int, [P, I64, PI32, PI32, PI32, PF64, F64, I32]),
    "pf_system_fix_eig": (C.c_int, [P, I64, PI32]),
    "pf_system_eig_count": (C.c_int, [P, C.POINTER(C.c_int32)]),
    "pf_system_eig": (C.c_int, [P, C.c_int32, F64, C.c_int32, PF64, PF64, C.POINTER(C.c_int32),
                                C.POINTER(Stats)]),
    "pf_element_stress": (C.c_int, [P, I64, I64, PF64, PI32, PF64, PF64, I64, PF64, I64, I32, PF64]),
    "pf_system_element_stress": (C.c_int, [P, PF64, I64, PF64, I64, I32, PF64]),
    "pf_system_stiffness_dev": (C.c_int, [P, P, P, I32]),
    "pf_system_solve_multi_dev": (C.c_int, [P, I32, P, P, C.POINTER(SolveOpts), P, C.POINTER(Stats)]),
    "pf_system_elem_sensitivity_dev": (C.c_int, [P, I32, P, P, P]),
    "pf_system_solve_direct": (C.c_int, [P, I32, PF64, C.POINTER(Stats)]),
    "pf_system_batch_direct": (C.c_int, [P, I32, PF64, P, C.POINTER(Stats)]),
    "pf_system_direct_info": (C.c_int, [P, I32, PI64, PI32, PI32, PI32, PF64]),
    "pf_system_direct_kernel": (C.c_int, [P, I32]),
    "pf_system_ke_kernel": (C.c_int, [P, I32]),
}


class PFError(RuntimeError):
    def __init__(self, code: int, where: str, msg: str):
        super().__init__(f"{where}: {STATUS_NAMES.get(code, code)} ({code}) {msg}")
        self.code = code


_lib = None


def load() -> C.CDLL:
    """Load libpfgpu.so; raises if it has not been built (no fallback path exists)."""
    global _lib
    if _lib is not None:
        return _lib
    if not LIB_PATH.exists():
        raise RuntimeError(f"{LIB_PATH} is not built; run `python -m plantfem_amd.build` "
                           "(there is no CPU fallback)")
    lib = C.CDLL(str(LIB_PATH))
    for name, (res, args) in SIGNATURES.items():
        fn = getattr(lib, name)
        fn.restype = res
        fn.argtypes = args
    _lib = lib
    return lib


def last_error() -> str:
    return (load().pf_last_error() or b"").decode(errors="replace")


def check(code: int, where: str, ok=(PF_OK,)) -> int:
    if code not in ok:
        raise PFError(code, where, last_error())
    return code


def default_opts(**kw) -> SolveOpts:
    o = SolveOpts()
    load().pf_solve_opts_default(C.byref(o))
    for k, v in kw.items():
        if not hasattr(o, k):
            raise KeyError(k)
        setattr(o, k, v)
    return o


def ptr(a: np.ndarray, ctype):
    """Pointer to a contiguous numpy array (None passes NULL)."""
    if a is None:
        return None
    assert a.flags["C_CONTIGUOUS"] or a.flags["F_CONTIGUOUS"]
    return a.ctypes.data_as(C.POINTER(ctype))


class Context:
    """One device and one HIP stream (pf_ctx)."""

    def __init__(self, device: int = 0):
        # torch's wheel bundles its own HIP/HSA runtime beside the one this library links; in one
        # process torch's must come up first or torch later finds no device, so start it now
        # when torch is in use
        torch = sys.modules.get("torch")
        if torch is not None and torch.cuda.is_available():
            torch.cuda.init()
        lib = load()
        h = P()
        check(lib.pf_ctx_create(int(device), C.byref(h)), "pf_ctx_create")
        self.h = h
        self.device = device

    def synchronize(self):
        check(load().pf_ctx_synchronize(self.h), "pf_ctx_synchronize")

    def close(self):
        if getattr(self, "h", None):
            load().pf_ctx_destroy(self.h)
            self.h = None

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass


_default_ctx: Context | None = None


def default_context() -> Context:
    global _default_ctx
    if _default_ctx is None:
        _default_ctx = Context(int(os.environ.get("PF_DEVICE", "0")))
    return _default_ctx
