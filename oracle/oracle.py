"""ctypes wrapper of the CPU oracle (oracle/_build/liboracle.so).

TEST INFRASTRUCTURE ONLY: imported by tests/, __graft_entry__.smoke() and bench.py's
cpu_baseline leg.  The product package never imports this module.
"""
from __future__ import annotations

import ctypes as C
from pathlib import Path

import numpy as np

LIB = Path(__file__).resolve().parent / "_build" / "liboracle.so"
_lib = None

P64, P32, PD, PU8 = C.POINTER(C.c_int64), C.POINTER(C.c_int32), C.POINTER(C.c_double), C.POINTER(C.c_uint8)
I64, I32, D = C.c_int64, C.c_int, C.c_double

_SIG = {
    "orc_cube3d": (None, [I32, I32, I32, D, D, D, PD, P32]),
    "orc_zero_matrix": (I64, [I64, I64, I32, P32, I32, P64, P32]),
    "orc_zero_matrix_legacy": (I64, [I64, I64, I32, P32, I32, P64, P32]),
    "orc_hex8_stiffness": (None, [PD, D, D, I32, PD]),
    "orc_hex8_stiffness_all": (None, [I64, I64, PD, P32, PD, PD, I32, PD]),
    "orc_assemble_stiffness": (I32, [I64, I64, PD, P32, PD, PD, I32, P64, P32, PD]),
    "orc_set_element_matrix": (I32, [I64, I32, P32, I64, I32, PD, P64, P32, PD]),
    "orc_traction_box": (None, [I64, I64, PD, P32, PD, D, PD]),
    "orc_select_box": (I64, [I64, PD, PD, P32]),
    "orc_hex8_mass_vector": (None, [PD, D, PD, PD]),
    "orc_body_force": (None, [I64, I64, PD, P32, PD, PD, PD]),
    "orc_hex8_cell_stress": (None, [PD, PD, D, D, I32, PD]),
    "orc_cell_stress_all": (None, [I64, I64, PD, P32, PD, PD, PD, I32, PD]),
    "orc_apply_dirichlet": (None, [I64, P64, P32, PD, PD, PU8, PD]),
    "orc_spmv": (None, [I64, P64, P32, PD, PD, PD, I32]),
    "orc_bicgstab_crs2": (I32, [I64, P64, P32, PD, PD, PD, I32, D, D, I32, P32, PD, PD, PD]),
    "orc_solve": (I32, [I64, P64, P32, PD, PD, PU8, PD, PD, I32, D, D, I32, P32, PD, PD]),
    "orc_bicgstab_crs": (I32, [I64, P64, P32, PD, PD, PD, I32, D, I32, P32, PD, PD, PD]),
    "orc_crs_spmv_real64": (None, [PD, P64, P32, PD, PD, I64]),
    "orc_crs_spmv_real32": (None, [C.POINTER(C.c_float), P64, P32, C.POINTER(C.c_float), C.POINTER(C.c_float), I64]),
    "orc_assemble_stiffness_rows": (I32, [I64, I64, PD, P32, PD, PD, I32, PU8, P64, P32, PD]),
    "orc_jacobi_scale": (None, [I64, P64, P32, PD, PD]),
    "orc_hex8_mass_matrix": (None, [PD, D, PD]),
    "orc_mass_matrix_crs": (I32, [I64, I64, PD, P32, PD, P64, P32, PD]),
    "orc_ilu0": (None, [I64, P64, P32, PD]),
    "orc_ilu_apply": (None, [I64, P64, P32, PD, PD, PD]),
    "orc_bicgstab_ilu": (I32, [I64, P64, P32, PD, PD, PD, PD, I32, D, D, I32, P32, PD, PD]),
    "orc_gauss_point_position": (None, [PD, I32, PD]),
    "orc_inside_hex8": (I32, [PD, PD]),
    "orc_nearest_element": (I64, [I64, I64, PD, P32, PD]),
    "orc_overset_gpp": (I64, [I64, I64, PD, P32, I64, I64, PD, P32, P32, P32, P32, PD, I64]),
    "orc_local_coordinate": (None, [PD, PD, PD]),
    "orc_overset_p2p": (I64, [I64, PD, I64, I64, PD, P32, P32, P32, PD, I64]),
    "orc_connect_coeffs_p2p": (None, [PD, PD, PD]),
    "orc_connect_coeffs": (None, [PD, I32, PD, PD, PD, PD]),
    "orc_multi_crs": (I64, [I64, I64, P32, I32, I64, P32, P64, P32]),
    "orc_connect_coeffs_all": (None, [I64, I64, PD, P32, I64, P32, P32, P32, PD, PD, PD]),
    "orc_ebom": (I32, [I64, P64, P32, PD, I64, P32, PD, PD, D]),
    "orc_pbtrf_lower": (I32, [I64, I32, PD, I32]),
    "orc_pbtrs_lower": (None, [I64, I32, PD, I32, PD]),
    "orc_band_from_crs": (I32, [I64, P64, P32, PD, P32, I32, PD, I32]),
}


def load():
    global _lib
    if _lib is None:
        if not LIB.exists():
            raise RuntimeError(f"{LIB} missing: run `make -C oracle`")
        lib = C.CDLL(str(LIB))
        for k, (r, a) in _SIG.items():
            f = getattr(lib, k)
            f.restype = r
            f.argtypes = a
        _lib = lib
    return _lib


def _p(a, t):
    return None if a is None else a.ctypes.data_as(C.POINTER(t))


def _c(a, dt):
    return np.ascontiguousarray(a, dtype=dt)


def _fl(a, dt):
    return np.asfortranarray(a, dtype=dt).ravel(order="F")


def cube3d(xn, yn, zn, lx=1.0, ly=1.0, lz=1.0):
    nn = (xn + 1) * (yn + 1) * (zn + 1)
    ne = xn * yn * zn
    xyz = np.empty(nn * 3)
    en = np.empty(ne * 8, dtype=np.int32)
    load().orc_cube3d(xn, yn, zn, lx, ly, lz, _p(xyz, C.c_double), _p(en, C.c_int32))
    return xyz.reshape(3, nn).T.copy(order="F"), en.reshape(8, ne).T.copy(order="F")


def zero_matrix(elemnod, nn, dof=3, legacy=False):
    en = _fl(elemnod, np.int32)
    ne, nne = elemnod.shape
    f = load().orc_zero_matrix_legacy if legacy else load().orc_zero_matrix
    nnz = f(nn, ne, nne, _p(en, C.c_int32), dof, None, None)
    if nnz < 0:
        raise ValueError("invalid connectivity")
    rp = np.empty(dof * nn + 1, dtype=np.int64)
    ci = np.empty(nnz, dtype=np.int32)
    f(nn, ne, nne, _p(en, C.c_int32), dof, _p(rp, C.c_int64), _p(ci, C.c_int32))
    return rp, ci


def hex8_stiffness(X, E, nu, jmode=0):
    """Ke[r, c] (24x24) of one element with coordinates X (8,3)."""
    x = _c(X, np.float64).ravel()
    ke = np.empty(576)
    load().orc_hex8_stiffness(_p(x, C.c_double), float(E), float(nu), int(jmode), _p(ke, C.c_double))
    return ke.reshape(24, 24).T.copy()


def hex8_stiffness_all(nodcoord, elemnod, E, nu, jmode=0):
    xyz = _fl(nodcoord, np.float64)
    en = _fl(elemnod, np.int32)
    ne = elemnod.shape[0]
    ke = np.empty(ne * 576)
    load().orc_hex8_stiffness_all(nodcoord.shape[0], ne, _p(xyz, C.c_double), _p(en, C.c_int32),
                                  _p(_c(E, np.float64), C.c_double), _p(_c(nu, np.float64), C.c_double), jmode,
                                  _p(ke, C.c_double))
    return ke.reshape(ne, 24, 24).transpose(0, 2, 1)


def assemble(nodcoord, elemnod, E, nu, row_ptr, col_idx, jmode=0):
    xyz = _fl(nodcoord, np.float64)
    en = _fl(elemnod, np.int32)
    val = np.empty(col_idx.size)
    rc = load().orc_assemble_stiffness(nodcoord.shape[0], elemnod.shape[0], _p(xyz, C.c_double), _p(en, C.c_int32),
                                       _p(_c(E, np.float64), C.c_double), _p(_c(nu, np.float64), C.c_double), jmode,
                                       _p(_c(row_ptr, np.int64), C.c_int64), _p(_c(col_idx, np.int32), C.c_int32),
                                       _p(val, C.c_double))
    if rc:
        raise ValueError("assembly: missing CRS slot")
    return val


def set_element_matrix(elemnod, elem, dof, ke, row_ptr, col_idx, val):
    en = _fl(elemnod, np.int32)
    k = np.asfortranarray(ke, dtype=np.float64).ravel(order="F")
    return load().orc_set_element_matrix(elemnod.shape[0], elemnod.shape[1], _p(en, C.c_int32), int(elem), dof,
                                         _p(k, C.c_double), _p(row_ptr, C.c_int64), _p(col_idx, C.c_int32),
                                         _p(val, C.c_double))


def traction_box(nodcoord, elemnod, box, pressure, rhs=None):
    nn = nodcoord.shape[0]
    out = np.zeros(3 * nn) if rhs is None else np.array(rhs, dtype=np.float64)
    load().orc_traction_box(nn, elemnod.shape[0], _p(_fl(nodcoord, np.float64), C.c_double),
                            _p(_fl(elemnod, np.int32), C.c_int32), _p(_c(box, np.float64), C.c_double),
                            float(pressure), _p(out, C.c_double))
    return out


def hex8_mass_vector(X, rho, accel):
    mv = np.empty(24)
    load().orc_hex8_mass_vector(_p(_c(X, np.float64), C.c_double), float(rho),
                                _p(_c(accel, np.float64), C.c_double), _p(mv, C.c_double))
    return mv


def body_force(nodcoord, elemnod, rho, accel, rhs=None):
    nn, ne = nodcoord.shape[0], elemnod.shape[0]
    out = np.zeros(3 * nn) if rhs is None else np.array(rhs, dtype=np.float64)
    rho = np.broadcast_to(np.asarray(rho, dtype=np.float64), (ne,)).copy()
    load().orc_body_force(nn, ne, _p(_fl(nodcoord, np.float64), C.c_double), _p(_fl(elemnod, np.int32), C.c_int32),
                          _p(rho, C.c_double), _p(_c(accel, np.float64), C.c_double), _p(out, C.c_double))
    return out


def cell_stress(nodcoord, elemnod, disp, E, nu, jmode=0):
    """(ne, 3, 3) cell-averaged Cauchy stress per element."""
    nn, ne = nodcoord.shape[0], elemnod.shape[0]
    E = np.broadcast_to(np.asarray(E, dtype=np.float64), (ne,)).copy()
    nu = np.broadcast_to(np.asarray(nu, dtype=np.float64), (ne,)).copy()
    out = np.empty(ne * 9)
    load().orc_cell_stress_all(nn, ne, _p(_fl(nodcoord, np.float64), C.c_double), _p(_fl(elemnod, np.int32), C.c_int32),
                               _p(_c(disp, np.float64), C.c_double), _p(E, C.c_double), _p(nu, C.c_double), jmode,
                               _p(out, C.c_double))
    return out.reshape(ne, 3, 3).transpose(0, 2, 1)


def select_box(nodcoord, box):
    xyz = _fl(nodcoord, np.float64)
    b = _c(box, np.float64)
    m = load().orc_select_box(nodcoord.shape[0], _p(xyz, C.c_double), _p(b, C.c_double), None)
    ids = np.empty(m, dtype=np.int32)
    load().orc_select_box(nodcoord.shape[0], _p(xyz, C.c_double), _p(b, C.c_double), _p(ids, C.c_int32))
    return ids


def apply_dirichlet(row_ptr, col_idx, val, rhs, fixed, ubar):
    val = np.array(val, dtype=np.float64)
    rhs = np.array(rhs, dtype=np.float64)
    load().orc_apply_dirichlet(rhs.size, _p(_c(row_ptr, np.int64), C.c_int64), _p(_c(col_idx, np.int32), C.c_int32),
                               _p(val, C.c_double), _p(rhs, C.c_double), _p(_c(fixed, np.uint8), C.c_uint8),
                               _p(_c(ubar, np.float64), C.c_double))
    return val, rhs


def spmv(row_ptr, col_idx, val, x, nthreads=0):
    n = row_ptr.size - 1
    y = np.empty(n)
    load().orc_spmv(n, _p(_c(row_ptr, np.int64), C.c_int64), _p(_c(col_idx, np.int32), C.c_int32),
                    _p(_c(val, np.float64), C.c_double), _p(_c(x, np.float64), C.c_double), _p(y, C.c_double),
                    nthreads)
    return y


def bicgstab(row_ptr, col_idx, val, b, x0=None, itrmax=100000, er0=None, relative_er=None, nthreads=0, hist=False):
    """bicgstab_CRS_2; returns (x, status, iters, rr, rr1[, hist])."""
    er = float(np.float32(1e-10))
    er0 = er if er0 is None else er0
    relative_er = er if relative_er is None else relative_er
    n = b.size
    x = np.zeros(n) if x0 is None else np.array(x0, dtype=np.float64)
    it = C.c_int(0)
    rr = C.c_double(0)
    rr1 = C.c_double(0)
    h = np.zeros(max(itrmax, 1)) if hist else None
    st = load().orc_bicgstab_crs2(n, _p(_c(row_ptr, np.int64), C.c_int64), _p(_c(col_idx, np.int32), C.c_int32),
                                  _p(_c(val, np.float64), C.c_double), _p(_c(b, np.float64), C.c_double),
                                  _p(x, C.c_double), itrmax, er0, relative_er, nthreads, C.byref(it), C.byref(rr),
                                  C.byref(rr1), _p(h, C.c_double))
    out = (x, st, it.value, rr.value, rr1.value)
    return out + (h[:it.value],) if hist else out


def bicgstab_crs(row_ptr, col_idx, val, b, er, x0=None, itrmax=100000, nthreads=0, hist=False):
    """bicgstab_CRS (LinearSolver_ form); returns (x, status, iters, rr, init_rr[, hist])."""
    n = b.size
    x = np.zeros(n) if x0 is None else np.array(x0, dtype=np.float64)
    it = C.c_int(0)
    rr = C.c_double(0)
    rr0 = C.c_double(0)
    h = np.zeros(max(itrmax, 1)) if hist else None
    st = load().orc_bicgstab_crs(n, _p(_c(row_ptr, np.int64), C.c_int64), _p(_c(col_idx, np.int32), C.c_int32),
                                 _p(_c(val, np.float64), C.c_double), _p(_c(b, np.float64), C.c_double),
                                 _p(x, C.c_double), itrmax, float(er), nthreads, C.byref(it), C.byref(rr),
                                 C.byref(rr0), _p(h, C.c_double))
    out = (x, st, it.value, rr.value, rr0.value)
    return out + (h[:it.value],) if hist else out


def crs_spmv(row_ptr, col_idx, val, old_vector, new_vector):
    """crs_spmv_real64 / crs_spmv_real32 (by the dtype of val): new_vector + A old_vector."""
    f32 = np.asarray(val).dtype == np.float32
    dt = np.float32 if f32 else np.float64
    ct = C.c_float if f32 else C.c_double
    y = np.array(new_vector, dtype=dt)
    fn = load().orc_crs_spmv_real32 if f32 else load().orc_crs_spmv_real64
    fn(_p(_c(val, dt), ct), _p(_c(row_ptr, np.int64), C.c_int64), _p(_c(col_idx, np.int32), C.c_int32),
       _p(_c(old_vector, dt), ct), _p(y, ct), y.size)
    return y


def assemble_rows(nodcoord, elemnod, E, nu, row_ptr, col_idx, nodes, jmode=0):
    """Stiffness CRS values of the DOF rows of the given 0-based nodes (others zero)."""
    nn = nodcoord.shape[0]
    mark = np.zeros(nn, np.uint8)
    mark[np.asarray(nodes)] = 1
    val = np.empty(col_idx.size)
    rc = load().orc_assemble_stiffness_rows(nn, elemnod.shape[0], _p(_fl(nodcoord, np.float64), C.c_double),
                                            _p(_fl(elemnod, np.int32), C.c_int32),
                                            _p(_c(E, np.float64), C.c_double), _p(_c(nu, np.float64), C.c_double),
                                            jmode, _p(mark, C.c_uint8), _p(_c(row_ptr, np.int64), C.c_int64),
                                            _p(_c(col_idx, np.int32), C.c_int32), _p(val, C.c_double))
    if rc:
        raise ValueError("assembly: missing CRS slot")
    return val


def solve(row_ptr, col_idx, val, rhs, fixed, ubar, x0=None, itrmax=100000, er0=None, relative_er=None, nthreads=0):
    """solveFEMSolver on copies; returns (x, status, iters, rr, rr1)."""
    er = float(np.float32(1e-10))
    er0 = er if er0 is None else er0
    relative_er = er if relative_er is None else relative_er
    val = np.array(val, dtype=np.float64)
    rhs = np.array(rhs, dtype=np.float64)
    n = rhs.size
    x = np.zeros(n) if x0 is None else np.array(x0, dtype=np.float64)
    it = C.c_int(0)
    rr = C.c_double(0)
    rr1 = C.c_double(0)
    st = load().orc_solve(n, _p(_c(row_ptr, np.int64), C.c_int64), _p(_c(col_idx, np.int32), C.c_int32),
                          _p(val, C.c_double), _p(rhs, C.c_double), _p(_c(fixed, np.uint8), C.c_uint8),
                          _p(_c(ubar, np.float64), C.c_double), _p(x, C.c_double), itrmax, er0, relative_er,
                          nthreads, C.byref(it), C.byref(rr), C.byref(rr1))
    return x, st, it.value, rr.value, rr1.value


def hex8_mass_matrix(X, rho):
    """MassMatrixFEMDomain of one hex8 (X: 8x3 node coordinates): 24x24."""
    me = np.zeros(576)
    load().orc_hex8_mass_matrix(_p(_c(X, np.float64).ravel(), C.c_double), float(rho), _p(me, C.c_double))
    return me.reshape(24, 24, order="F")


def mass_matrix_crs(nodcoord, elemnod, rho, row_ptr, col_idx):
    """MassMatrix_as_CRS (rho: one value or one per element) on the given ZeroMatrix pattern."""
    xyz = _fl(nodcoord, np.float64)
    en = _fl(elemnod, np.int32)
    nn, ne = np.asarray(nodcoord).shape[0], np.asarray(elemnod).shape[0]
    r = np.broadcast_to(np.asarray(rho, dtype=np.float64), (ne,)).copy()
    val = np.zeros(int(row_ptr[-1] - 1))
    st = load().orc_mass_matrix_crs(nn, ne, _p(xyz, C.c_double), _p(en, C.c_int32), _p(r, C.c_double),
                                    _p(_c(row_ptr, np.int64), C.c_int64), _p(_c(col_idx, np.int32), C.c_int32),
                                    _p(val, C.c_double))
    if st != 0:
        raise ValueError("mass matrix entry outside the pattern")
    return val


def jacobi_scale(row_ptr, col_idx, val, rhs):
    """JacobiPreconditionerCRS on copies; returns (val, rhs)."""
    val = np.array(val, dtype=np.float64)
    rhs = np.array(rhs, dtype=np.float64)
    load().orc_jacobi_scale(rhs.size, _p(_c(row_ptr, np.int64), C.c_int64), _p(_c(col_idx, np.int32), C.c_int32),
                            _p(val, C.c_double), _p(rhs, C.c_double))
    return val, rhs


def ilu0(row_ptr, col_idx, val):
    """ILUCRS(0) on a copy of val: the factor values on the same pattern."""
    f = np.array(val, dtype=np.float64)
    load().orc_ilu0(row_ptr.size - 1, _p(_c(row_ptr, np.int64), C.c_int64), _p(_c(col_idx, np.int32), C.c_int32),
                    _p(f, C.c_double))
    return f


def ilu_apply(row_ptr, col_idx, fval, v):
    """ILU_matvecCRS of the factor fval applied to v."""
    n = row_ptr.size - 1
    y = np.empty(n)
    load().orc_ilu_apply(n, _p(_c(row_ptr, np.int64), C.c_int64), _p(_c(col_idx, np.int32), C.c_int32),
                         _p(_c(fval, np.float64), C.c_double), _p(_c(v, np.float64), C.c_double), _p(y, C.c_double))
    return y


def bicgstab_ilu(row_ptr, col_idx, val, b, fval=None, x0=None, itrmax=100000, er0=None, relative_er=None,
                 nthreads=0):
    """bicgstab_CRS_ILU; fval = ILU(0) factor (computed here when None).  Returns (x, status, iters, rr, rr1)."""
    er = float(np.float32(1e-10))
    er0 = er if er0 is None else er0
    relative_er = er if relative_er is None else relative_er
    if fval is None:
        fval = ilu0(row_ptr, col_idx, val)
    n = b.size
    x = np.zeros(n) if x0 is None else np.array(x0, dtype=np.float64)
    it = C.c_int(0)
    rr = C.c_double(0)
    rr1 = C.c_double(0)
    st = load().orc_bicgstab_ilu(n, _p(_c(row_ptr, np.int64), C.c_int64), _p(_c(col_idx, np.int32), C.c_int32),
                                 _p(_c(val, np.float64), C.c_double), _p(_c(fval, np.float64), C.c_double),
                                 _p(_c(b, np.float64), C.c_double), _p(x, C.c_double), itrmax, er0, relative_er,
                                 nthreads, C.byref(it), C.byref(rr), C.byref(rr1))
    return x, st, it.value, rr.value, rr1.value


def build_system(w):
    """Oracle CRS system of a Workload: (row_ptr, col_idx, val, rhs, fixed(uint8), ubar)."""
    rp, ci = zero_matrix(w.elemnod, w.nn, 3)
    val = assemble(w.nodcoord, w.elemnod, w.E, w.nu, rp, ci)
    rhs = traction_box(w.nodcoord, w.elemnod, w.traction_box, w.pressure)
    fixed = np.zeros(3 * w.nn, dtype=np.uint8)
    fixed[w.fixed_dofs - 1] = 1
    return rp, ci, val, rhs, fixed, np.zeros(3 * w.nn)


# ---- multi-domain assembly with overset penalties ----------------------------------------
def gauss_point_position(X, gp):
    """GlobalPositionOfGaussPoint of Gauss point gp (1-based) of an element with nodes X (8,3)."""
    out = np.empty(3)
    load().orc_gauss_point_position(_p(_c(X, np.float64).ravel(), C.c_double), int(gp) - 1, _p(out, C.c_double))
    return out


def inside_hex8(X, p):
    return bool(load().orc_inside_hex8(_p(_c(X, np.float64).ravel(), C.c_double),
                                       _p(_c(p, np.float64), C.c_double)))


def nearest_element(nodcoord, elemnod, p):
    """Mesh%nearestElementID: 1-based element id or -1."""
    xyz = _fl(nodcoord, np.float64)
    en = _fl(elemnod, np.int32)
    e = load().orc_nearest_element(nodcoord.shape[0], elemnod.shape[0], _p(xyz, C.c_double), _p(en, C.c_int32),
                                   _p(_c(p, np.float64), C.c_double))
    return int(e) + 1 if e >= 0 else -1


def overset_gpp(nodA, enA, nodB, enB):
    """FEMDomain%overset(GPP) of mesh A onto mesh B: (elem, gp, pair_elem, position), 1-based ids."""
    xa, ea, xb, eb = _fl(nodA, np.float64), _fl(enA, np.int32), _fl(nodB, np.float64), _fl(enB, np.int32)
    args = (nodA.shape[0], enA.shape[0], _p(xa, C.c_double), _p(ea, C.c_int32), nodB.shape[0], enB.shape[0],
            _p(xb, C.c_double), _p(eb, C.c_int32))
    cap = enA.shape[0] * 8
    me, gp, pe = (np.empty(cap, np.int32) for _ in range(3))
    pos = np.empty(cap * 3)
    k = load().orc_overset_gpp(*args, _p(me, C.c_int32), _p(gp, C.c_int32), _p(pe, C.c_int32),
                               _p(pos, C.c_double), cap)
    return me[:k].copy(), gp[:k].copy(), pe[:k].copy(), pos[:3 * k].reshape(k, 3).copy()


def overset_p2p(nodA, enA, nodB, enB):
    """FEMDomain%overset(P2P) of mesh A onto mesh B: (node, pair_elem, position), 1-based ids."""
    xa, xb, eb = _fl(nodA, np.float64), _fl(nodB, np.float64), _fl(enB, np.int32)
    cap = nodA.shape[0]
    nd, pe = np.empty(cap, np.int32), np.empty(cap, np.int32)
    pos = np.empty(cap * 3)
    k = load().orc_overset_p2p(nodA.shape[0], _p(xa, C.c_double), nodB.shape[0], enB.shape[0], _p(xb, C.c_double),
                               _p(eb, C.c_int32), _p(nd, C.c_int32), _p(pe, C.c_int32), _p(pos, C.c_double), cap)
    return nd[:k].copy(), pe[:k].copy(), pos[:3 * k].reshape(k, 3).copy()


def local_coordinate(X, p):
    xi = np.empty(3)
    load().orc_local_coordinate(_p(_c(X, np.float64).ravel(), C.c_double), _p(_c(p, np.float64), C.c_double),
                                _p(xi, C.c_double))
    return xi


def connect_coeffs(nodcoord, elemnod, elem, gp, pair_elem, position):
    """Per connect: c16 (16) = [N_own(gp), -N_pair(xi)] and detJ of the own Gauss point; a P2P
    connect (gp 0, elem = own node) gets [1, -N_pair(xi), 0 ...] and detJ 1."""
    k = len(elem)
    c16 = np.zeros((k, 16))
    dj = np.empty(k)
    if k == 0:
        return c16, dj
    xyz = _fl(nodcoord, np.float64)
    en = _fl(elemnod, np.int32)
    load().orc_connect_coeffs_all(np.asarray(nodcoord).shape[0], np.asarray(elemnod).shape[0], _p(xyz, C.c_double),
                                  _p(en, C.c_int32), k, _p(_c(elem, np.int32), C.c_int32),
                                  _p(_c(gp, np.int32), C.c_int32), _p(_c(pair_elem, np.int32), C.c_int32),
                                  _p(_c(position, np.float64).ravel(), C.c_double), _p(c16, C.c_double),
                                  _p(dj, C.c_double))
    return c16, dj


def connect_nodes(elemnod, elem, pair_elem, gp=None):
    """16 node ids per connect (1-based); P2P connects (gp 0): own node, 8 paired nodes, 0s."""
    en = np.asarray(elemnod, dtype=np.int32)
    elem, pair = np.asarray(elem), np.asarray(pair_elem)
    p2p = np.zeros(elem.size, bool) if gp is None else np.asarray(gp) == 0
    g = ~p2p
    out = np.zeros((elem.size, 16), np.int32)
    out[g, :8] = en[elem[g] - 1]
    out[g, 8:] = en[pair[g] - 1]
    out[p2p, 0] = elem[p2p]
    out[p2p, 1:9] = en[pair[p2p] - 1]
    return out


def multi_crs(nn, elemnod, any_overset, elem, pair_elem, gp=None):
    """Multi-domain setCRS(DOF=3) pattern of the concatenated mesh: (row_ptr, col_idx)."""
    en = _fl(elemnod, np.int32)
    nodes = _c(connect_nodes(elemnod, elem, pair_elem, gp), np.int32).ravel() if len(elem) else np.zeros(1, np.int32)
    ne = elemnod.shape[0]
    rp = np.empty(3 * nn + 1, np.int64)
    nnz = load().orc_multi_crs(nn, ne, _p(en, C.c_int32), int(bool(any_overset)), len(elem),
                               _p(nodes, C.c_int32), _p(rp, C.c_int64), None)
    ci = np.empty(max(nnz, 1), np.int32)
    load().orc_multi_crs(nn, ne, _p(en, C.c_int32), int(bool(any_overset)), len(elem), _p(nodes, C.c_int32),
                         _p(rp, C.c_int64), _p(ci, C.c_int32))
    return rp, ci[:nnz].copy()


def ebom(nodcoord, elemnod, row_ptr, col_idx, val, elem, gp, pair_elem, position, penalty):
    """setEbOM(penalty, DOF=3) added into a copy of val."""
    val = np.array(val, dtype=np.float64)
    k = len(elem)
    if k == 0:
        return val
    c16, dj = connect_coeffs(nodcoord, elemnod, elem, gp, pair_elem, position)
    nodes = _c(connect_nodes(elemnod, elem, pair_elem, gp), np.int32).ravel()
    st = load().orc_ebom(row_ptr.size - 1, _p(_c(row_ptr, np.int64), C.c_int64),
                         _p(_c(col_idx, np.int32), C.c_int32), _p(val, C.c_double), k, _p(nodes, C.c_int32),
                         _p(_c(c16, np.float64).ravel(), C.c_double), _p(dj, C.c_double), float(penalty))
    if st != 0:
        raise ValueError("connect entry outside the pattern")
    return val


def band_cholesky(ab):
    """orc_pbtrf_lower (LAPACK dpbtf2, lower) of a band matrix given as an (n, kd+1) array whose
    row j is column j of the band (ab[j, d] = A[j+d][j]).  Returns (info, L in the same layout);
    info = j+1 when the pivot of column j is not positive."""
    a = np.array(ab, dtype=np.float64, order="C", copy=True)
    n, ldab = a.shape
    info = load().orc_pbtrf_lower(n, ldab - 1, _p(a, C.c_double), ldab)
    return int(info), a


def band_solve(lfac, b):
    """orc_pbtrs_lower: (L L^T)^-1 b with L from band_cholesky."""
    a = np.ascontiguousarray(lfac, dtype=np.float64)
    x = np.array(b, dtype=np.float64, copy=True)
    n, ldab = a.shape
    load().orc_pbtrs_lower(n, ldab - 1, _p(a, C.c_double), ldab, _p(x, C.c_double))
    return x


def band_from_crs(row_ptr, col_idx, val, perm, kd):
    """Lower band (n, kd+1) of P A P^T, perm[q] = original row of band row q (0-based)."""
    rp, ci, va = _c(row_ptr, np.int64), _c(col_idx, np.int32), _c(val, np.float64)
    pm = _c(perm, np.int32)
    n = rp.size - 1
    ab = np.zeros((n, kd + 1))
    bad = load().orc_band_from_crs(n, _p(rp, C.c_int64), _p(ci, C.c_int32), _p(va, C.c_double),
                                   _p(pm, C.c_int32), kd, _p(ab, C.c_double), kd + 1)
    if bad:
        raise ValueError("band_from_crs: an entry lies outside the band")
    return ab


def direct_solve(row_ptr, col_idx, val, rhs, perm, kd):
    """Direct solve of A x = rhs through the band Cholesky of P A P^T (lower triangle used)."""
    ab = band_from_crs(row_ptr, col_idx, val, perm, kd)
    info, lf = band_cholesky(ab)
    if info:
        raise ValueError(f"direct_solve: pivot of band column {info - 1} is not positive")
    pm = np.asarray(perm)
    y = band_solve(lf, np.asarray(rhs, dtype=np.float64)[pm])
    x = np.empty_like(y)
    x[pm] = y
    return x
