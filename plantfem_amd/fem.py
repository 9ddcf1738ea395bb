"""Python mirror of the reference's FEMDomain_/FEMSolver_ surface for the static
linear-elastic path, dispatching every compute step to libpfgpu.so.

Names, argument meanings and error behaviour follow plantFEM:
  FEMDomain_%create("Cube3D", x_num, y_num, z_num, x_len, y_len, z_len)  FEMDomainClass.f90:7215-7310
  FEMDomain_%ZeroMatrix(DOF)                                             FEMDomainClass.f90:10581-10840
  FEMDomain_%StiffnessMatrix(ElementID, E, v)                            FEMDomainClass.f90:11457-11515
  FEMDomain_%StiffnessMatrix(YoungModulus(:), PoissonRatio(:))           FEMDomainClass.f90:11123-11232
  FEMDomain_%select(x_min, x_max, ...)                                   FEMDomainClass.f90:12854-12926
  FEMSolver_%init/setDomain/setCRS/setRHS/setMatrix/setVector/fix/solve  FEMSolverClass.f90:84-163
Indices are 1-based as in the reference (node ids, element ids, DOF ids).
"""
from __future__ import annotations

import ctypes as C
from dataclasses import dataclass

import math

import numpy as np

from . import _lib as L


# ----------------------------------------------------------------------------- meshes
def cube3d(x_num: int = 10, y_num: int = 10, z_num: int = 10, x_len: float = 1.0, y_len: float = 1.0,
           z_len: float = 1.0):
    """Cube3D numbering (MeshClass.f90:6348-6361, :6745-6808, :6992-7090).

    Returns (NodCoord(nn,3) float64 Fortran-ordered, ElemNod(ne,8) int32 Fortran-ordered, 1-based).
    Node id = i + (xn+1)(j-1) + (xn+1)(yn+1)(k-1), x fastest; coordinates lx/dble(xn)*dble(i-1).
    """
    xn, yn, zn = int(x_num), int(y_num), int(z_num)
    i = np.arange(xn + 1, dtype=np.float64)
    j = np.arange(yn + 1, dtype=np.float64)
    k = np.arange(zn + 1, dtype=np.float64)
    xs = (float(x_len) / float(xn)) * i
    ys = (float(y_len) / float(yn)) * j
    zs = (float(z_len) / float(zn)) * k
    Z, Y, X = np.meshgrid(zs, ys, xs, indexing="ij")
    nod = np.empty((X.size, 3), dtype=np.float64, order="F")
    nod[:, 0] = X.ravel()
    nod[:, 1] = Y.ravel()
    nod[:, 2] = Z.ravel()
    nn2 = (xn + 1) * (yn + 1)
    ii, jj = np.meshgrid(np.arange(1, xn + 1), np.arange(1, yn + 1), indexing="xy")  # x fastest
    ii = ii.ravel()
    jj = jj.ravel()
    q = np.stack([ii + (jj - 1) * (xn + 1), ii + 1 + (jj - 1) * (xn + 1),
                  xn + 2 + ii + (jj - 1) * (xn + 1), xn + 1 + ii + (jj - 1) * (xn + 1)], axis=1)
    layers = np.arange(zn)[:, None, None] * nn2
    bottom = (q[None, :, :] + layers).reshape(-1, 4)
    top = bottom + nn2
    elem = np.asfortranarray(np.concatenate([bottom, top], axis=1).astype(np.int32))
    return nod, elem


PF_GLYCINE_MAX = 1  # MeshClass.f90:6


def leaf3d(x_num: int = 10, y_num: int = 10, z_num: int = 10, x_len: float = 1.0, y_len: float = 1.0,
           z_len: float = 1.0, shaperatio: float = 0.3, species=None, SoyWidthRatio: float = 0.5):
    """Mesh%create(meshtype="Leaf3D") (MeshClass.f90:6017-6276): the rectangular3D mesh (Cube3D
    numbering; x = width, y = thickness, z = length) centred in x and y, then shaped along z.

    Default (MeshClass.f90:6246-6272): half-width ratio 1/10 up to z = zl/20 (where the
    thickness grows to width/10), rising linearly to 1 at z = shaperatio*zl, falling linearly to
    1/10 at the tip.  species=PF_GLYCINE_MAX (:6033-6094): the soybean outline (circular arc to
    the widest point alpha = SoyWidthRatio*width on one side, (1-SoyWidthRatio)*width on the
    other, an arc of radius 1.2 r to the tip, each blended half-and-half with a straight line),
    a fat base (x/10, thickness x5 below z = length/25) and the curl y -= 0.1 theta (theta - 1).
    The soybean branch decides each node's side from the x extent of the nodes as already
    moved, so it runs node by node in id order."""
    nod, el = cube3d(x_num, y_num, z_num, x_len, y_len, z_len)
    x, y, z = nod[:, 0].copy(), nod[:, 1].copy(), nod[:, 2].copy()
    x = x - (x.max() - x.min()) * 0.50
    y = y - (y.max() - y.min()) * 0.50
    if species is None:
        thickness_ = y.max() - y.min()
        width = x.max() - x.min()
        thickness_ratio = (width * 1.0 / 10.0) / thickness_
        zl = z.max() - z.min()
        s = float(shaperatio)
        z20 = 1.0 / 20.0 * zl
        base = z <= z20
        rise = (~base) & (z <= zl * s)
        ratio = np.where(base, 1.0 / 10.0,
                         np.where(rise, 1.0 / 10.0 + 0.90 / (zl * s - z20) * (z - z20),
                                  1.0 - 0.90 / (zl - s * zl) * (z - s * zl)))
        y = np.where(base, y * thickness_ratio, y)
        x = x * ratio
    elif species == PF_GLYCINE_MAX:
        lin_curve_ratio = 0.50
        length = z.max() - z.min()
        width = x.max() - x.min()
        sw = float(SoyWidthRatio)
        if sw >= 1.0 or sw <= 0.0:
            raise ValueError(f"mesh%create(leaf3d, PF_SOYBEAN) >> invalid SoyWidthRatio {SoyWidthRatio}")
        thickness_ratio = 5.0
        xmax, xmin = x.max(), x.min()
        for i in range(x.size):
            xx = z[i]
            alpha = sw * width if x[i] <= (xmax + xmin) * 0.50 else (1.0 - sw) * width
            r = (alpha * alpha + (length - alpha) * (length - alpha)) / (2 * alpha) * 1.20
            if xx <= 1.0 / 25.0 * length:
                x[i] = x[i] * 1.0 / 10.0
                y[i] = y[i] * thickness_ratio
            else:
                if xx < alpha:
                    yy = np.sqrt(alpha * alpha - (xx - alpha) * (xx - alpha))
                    yy_ = xx
                else:
                    yy_ = alpha + (-alpha) / (length - alpha) * (xx - alpha)
                    yy = alpha - r + np.sqrt(r * r - (xx - alpha) * (xx - alpha))
                yy = abs(lin_curve_ratio * yy + (1.0 - lin_curve_ratio) * yy_)
                x[i] = x[i] * (yy / alpha)
            # the reference re-reads maxval/minval of the moved coordinates for the next node
            xmax, xmin = x.max(), x.min()
        theta = np.abs(x / (width / 2.0))
        y = y - 0.10 * theta * (theta - 1.0)
    else:
        raise ValueError(f"Mesh%create => No such species as {species}")
    out = np.empty_like(nod, order="F")
    out[:, 0], out[:, 1], out[:, 2] = x, y, z
    return out, el


def _rect2d(xn: int, yn: int, lx: float, ly: float):
    """rectangular2D (MeshClass.f90:6745-6808): x fastest, lx/dble(xn)*dble(i-1); quads."""
    i = np.arange(xn + 1, dtype=np.float64)
    j = np.arange(yn + 1, dtype=np.float64)
    Y, X = np.meshgrid(ly / float(yn) * j, lx / float(xn) * i, indexing="ij")
    nod = np.stack([X.ravel(), Y.ravel()], axis=1)
    ii, jj = np.meshgrid(np.arange(1, xn + 1), np.arange(1, yn + 1), indexing="xy")
    ii, jj = ii.ravel(), jj.ravel()
    el = np.stack([ii + (jj - 1) * (xn + 1), ii + 1 + (jj - 1) * (xn + 1), xn + 2 + ii + (jj - 1) * (xn + 1),
                   xn + 1 + ii + (jj - 1) * (xn + 1)], axis=1).astype(np.int32)
    return nod, el


def circle2d(x_num: int = 10, y_num: int = 10, x_len: float | None = None, y_len: float | None = None):
    """Circle2D (MeshClass.f90:6524-6743): a (2m)^2 quad square on [-1,1]^2 rounded towards a
    disc (node scale alpha(1/R) with R = sqrt(1 + (minor/major)^2), times sqrt(2)*0.95, then
    (2m-1)/(2m)/sqrt(2)), surrounded by 2m rings of 8m nodes at radii 1 + i/(2m); the square's
    boundary is tied to ring 1 and rings 1..m+2 by 8m quads per layer (the outer rings stay
    unused), m = max(x_num, y_num)//2 + 1; resized to x_len, y_len over all nodes.  2-D nodes
    (nn, 2) and quads (ne, 4), 1-based."""
    xn = x_num // 2 + 1
    yn = y_num // 2 + 1
    xn = yn = max(xn, yn)
    if xn < 2:  # the ring layers would index past the nodes that exist
        raise ValueError(f"Circle2D / Cylinder: x_num or y_num must be >= 2 (got {x_num}, {y_num})")
    nod, el = _rect2d(2 * xn, 2 * yn, 2.0, 2.0)
    nod = nod - 1.0
    x, y = nod[:, 0].copy(), nod[:, 1].copy()
    s2 = math.sqrt(2.0)
    theta = (1.0 - 0.80) / (1.0 - 1.0 / s2)
    for k in range(x.size):
        xx, yy = x[k], y[k]
        RR = 1.00
        if xx != 0.0:
            if abs(yy / xx) > 1.0:
                if yy != 0.0:
                    RR = math.sqrt((xx / yy) * (xx / yy) + 1)
            else:
                RR = math.sqrt((yy / xx) * (yy / xx) + 1)
        alpha = theta * ((1 / RR) - 1.0 / s2) + 0.790
        x[k] = (xx * alpha) * s2 * 0.95
        y[k] = (yy * alpha) * s2 * 0.95
    c = float(2 * xn - 1) / float(2 * xn)
    x = c * x / s2
    y = c * y / s2
    m = 2 * xn
    sq = (m + 1) * (m + 1)
    ring = np.empty((m * 4 * m, 2))
    for i in range(1, m + 1):
        r = 1.0 + float(i) * (1.0 / float(m))
        for j in range(1, 4 * m + 1):
            a = 2.0 * math.pi / 4.0 / float(m) * float(j - 1)
            ring[(i - 1) * 4 * m + j - 1] = (r * math.cos(a), r * math.sin(a))
    e1 = []
    j = 0
    for i in range(1, xn + 1):
        j += 1
        e1.append([(m + 1) * (xn + i), sq + j, sq + j + 1, (m + 1) * (xn + i + 1)])
    for i in range(1, m + 1):
        j += 1
        e1.append([sq - i + 1, sq + j, sq + j + 1, sq - i])
    for i in range(1, m + 1):
        j += 1
        e1.append([sq - (m + 1) + 1 - (i - 1) * (m + 1), sq + j, sq + j + 1, sq - (m + 1) + 1 - i * (m + 1)])
    for i in range(1, m + 1):
        j += 1
        e1.append([i, sq + j, sq + j + 1, i + 1])
    for i in range(1, xn + 1):
        j += 1
        e1.append([(m + 1) * i, sq + j, sq + j + 1, (m + 1) * (i + 1)])
    e1[j - 1][2] = sq + 1
    for i in range(1, xn + 2):
        ini = j + 1
        for k in range(1, 8 * xn):
            j += 1
            e1.append([sq + j - 8 * xn, sq + j, sq + j + 1, sq + j + 1 - 8 * xn])
        j += 1
        e1.append([sq + j - 8 * xn, sq + j, sq + ini, sq + ini - 8 * xn])
    nod = np.concatenate([np.stack([x, y], axis=1), ring])
    el = np.concatenate([el, np.array(e1, dtype=np.int32)]).astype(np.int32)
    if x_len is not None:
        nod[:, 0] = x_len / (nod[:, 0].max() - nod[:, 0].min()) * nod[:, 0]
    if y_len is not None:
        nod[:, 1] = y_len / (nod[:, 1].max() - nod[:, 1].min()) * nod[:, 1]
    return nod, el


def clean_mesh(nod, el):
    """cleanMesh (MeshClass.f90:8571-8605): drop nodes no element uses, renumber the rest."""
    used = np.zeros(nod.shape[0], bool)
    used[el.ravel() - 1] = True
    removed_before = np.cumsum(~used)  # removed nodes with id <= k
    return nod[used].copy(), (el - removed_before[el - 1]).astype(np.int32)


def convert_2d_to_3d(nod, el, thickness: float = 1.0, division: int = 1):
    """Convert2Dto3D (MeshClass.f90:6992-7090): layer j of the 2-D nodes at z = Tn/dble(N)*dble(j-1),
    element e of layer j = (2-D nodes + n(j-1), 2-D nodes + n j)."""
    n = nod.shape[0]
    z = float(thickness) / float(division) * np.arange(division + 1, dtype=np.float64)
    out = np.empty((n * (division + 1), 3), order="F")
    out[:, 0] = np.tile(nod[:, 0], division + 1)
    out[:, 1] = np.tile(nod[:, 1], division + 1)
    out[:, 2] = np.repeat(z, n)
    layers = np.arange(division)[:, None, None] * n
    e3 = np.concatenate([(el[None] + layers).reshape(-1, el.shape[1]), (el[None] + layers + n).reshape(-1, el.shape[1])],
                        axis=1)
    return out, np.asfortranarray(e3.astype(np.int32))


def cylinder3d(x_num: int = 10, y_num: int = 10, z_num: int = 10, x_len: float = 1.0, y_len: float = 1.0,
               z_len: float = 1.0):
    """Mesh%create(meshtype="Cylinder") (MeshClass.f90:6506-6522), the stem mesh of StemClass:
    Circle2D on a unit disc, clean, Convert2Dto3D(thickness=z_len, division=z_num), resized to
    x_len, y_len over the remaining nodes."""
    nod, el = circle2d(x_num, y_num, 1.0, 1.0)
    nod, el = clean_mesh(nod, el)
    nod, el = convert_2d_to_3d(nod, el, z_len, z_num)
    nod[:, 0] = x_len / (nod[:, 0].max() - nod[:, 0].min()) * nod[:, 0]
    nod[:, 1] = y_len / (nod[:, 1].max() - nod[:, 1].min()) * nod[:, 1]
    return np.asfortranarray(nod), el


@dataclass
class CRS:
    """CRS_ (SparseClass.f90:128-131): 1-based int64 row_ptr, int32 col_idx, float64 val."""
    row_ptr: np.ndarray
    col_idx: np.ndarray
    val: np.ndarray

    @property
    def n(self) -> int:
        return int(self.row_ptr.size - 1)

    @property
    def nnz(self) -> int:
        return int(self.col_idx.size)

    def to_scipy(self):
        import scipy.sparse as sp
        return sp.csr_matrix((self.val, self.col_idx.astype(np.int64) - 1, self.row_ptr - 1),
                             shape=(self.n, self.n))


def _f64(a):
    return np.ascontiguousarray(a, dtype=np.float64)


def _i32(a):
    return np.ascontiguousarray(a, dtype=np.int32)


def _i64(a):
    return np.ascontiguousarray(a, dtype=np.int64)


FEMDomain_Overset_GPP = 1  # FEMDomainClass.f90:40-41
FEMDomain_Overset_P2P = 2


class OversetConnect:
    """The OversetConnect_ list of one domain as arrays: ElementID (own element, GPP; 0 for P2P),
    NodeID (own node, P2P; 0 for GPP), GaussPointID (1-based; 0 marks P2P as in the reference),
    the paired element (1-based, in the paired domain), position (k, 3) and DomainIDs12 as
    (own, paired) per connect; all connects are active."""

    def __init__(self):
        self.NodeID = np.zeros(0, np.int32)
        self.ElementID = np.zeros(0, np.int32)
        self.GaussPointID = np.zeros(0, np.int32)
        self.PairElementID = np.zeros(0, np.int32)
        self.position = np.zeros((0, 3))
        self.DomainIDs12 = np.zeros((0, 2), np.int32)

    @property
    def n(self) -> int:
        return int(self.ElementID.size)

    def append(self, elem, gp, pair, pos, my_id: int, pair_id: int, node=None):
        k = len(pair)
        elem = np.zeros(k, np.int32) if elem is None else elem
        node = np.zeros(k, np.int32) if node is None else node
        self.NodeID = np.concatenate([self.NodeID, np.asarray(node, np.int32)])
        self.ElementID = np.concatenate([self.ElementID, np.asarray(elem, np.int32)])
        self.GaussPointID = np.concatenate([self.GaussPointID, np.asarray(gp, np.int32)])
        self.PairElementID = np.concatenate([self.PairElementID, np.asarray(pair, np.int32)])
        self.position = np.concatenate([self.position, np.asarray(pos, np.float64).reshape(k, 3)])
        ids = np.tile(np.array([my_id, pair_id], np.int32), (k, 1))
        self.DomainIDs12 = np.concatenate([self.DomainIDs12, ids])


class FEMDomain:
    """Subset of FEMDomain_ used by the static solve path."""

    def __init__(self, nodcoord=None, elemnod=None, ctx: L.Context | None = None):
        self.ctx = ctx
        self.nodcoord = None if nodcoord is None else np.asfortranarray(nodcoord, dtype=np.float64)
        self.elemnod = None if elemnod is None else np.asfortranarray(elemnod, dtype=np.int32)
        # OversetConnect (FEMDomainClass.f90:14606-14700): None until the first overset call
        self.OversetConnect: OversetConnect | None = None

    def overset(self, FEMDomain: "FEMDomain", DomainID: int, algorithm=None, MyDomainID: int | None = None):
        """overset(FEMDomain, DomainID, algorithm=FEMDomain_Overset_GPP, MyDomainID)
        (FEMDomainClass.f90:14606-14700): every Gauss point of this mesh that lies in an element
        of FEMDomain becomes one connect (own element, Gauss point, paired element, position,
        domain ids [MyDomainID (default 1), DomainID]), appended in the reference's order.  The
        search runs on the GPU (pf_overset_gpp)."""
        alg = FEMDomain_Overset_GPP if algorithm is None else algorithm
        p2p = alg in ("P2P", FEMDomain_Overset_P2P)
        if not p2p and alg not in ("GPP", FEMDomain_Overset_GPP):
            raise ValueError(f"overset: unknown algorithm {algorithm!r}")
        my_id = 1 if MyDomainID is None else int(MyDomainID)
        if self.OversetConnect is None:
            self.OversetConnect = OversetConnect()
            self._overset_p2p = p2p
        elif self._overset_p2p != p2p:  # FEMDomainClass.f90:14656-14662, :14717-14723
            raise ValueError("overset: P2P & GPP cannot be mixed on one domain")
        xa, ea = self._flat()
        xb, eb = FEMDomain._flat()
        lib = L.load()
        cap = self.nn() if p2p else 8 * self.ne()
        elem, gp, pe = (np.empty(max(cap, 1), np.int32) for _ in range(3))
        pos = np.empty(max(cap, 1) * 3)
        count = C.c_int64(0)
        meshes = (self._ctx().h, self.nn(), self.ne(), L.ptr(xa, C.c_double), L.ptr(ea, C.c_int32),
                  FEMDomain.nn(), FEMDomain.ne(), L.ptr(xb, C.c_double), L.ptr(eb, C.c_int32), cap, C.byref(count))
        if p2p:
            L.check(lib.pf_overset_p2p(*meshes, L.ptr(elem, C.c_int32), L.ptr(pe, C.c_int32),
                                       L.ptr(pos, C.c_double)), "overset")
            k = count.value
            self.OversetConnect.append(None, np.zeros(k, np.int32), pe[:k], pos[:3 * k].reshape(k, 3), my_id,
                                       int(DomainID), node=elem[:k])
        else:
            L.check(lib.pf_overset_gpp(*meshes, L.ptr(elem, C.c_int32), L.ptr(gp, C.c_int32), L.ptr(pe, C.c_int32),
                                       L.ptr(pos, C.c_double)), "overset")
            k = count.value
            self.OversetConnect.append(elem[:k], gp[:k], pe[:k], pos[:3 * k].reshape(k, 3), my_id, int(DomainID))
        return self

    def numOversetElements(self) -> int:
        return 0 if self.OversetConnect is None else self.OversetConnect.n

    def _ctx(self):
        return self.ctx or L.default_context()

    def create(self, meshtype: str = "Cube3D", x_num=None, y_num=None, z_num=None, x_len=None, y_len=None,
               z_len=None, shaperatio: float = 0.3, species=None, SoyWidthRatio: float = 0.5):
        """create(meshtype, x_num, y_num, z_num, x_len, y_len, z_len) (FEMDomainClass.f90:7215-7310):
        "Cube3D" / "Cube" (the same numbering), "Cylinder" (StemClass's round stem) or "Leaf3D"
        (organ mesh of LeafClass, with shaperatio or species=PF_GLYCINE_MAX and SoyWidthRatio)."""
        args = (x_num or 10, y_num or 10, z_num or 10, x_len or 1.0, y_len or 1.0, z_len or 1.0)
        if meshtype in ("Cube3D", "Cube", "cube3D", "rectangular3D"):
            self.nodcoord, self.elemnod = cube3d(*args)
        elif meshtype == "Cylinder":
            self.nodcoord, self.elemnod = cylinder3d(*args)
        elif meshtype == "Cylinder3D":
            # FEMDomain_%create sends "Cylinder3D" to to_cylinder (FEMDomainClass.f90:7279-7282,
            # :16136-16157), a different mesh from Mesh%create("Cylinder"); not built here
            raise ValueError("meshtype 'Cylinder3D' (to_cylinder) is not on this path; 'Cylinder' is StemClass's "
                             "Mesh%create route")
        elif meshtype == "Leaf3D":
            self.nodcoord, self.elemnod = leaf3d(*args, shaperatio=shaperatio, species=species,
                                                 SoyWidthRatio=SoyWidthRatio)
        else:
            raise ValueError(f"meshtype {meshtype!r}: Cube3D, Cube, Cylinder and Leaf3D are on this path")
        return self

    def move(self, x: float = 0.0, y: float = 0.0, z: float = 0.0):
        """move(x, y, z): translate every node."""
        self.nodcoord = np.asfortranarray(self.nodcoord + np.array([x, y, z], dtype=np.float64))
        return self

    # sizes (FEMDomainClass.f90:10117-10173)
    def nn(self) -> int:
        return int(self.nodcoord.shape[0])

    def ne(self) -> int:
        return int(self.elemnod.shape[0])

    def nne(self) -> int:
        return int(self.elemnod.shape[1])

    def nd(self) -> int:
        return int(self.nodcoord.shape[1])

    def x_min(self): return float(self.nodcoord[:, 0].min())
    def x_max(self): return float(self.nodcoord[:, 0].max())
    def y_min(self): return float(self.nodcoord[:, 1].min())
    def y_max(self): return float(self.nodcoord[:, 1].max())
    def z_min(self): return float(self.nodcoord[:, 2].min())
    def z_max(self): return float(self.nodcoord[:, 2].max())

    def select(self, x_min=None, x_max=None, y_min=None, y_max=None, z_min=None, z_max=None) -> np.ndarray:
        """Node ids inside the closed box (InOrOut, ArrayClass.f90:2458-2487); defaults +-1e14."""
        lo = [-1.0e14 if v is None else v for v in (x_min, y_min, z_min)]
        hi = [1.0e14 if v is None else v for v in (x_max, y_max, z_max)]
        X = self.nodcoord
        inside = np.ones(X.shape[0], dtype=bool)
        for d in range(3):
            inside &= (lo[d] <= X[:, d]) & (X[:, d] <= hi[d])
        return (np.nonzero(inside)[0] + 1).astype(np.int32)

    def ZeroMatrix(self, DOF: int = 3) -> CRS:
        lib = L.load()
        en = np.asfortranarray(self.elemnod, dtype=np.int32).ravel(order="F")
        nnz = C.c_int64(0)
        L.check(lib.pf_crs_graph(self._ctx().h, self.nn(), self.ne(), self.nne(), L.ptr(en, C.c_int32), DOF,
                                 C.byref(nnz), None, None), "ZeroMatrix")
        rp = np.empty(DOF * self.nn() + 1, dtype=np.int64)
        ci = np.empty(nnz.value, dtype=np.int32)
        L.check(lib.pf_crs_graph(self._ctx().h, self.nn(), self.ne(), self.nne(), L.ptr(en, C.c_int32), DOF,
                                 C.byref(nnz), L.ptr(rp, C.c_int64), L.ptr(ci, C.c_int32)), "ZeroMatrix")
        return CRS(rp, ci, np.zeros(nnz.value, dtype=np.float64))

    def _flat(self):
        return (np.asfortranarray(self.nodcoord).ravel(order="F"), np.asfortranarray(self.elemnod).ravel(order="F"))

    def StiffnessMatrix(self, ElementID=None, E=None, v=None, YoungModulus=None, PoissonRatio=None, jmode: int = 0):
        """Element form (ElementID, E, v) -> Ke(24,24); array form (YoungModulus(:), PoissonRatio(:)) -> CRS."""
        lib = L.load()
        xyz, en = self._flat()
        if YoungModulus is not None:
            Ea = _f64(YoungModulus)
            va = _f64(PoissonRatio)
            crs = self.ZeroMatrix(3)
            L.check(lib.pf_hex8_stiffness_crs(self._ctx().h, self.nn(), self.ne(), L.ptr(xyz, C.c_double),
                                              L.ptr(en, C.c_int32), L.ptr(Ea, C.c_double), L.ptr(va, C.c_double),
                                              jmode, L.ptr(crs.row_ptr, C.c_int64), L.ptr(crs.col_idx, C.c_int32),
                                              L.ptr(crs.val, C.c_double)), "StiffnessMatrix")
            return crs
        ids = np.atleast_1d(np.asarray(ElementID, dtype=np.int64))
        Ea = np.zeros(self.ne())
        va = np.zeros(self.ne())
        Ea[ids - 1] = E
        va[ids - 1] = v
        out = np.empty((ids.size, 24, 24), dtype=np.float64)
        # contiguous runs of element ids share one launch
        for q, e in enumerate(ids):
            ke = np.empty(576, dtype=np.float64)
            L.check(lib.pf_hex8_stiffness(self._ctx().h, self.nn(), self.ne(), L.ptr(xyz, C.c_double),
                                          L.ptr(en, C.c_int32), L.ptr(Ea, C.c_double), L.ptr(va, C.c_double), jmode,
                                          int(e), 1, L.ptr(ke, C.c_double)), "StiffnessMatrix")
            out[q] = ke.reshape(24, 24).T  # column-major -> Ke[r, c]
        return out[0] if np.ndim(ElementID) == 0 else out

    def StiffnessMatrices(self, E, v, first: int = 1, count=None, jmode: int = 0) -> np.ndarray:
        """Batched element form: Ke of elements first..first+count-1 as (count,24,24) with Ke[q, r, c]."""
        lib = L.load()
        xyz, en = self._flat()
        count = self.ne() - first + 1 if count is None else int(count)
        ke = np.empty(count * 576, dtype=np.float64)
        L.check(lib.pf_hex8_stiffness(self._ctx().h, self.nn(), self.ne(), L.ptr(xyz, C.c_double), L.ptr(en, C.c_int32),
                                      L.ptr(_f64(E), C.c_double), L.ptr(_f64(v), C.c_double), jmode, int(first), count,
                                      L.ptr(ke, C.c_double)), "StiffnessMatrices")
        return ke.reshape(count, 24, 24).transpose(0, 2, 1)


    def resize(self, x=None, y=None, z=None):
        """resizeFEMDomain(x=, y=, z=) -> resizeMeshobj x_len/y_len/z_len (MeshClass.f90:892-933):
        rate = new length / (max - min), coordinates = rate * coordinates."""
        X = np.array(self.nodcoord, dtype=np.float64, order="F")
        for d, L_new in enumerate((x, y, z)):
            if L_new is None:
                continue
            old = X[:, d].max() - X[:, d].min()
            X[:, d] = (float(L_new) / old) * X[:, d]
        self.nodcoord = X
        return self

    def MassVector(self, ElementID, Density: float = 1.0, DOF: int = 3, Accel=(0.0, 0.0, 0.0)) -> np.ndarray:
        """MassVectorFEMDomain (FEMDomainClass.f90:10993-11073): rho sum_gp N accel detJ, 24 values
        per element (one element id) or (count, 24) for an array of ids."""
        if DOF != 3:
            raise ValueError("DOF=3 only on this path")
        lib = L.load()
        xyz, en = self._flat()
        ids = np.atleast_1d(np.asarray(ElementID, dtype=np.int64))
        acc = _f64(Accel)
        out = np.empty((ids.size, 24))
        for q, e in enumerate(ids):
            L.check(lib.pf_hex8_mass_vector(self._ctx().h, self.nn(), self.ne(), L.ptr(xyz, C.c_double),
                                            L.ptr(en, C.c_int32), float(Density), L.ptr(acc, C.c_double), int(e), 1,
                                            L.ptr(out[q], C.c_double)), "MassVector")
        return out[0] if np.ndim(ElementID) == 0 else out

    def MassMatrix(self, ElementID, Density=1.0, DOF: int = 3) -> np.ndarray:
        """MassMatrixFEMDomain (FEMDomainClass.f90:10874-10945): rho sum_gp N N^T det J, a 24x24
        matrix (one element id) or (count, 24, 24) for an array of ids; Density one value or one
        per element of the mesh."""
        if DOF != 3:
            raise ValueError("DOF=3 only on this path")
        xyz, en = self._flat()
        ids = np.atleast_1d(np.asarray(ElementID, dtype=np.int64))
        rho = _f64(np.atleast_1d(Density))
        out = np.empty((ids.size, 576))
        for q, e in enumerate(ids):
            L.check(L.load().pf_hex8_mass_matrix(self._ctx().h, self.nn(), self.ne(), L.ptr(xyz, C.c_double),
                                                 L.ptr(en, C.c_int32), L.ptr(rho, C.c_double), rho.size, int(e), 1,
                                                 L.ptr(out[q], C.c_double)), "MassMatrix")
        M = out.reshape(ids.size, 24, 24).transpose(0, 2, 1)  # column-major storage
        return M[0] if np.ndim(ElementID) == 0 else M

    def StressMatrices(self, displacement, E, v, jmode: int = 0) -> np.ndarray:
        """Cell-averaged Cauchy stress of every element, (ne, 3, 3) (StressMatrixFEMDomain without
        GaussPoint, FEMDomainClass.f90:11912-12032)."""
        lib = L.load()
        xyz, en = self._flat()
        Ea, va = _f64(np.atleast_1d(E)), _f64(np.atleast_1d(v))
        sig = np.empty(self.ne() * 9)
        L.check(lib.pf_element_stress(self._ctx().h, self.nn(), self.ne(), L.ptr(xyz, C.c_double), L.ptr(en, C.c_int32),
                                      L.ptr(_f64(displacement), C.c_double), L.ptr(Ea, C.c_double), Ea.size,
                                      L.ptr(va, C.c_double), va.size, jmode, L.ptr(sig, C.c_double)),
                "StressMatrices")
        return sig.reshape(self.ne(), 3, 3).transpose(0, 2, 1)

    def getElementCauchyStress(self, displacement, E, v, i=None, j=None, option=None) -> np.ndarray:
        """getElementCauchyStressFEMDomain (FEMDomainClass.f90:15249-15380): per element sigma(i,j) or
        an invariant ("p", "I1", "I2", "I3", "J1", "J2", "J3", "1,1", ...) of the cell stress."""
        S = self.StressMatrices(displacement, E, v)
        return stress_option(S, i, j, option)


def loadMatrix(name: str) -> "CRS":
    """The three files saveMatrix writes, back into a CRS (any whitespace-separated numbers,
    Fortran D exponents accepted)."""
    def nums(path, dtype):
        with open(path) as f:
            txt = f.read().replace("D", "E").replace("d", "e")
        return np.array(txt.split(), dtype=np.float64).astype(dtype)
    return CRS(nums(name + "_indptr.txt", np.int64), nums(name + "_indices.txt", np.int32),
               nums(name + "_data.txt", np.float64))


def _trace(S):
    return (S[:, 0, 0] + S[:, 1, 1]) + S[:, 2, 2]


def _det3(S):
    # det_mat cofactor expansion along column 1 (MathClass.f90:1460-1476)
    m1 = S[:, 1, 1] * S[:, 2, 2] - S[:, 2, 1] * S[:, 1, 2]
    m2 = S[:, 0, 1] * S[:, 2, 2] - S[:, 2, 1] * S[:, 0, 2]
    m3 = S[:, 0, 1] * S[:, 1, 2] - S[:, 1, 1] * S[:, 0, 2]
    return ((0.0 + S[:, 0, 0] * m1) + (-S[:, 1, 0]) * m2) + S[:, 2, 0] * m3


def stress_option(S, i=None, j=None, option=None) -> np.ndarray:
    """The scalar getElementCauchyStress returns for each cell stress tensor S (ne, 3, 3)."""
    if i is not None and j is not None:
        return S[:, i - 1, j - 1].copy()
    o = option
    if o in ("p", "P"):
        return _trace(S) / 3.0
    if o in ("I", "I1", "i1", "trace", "tr", "TRACE", "TR"):
        return _trace(S)
    if o in ("II", "I2", "i2", "ii"):
        return (_trace(S) * _trace(S) - _trace(S @ S)) * 0.5
    if o in ("III", "I3", "i3", "iii"):
        return _det3(S)
    if o in ("J", "J1", "j1", "j", "JJ", "J2", "j2", "jj", "JJJ", "J3", "j3", "jjj"):
        dev = S - (_trace(S) / 3.0)[:, None, None] * np.eye(3)
        if o in ("J", "J1", "j1", "j"):
            return _trace(dev)
        if o in ("JJ", "J2", "j2", "jj"):
            return (_trace(dev) * _trace(dev) - _trace(dev @ dev)) * 0.5
        return _det3(dev)
    comp = {"1,1": (0, 0), "(1,1)": (0, 0), "_{1,1}": (0, 0), "2,2": (1, 1), "(2,2)": (1, 1), "_{2,2}": (1, 1),
            "3,3": (2, 2), "(3,3)": (2, 2), "_{3,3}": (2, 2)}
    for key, ij in (("1,2", (0, 1)), ("1,3", (0, 2)), ("3,2", (1, 2))):
        a, b = key.split(",")
        for k in (f"{a},{b}", f"({a},{b})", f"_{{{a},{b}}}", f"{b},{a}", f"({b},{a})", f"_{{{b},{a}}}"):
            comp[k] = ij
    if o in comp:
        p, q = comp[o]
        return S[:, p, q].copy()
    raise ValueError(f"unknown stress option {o!r}")


class FEMSolver:
    """FEMSolver_ (FEMSolverClass.f90:13-164) for one domain, state resident on the GPU."""

    def __init__(self, ctx: L.Context | None = None):
        self.ctx = ctx or L.default_context()
        lib = L.load()
        h = C.c_void_p()
        L.check(lib.pf_system_create(self.ctx.h, C.byref(h)), "pf_system_create")
        self.h = h
        self.itrmax = 100000
        self.er0 = L.ER0_DEFAULT
        self.relative_er = L.ER0_DEFAULT
        self.debug = False
        self.domain: FEMDomain | None = None
        self.DOF = 0
        self.n = 0
        self.nnz = 0
        self.last_stats: dict | None = None
        self._pending_ids: list = []
        self._pending_mats: list = []

    def close(self):
        if getattr(self, "h", None):
            L.load().pf_system_destroy(self.h)
            self.h = None

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass

    # (1)-(2)
    def init(self, NumDomain: int = 1):
        """init(NumDomain) (FEMSolverClass.f90:167-210).  With NumDomain >= 2 the domains are
        assembled as one system (setDomain per slot, then the multi-domain setCRS)."""
        if NumDomain < 1:
            raise ValueError("init: NumDomain >= 1")
        self.FEMDomains: list = [None] * int(NumDomain)
        self.CRS_ID_Starts_From = None
        return self

    @property
    def _multi(self) -> bool:
        return len(getattr(self, "FEMDomains", [None])) >= 2

    def setDomain(self, FEMDomain: FEMDomain | None = None, DomainID: int = 1, FEMDomains=None, DomainIDs=None,
                  FEMDomainPointers=None):
        """setDomain (FEMSolverClass.f90:329-419): one domain into slot DomainID, a list into
        DomainIDs (or slots 1..n for FEMDomainPointers)."""
        if FEMDomainPointers is not None:
            FEMDomains, DomainIDs = list(FEMDomainPointers), list(range(1, len(FEMDomainPointers) + 1))
        if FEMDomains is not None:
            if DomainIDs is None:
                DomainIDs = list(range(1, len(FEMDomains) + 1))
            for d, i in zip(FEMDomains, DomainIDs):
                self.setDomain(FEMDomain=d, DomainID=int(i))
            return self
        if not self._multi:
            if DomainID != 1:
                raise ValueError("setDomain: DomainID outside 1..NumDomain (call init(NumDomain) first)")
            self.domain = FEMDomain
            xyz, en = FEMDomain._flat()
            L.check(L.load().pf_system_set_mesh(self.h, FEMDomain.nn(), FEMDomain.ne(), FEMDomain.nne(),
                                                L.ptr(xyz, C.c_double), L.ptr(en, C.c_int32)), "setDomain")
            return self
        if not 1 <= DomainID <= len(self.FEMDomains):
            raise ValueError("setDomain: invalid domain ID")  # "invalid domain IDs" stop
        self.FEMDomains[DomainID - 1] = FEMDomain
        return self

    def _union(self):
        """The domains concatenated: node / element offsets per slot and the joint mesh."""
        doms = self.FEMDomains
        if any(d is None or d.nn() == 0 for d in doms):
            raise ValueError("setCRS: every domain slot needs a non-empty mesh")
        nn = np.array([d.nn() for d in doms], np.int64)
        ne = np.array([d.ne() for d in doms], np.int64)
        self._node_off = np.concatenate([[0], np.cumsum(nn)])
        self._elem_off = np.concatenate([[0], np.cumsum(ne)])
        xyz = np.concatenate([d.nodcoord for d in doms], axis=0)
        en = np.concatenate([d.elemnod + int(o) for d, o in zip(doms, self._node_off[:-1])], axis=0)
        self.domain = FEMDomain(xyz, en, self.ctx)
        return self.domain

    def _connects(self):
        """Every active connect of every domain in domain order, as element ids of the joint mesh:
        (own, Gauss point, paired, positions); any_overset = some domain holds OversetConnect."""
        own, gp, pr, pos = [], [], [], []
        any_overset = False
        for d in self.FEMDomains:
            oc = d.OversetConnect
            if oc is None:
                continue
            any_overset = True
            if oc.n == 0:
                continue
            p2p = oc.GaussPointID == 0  # own node of the joint mesh instead of the own element
            own.append(np.where(p2p, oc.NodeID + self._node_off[oc.DomainIDs12[:, 0] - 1],
                                oc.ElementID + self._elem_off[oc.DomainIDs12[:, 0] - 1]).astype(np.int32))
            pr.append(oc.PairElementID + self._elem_off[oc.DomainIDs12[:, 1] - 1].astype(np.int32))
            gp.append(oc.GaussPointID)
            pos.append(oc.position)
        if not own:
            z = np.zeros(0, np.int32)
            return any_overset, z, z, z, np.zeros((0, 3))
        return (any_overset, _i32(np.concatenate(own)), _i32(np.concatenate(gp)), _i32(np.concatenate(pr)),
                _f64(np.concatenate(pos)))

    # (3)
    def setCRS(self, DOF: int | None = None, CRS: CRS | None = None):
        lib = L.load()
        if CRS is not None:
            L.check(lib.pf_system_set_crs(self.h, CRS.n, CRS.nnz, L.ptr(CRS.row_ptr, C.c_int64),
                                          L.ptr(CRS.col_idx, C.c_int32), L.ptr(_f64(CRS.val), C.c_double)), "setCRS")
            self.n, self.nnz, self.DOF = CRS.n, CRS.nnz, 0
            return self
        nnz = C.c_int64(0)
        if self._multi:
            # multi-domain setCRS (FEMSolverClass.f90:466-676): joint mesh, block pattern
            # merged with the overset node pairs, columns ascending
            u = self._union()
            xyz, en = u._flat()
            L.check(lib.pf_system_set_mesh(self.h, u.nn(), u.ne(), u.nne(), L.ptr(xyz, C.c_double),
                                           L.ptr(en, C.c_int32)), "setCRS")
            any_ov, own, gp, pr, _ = self._connects()
            L.check(lib.pf_system_multi_crs(self.h, int(DOF), int(any_ov), own.size, L.ptr(own, C.c_int32),
                                            L.ptr(gp, C.c_int32), L.ptr(pr, C.c_int32), C.byref(nnz)), "setCRS")
            self.CRS_ID_Starts_From = (1 + int(DOF) * self._node_off[:-1]).astype(np.int64)
        else:
            L.check(lib.pf_system_zero_matrix(self.h, int(DOF), C.byref(nnz)), "setCRS")
        self.DOF = int(DOF)
        self.n = self.DOF * self.domain.nn()
        self.nnz = nnz.value
        return self

    def setEbOM(self, penalty: float, DOF: int = 3):
        """setEbOM(penalty, DOF) (FEMSolverClass.f90:1891-1985): penalty * ConnectMatrix of every
        connect of every domain, in order, added on the GPU."""
        self._flush()
        if not self._multi:
            return self  # one domain: no connects to couple
        _, own, gp, pr, pos = self._connects()
        L.check(L.load().pf_system_ebom(self.h, own.size, L.ptr(own, C.c_int32), L.ptr(gp, C.c_int32),
                                        L.ptr(pr, C.c_int32), L.ptr(pos, C.c_double), float(penalty), int(DOF)),
                "setEbOM")
        return self

    def setRHS(self, RHS):
        self._flush()
        L.check(L.load().pf_system_set_rhs(self.h, L.ptr(_f64(RHS), C.c_double)), "setRHS")

    def getRHS(self) -> np.ndarray:
        out = np.empty(self.n)
        L.check(L.load().pf_system_get_rhs(self.h, L.ptr(out, C.c_double)), "getRHS")
        return out

    def getX(self) -> np.ndarray:
        out = np.empty(self.n)
        L.check(L.load().pf_system_get_x(self.h, L.ptr(out, C.c_double)), "getX")
        return out

    def spmv(self, kind: int, x=None, repeat: int = 1, download: bool = True):
        """y = A x with kernel kind (1 CRS FP64, 2 blocked FP64, 3 blocked FP32, 4 CRS FP64 with x staged
        in LDS per row block, bitwise kind 1); returns (y, ms/launch)."""
        y = np.empty(self.n) if download else None
        ms = C.c_double(0)
        L.check(L.load().pf_system_spmv(self.h, int(kind), L.ptr(None if x is None else _f64(x), C.c_double),
                                        L.ptr(y, C.c_double), int(repeat), C.byref(ms)), "spmv")
        return y, ms.value

    def getCRS(self) -> CRS:
        self._flush()
        rp = np.empty(self.n + 1, dtype=np.int64)
        ci = np.empty(self.nnz, dtype=np.int32)
        va = np.empty(self.nnz, dtype=np.float64)
        L.check(L.load().pf_system_get_crs(self.h, L.ptr(rp, C.c_int64), L.ptr(ci, C.c_int32), L.ptr(va, C.c_double)),
                "getCRS")
        return CRS(rp, ci, va)

    # (4) element contributions are buffered and scattered on the device in call order
    def setMatrix(self, DomainID: int, ElementID: int, DOF: int, Matrix):
        if self.n == 0:
            self.setCRS(DOF=DOF)
        if self._multi:  # setValue multi-domain branch: ids offset by CRS_ID_Starts_From
            ElementID = int(ElementID) + int(self._elem_off[int(DomainID) - 1])
        self._pending_ids.append(int(ElementID))
        self._pending_mats.append(np.asfortranarray(Matrix, dtype=np.float64).ravel(order="F"))
        if len(self._pending_ids) >= 65536:
            self._flush()

    def _flush(self):
        if not self._pending_ids:
            return
        ids = np.asarray(self._pending_ids, dtype=np.int32)
        ke = np.ascontiguousarray(np.concatenate(self._pending_mats))
        self._pending_ids, self._pending_mats = [], []
        L.check(L.load().pf_system_add_element_matrices(self.h, ids.size, L.ptr(ids, C.c_int32),
                                                        L.ptr(ke, C.c_double)), "setMatrix")

    def setVector(self, DomainID: int, ElementID, DOF: int, Vector):
        self._flush()
        ids = np.atleast_1d(np.asarray(ElementID, dtype=np.int32))
        if self._multi:
            ids = (ids + int(self._elem_off[int(DomainID) - 1])).astype(np.int32)
        vec = np.ascontiguousarray(np.asarray(Vector, dtype=np.float64).reshape(ids.size, -1))
        L.check(L.load().pf_system_add_element_vectors(self.h, ids.size, L.ptr(ids, C.c_int32),
                                                       L.ptr(vec, C.c_double)), "setVector")

    def setStiffness(self, YoungModulus=None, PoissonRatio=None, jmode: int = 0):
        """Device assembly of StiffnessMatrix(YoungModulus(:),PoissonRatio(:)) into this solver's CRS.
        Without arguments the materials of the previous call (resident on the device) are reused."""
        self._flush()
        if self._multi and isinstance(YoungModulus, (list, tuple)):
            YoungModulus = np.concatenate([np.broadcast_to(np.asarray(e, np.float64), (d.ne(),))
                                           for e, d in zip(YoungModulus, self.FEMDomains)])
        if self._multi and isinstance(PoissonRatio, (list, tuple)):
            PoissonRatio = np.concatenate([np.broadcast_to(np.asarray(e, np.float64), (d.ne(),))
                                           for e, d in zip(PoissonRatio, self.FEMDomains)])
        E = None if YoungModulus is None else _f64(YoungModulus)
        v = None if PoissonRatio is None else _f64(PoissonRatio)
        L.check(L.load().pf_system_stiffness(self.h, L.ptr(E, C.c_double), L.ptr(v, C.c_double), jmode),
                "setStiffness")
        self.DOF = 3
        self.n = 3 * self.domain.nn()
        n = C.c_int64(0)
        nnz = C.c_int64(0)
        L.check(L.load().pf_system_sizes(self.h, C.byref(n), C.byref(nnz), None, None), "sizes")
        self.nnz = nnz.value

    def addTraction(self, box, pressure: float):
        """Sum of TractionVector_by_elem over elements (open box) added to the RHS."""
        b = _f64(box)
        L.check(L.load().pf_system_add_traction(self.h, L.ptr(b, C.c_double), float(pressure)), "addTraction")

    # (5)
    def fix(self, DomainID: int = 1, IDs=None, FixValue: float | None = None, FixValues=None):
        """fix(IDs, FixValue|FixValues) (FEMSolverClass.f90:1073-1136): IDs are DOF ids of the
        whole system.  With several domains the reference's DomainID >= 2 form marks IDs but
        stores the values at IDs + (nodes of the earlier domains); that form is refused here."""
        if self._multi and DomainID is not None and int(DomainID) >= 2:
            raise NotImplementedError("fix(DomainID>=2): pass DOF ids of the whole system (offset by "
                                      "CRS_ID_Starts_From - 1) without DomainID")
        ids = _i32(IDs)
        vals = None if FixValues is None else _f64(FixValues)
        L.check(L.load().pf_system_fix(self.h, ids.size, L.ptr(ids, C.c_int32), L.ptr(vals, C.c_double),
                                       0.0 if FixValue is None else float(FixValue)), "fix")

    # (7-2)
    def solve(self, x0=None, precision: int = 64, raise_on_breakdown: bool = True, download: bool = True,
              preconditioning: str | None = None, jacobi_solve: bool = True, algorithm: str | None = None,
              refine: int = 1, **opts):
        """solve([algorithm], [preconditioning], [x0]) (FEMSolverClass.f90:1697-1825).
        download=False keeps x on the device (returns None).

        preconditioning (FEMSolverClass.f90:1768-1790): None -> bicgstab_CRS_2; "ILU" or
        "incompleteLU" -> bicgstab_CRS_ILU with the ILU(0) factor kept from the first such solve
        (iluReset() drops it); "PointJacobi" -> JacobiPreconditionerCRS on the eliminated system,
        then bicgstab_CRS_2 on the scaled system.  The reference's PointJacobi branch calls no
        solver after the scaling (its x is x0 with the fixed values); jacobi_solve=False does
        exactly that.

        algorithm (FEMSolverClass.f90:1793-1808, used when no preconditioning is given):
        "GaussJordan" -> a direct solve (the reference's gauss_jordan_crs, FEMSolverClass.f90:
        2413-2460, stops "Under construction"; here a banded Cholesky factorization after a
        bandwidth-reducing renumbering, kept while the matrix is unchanged, with `refine` steps
        of FP64 iterative refinement); any other value -> bicgstab_CRS_2."""
        self._flush()
        if preconditioning is None and algorithm == "GaussJordan":
            st = L.Stats()
            x = np.empty(self.n) if download else None
            code = L.load().pf_system_solve_direct(self.h, int(refine), L.ptr(x, C.c_double), C.byref(st))
            self.last_stats = st.as_dict()
            L.check(code, "solve (GaussJordan)")
            return x
        if preconditioning is not None:
            if preconditioning == "PointJacobi":
                opts["preconditioning"] = 1 if jacobi_solve else 3
            elif preconditioning in ("ILU", "incompleteLU"):
                opts["preconditioning"] = 2
            else:  # the reference prints "invalid preconditioning" and stops
                raise ValueError(f"FEMSolver :: invalid preconditioning {preconditioning!r}")
        o = L.default_opts(itrmax=int(self.itrmax), er0=float(self.er0), relative_er=float(self.relative_er),
                           debug=int(bool(self.debug)), precision=int(precision), **opts)
        st = L.Stats()
        x = np.empty(self.n) if download else None
        code = L.load().pf_system_solve(self.h, C.byref(o), L.ptr(None if x0 is None else _f64(x0), C.c_double),
                                        L.ptr(x, C.c_double), C.byref(st))
        self.last_stats = st.as_dict()
        ok = (L.PF_OK, L.PF_NOT_CONVERGED) if raise_on_breakdown else (
            L.PF_OK, L.PF_NOT_CONVERGED, L.PF_BREAKDOWN_VV, L.PF_BREAKDOWN_C2C3)
        L.check(code, "solve (Bicgstab devide by zero)" if code in (1, 2) else "solve", ok=ok)
        return x


    def setMassMatrix(self, Density=1.0):
        """MassMatrix_as_CRS(Density(:), DOF=3) assembled into this solver's CRS values
        (FEMDomainClass.f90:10364-10467); Density one value or one per element."""
        self._flush()
        if np.ndim(Density) == 0:
            L.check(L.load().pf_system_mass_matrix(self.h, None, float(Density)), "setMassMatrix")
        else:
            L.check(L.load().pf_system_mass_matrix(self.h, L.ptr(_f64(Density), C.c_double), 0.0), "setMassMatrix")
        if self.n == 0:
            self.DOF = 3
            self.n = 3 * self.domain.nn()
            nnz = C.c_int64(0)
            L.check(L.load().pf_system_sizes(self.h, None, C.byref(nnz), None, None), "sizes")
            self.nnz = nnz.value

    def zeros(self):
        """zerosFEMSolver: CRS values and RHS set to 0."""
        self._flush()
        L.check(L.load().pf_system_zeros(self.h), "zeros")

    def keepThisMatrixAs(self, As: str):
        """keepThisMatrixAs: keep a named copy of the CRS values ("A" stiffness, "B" mass for eig)."""
        self._flush()
        L.check(L.load().pf_system_keep_matrix(self.h, As.encode()), "keepThisMatrixAs")

    def fix_eig(self, IDs):
        """fix_eig: DOFs (1-based) held at zero amplitude in eig."""
        ids = _i32(IDs)
        L.check(L.load().pf_system_fix_eig(self.h, ids.size, L.ptr(ids, C.c_int32)), "fix_eig")

    def eig(self, num_eigen: int | None = None, tol: float = 1e-8, maxit: int = 10000):
        """eigFEMSolver: (eigen_value, eigen_vectors) of A u = lambda B u (kept "A", "B"), the
        fix_eig DOFs removed.  num_eigen None: every pair (dense, as the reference's LAPACK
        path); an integer: the num_eigen smallest by LOBPCG.  eigen_vectors is (n, count),
        B-orthonormal, zero at the removed DOFs."""
        lib = L.load()
        if num_eigen is None:
            c = C.c_int32(0)
            L.check(lib.pf_system_eig_count(self.h, C.byref(c)), "eig")
            cnt, nev = c.value, 0
        else:
            cnt, nev = int(num_eigen), int(num_eigen)
        vals = np.empty(max(cnt, 1))
        vecs = np.empty((max(cnt, 1), self.n))  # row l = eigenvector l (column-major n x count)
        got = C.c_int32(0)
        st = L.Stats()
        code = lib.pf_system_eig(self.h, nev, float(tol), int(maxit), L.ptr(vals, C.c_double),
                                 L.ptr(vecs, C.c_double), C.byref(got), C.byref(st))
        self.last_stats = st.as_dict()
        L.check(code, "eig", ok=(L.PF_OK, L.PF_NOT_CONVERGED))
        k = got.value
        return vals[:k].copy(), np.ascontiguousarray(vecs[:k].T)

    def iluFactor(self):
        """(factor values on the CRS pattern, number of schedule levels) of the ILU(0) factor the
        solver keeps (computed now from the eliminated matrix when none is kept)."""
        self._flush()
        f = np.empty(self.nnz)
        lv = C.c_int32(0)
        L.check(L.load().pf_system_ilu_factor(self.h, L.ptr(f, C.c_double), C.byref(lv)), "iluFactor")
        return f, int(lv.value)

    def iluApply(self, v) -> np.ndarray:
        """ILU_matvec of the kept factor: D^-1 L^-1 v."""
        self._flush()
        y = np.empty(self.n)
        L.check(L.load().pf_system_ilu_apply(self.h, L.ptr(_f64(v), C.c_double), L.ptr(y, C.c_double)), "iluApply")
        return y

    def iluReset(self):
        """Drop the kept ILU(0) factor; the next ILU solve factors the current matrix."""
        L.check(L.load().pf_system_ilu_reset(self.h), "iluReset")

    def saveMatrix(self, name: str, CRS_as_dense: bool = False, zero_or_nonzero: bool = False):
        """saveMatrixFEMSolver (FEMSolverClass.f90:1225-1307): name_data.txt (one value per line),
        name_indices.txt and name_indptr.txt (one integer and a blank per line, 1-based), or
        name_dense.csv with CRS_as_dense ("0"/"*" per entry with zero_or_nonzero)."""
        crs = self.getCRS()
        if CRS_as_dense:
            with open(name + "_dense.csv", "w") as f:
                for i in range(crs.n):
                    row = np.zeros(crs.n)
                    k0, k1 = crs.row_ptr[i] - 1, crs.row_ptr[i + 1] - 1
                    row[crs.col_idx[k0:k1] - 1] = crs.val[k0:k1]
                    if zero_or_nonzero:
                        f.write("".join("0" if v == 0.0 else "*" for v in row) + "\n")
                    else:
                        f.write(",".join(repr(float(v)) for v in row) + "\n")
            return
        with open(name + "_data.txt", "w") as f:
            f.write("".join(f"  {float(v)!r}\n" for v in crs.val))
        with open(name + "_indices.txt", "w") as f:
            f.write("".join(f"{int(c)} \n" for c in crs.col_idx))
        with open(name + "_indptr.txt", "w") as f:
            f.write("".join(f"{int(r)} \n" for r in crs.row_ptr))

    def addBodyForce(self, Density=1.0, Accel=(0.0, 0.0, 0.0)):
        """setVector(MassVector(ElementID=e, Density, DOF=3, Accel)) for every element in
        ascending order (Density: one value or one per element)."""
        lib = L.load()
        acc = _f64(Accel)
        if np.ndim(Density) == 0:
            L.check(lib.pf_system_add_body_force(self.h, None, float(Density), L.ptr(acc, C.c_double)), "addBodyForce")
        else:
            d = _f64(Density)
            L.check(lib.pf_system_add_body_force(self.h, L.ptr(d, C.c_double), 0.0, L.ptr(acc, C.c_double)),
                    "addBodyForce")

    def elementStress(self, E, v, jmode: int = 0) -> np.ndarray:
        """Cell-averaged Cauchy stress (ne, 3, 3) of the current solution x."""
        lib = L.load()
        Ea, va = _f64(np.atleast_1d(E)), _f64(np.atleast_1d(v))
        ne = self.domain.ne()
        sig = np.empty(ne * 9)
        L.check(lib.pf_system_element_stress(self.h, L.ptr(Ea, C.c_double), Ea.size, L.ptr(va, C.c_double), va.size,
                                             jmode, L.ptr(sig, C.c_double)), "elementStress")
        return sig.reshape(ne, 3, 3).transpose(0, 2, 1)

    # ---- batched form: independent meshes concatenated into this system ----------------
    def addTractions(self, node_off, boxes, pressures):
        """One traction box and pressure per mesh (nodes node_off[m]..node_off[m+1]-1)."""
        off = _i64(node_off)
        bx = _f64(np.asarray(boxes, dtype=np.float64).reshape(-1, 6))
        pr = _f64(pressures)
        L.check(L.load().pf_system_add_tractions(self.h, off.size - 1, L.ptr(off, C.c_int64),
                                                 L.ptr(bx, C.c_double), L.ptr(pr, C.c_double)), "addTractions")

    def setBatch(self, node_off, cost=None):
        """Declare the mesh partition of a block-diagonal system (node offsets, 0-based)."""
        off = _i64(node_off)
        cst = None if cost is None else _f64(cost)
        self._nmesh = off.size - 1
        L.check(L.load().pf_system_batch_set(self.h, self._nmesh, L.ptr(off, C.c_int64),
                                             L.ptr(cst, C.c_double)), "setBatch")

    def autotuneBatch(self, iters: int = 20, precision: int = 64, **opts):
        """Time every launch shape of the persistent kernel; returns (best, [(shape, ms), ...])."""
        o = L.default_opts(itrmax=int(self.itrmax), er0=float(self.er0), relative_er=float(self.relative_er),
                           precision=int(precision), **opts)
        best = L.BatchTune()
        tried = (L.BatchTune * 32)()
        ms = np.zeros(32)
        nt = C.c_int32(0)
        L.check(L.load().pf_system_batch_autotune(self.h, C.byref(o), int(iters), C.byref(best), 32, tried,
                                                  L.ptr(ms, C.c_double), C.byref(nt)), "autotuneBatch")
        return best.as_dict(), [(tried[i].as_dict(), float(ms[i])) for i in range(min(nt.value, 32))]

    def solveBatch(self, precision: int = 64, tune: dict | None = None, download: bool = True,
                   raise_on_breakdown: bool = True, preconditioning: str | None = None,
                   algorithm: str | None = None, refine: int = 1, **opts):
        """Solve every mesh of the batch in one persistent kernel; returns (x, per-mesh stats).
        preconditioning="PointJacobi" scales the eliminated rows first (JacobiPreconditionerCRS,
        in place) and solves the scaled systems.  algorithm="GaussJordan": the direct solve of
        solve(algorithm="GaussJordan") for every mesh, one workgroup per mesh."""
        self._flush()
        if preconditioning is None and algorithm == "GaussJordan":
            st = np.zeros(self._nmesh, dtype=L.BATCH_STATS_DTYPE)
            tot = L.Stats()
            x = np.empty(self.n) if download else None
            code = L.load().pf_system_batch_direct(self.h, int(refine), L.ptr(x, C.c_double),
                                                   st.ctypes.data_as(C.c_void_p), C.byref(tot))
            self.last_stats = tot.as_dict()
            L.check(code, "solveBatch (GaussJordan)",
                    ok=(L.PF_OK,) if raise_on_breakdown else (L.PF_OK, L.PF_NOT_SPD))
            return x, st

        if preconditioning is not None:
            if preconditioning != "PointJacobi":
                raise ValueError(f"solveBatch :: invalid preconditioning {preconditioning!r}")
            opts["preconditioning"] = 1
        o = L.default_opts(itrmax=int(self.itrmax), er0=float(self.er0), relative_er=float(self.relative_er),
                           debug=int(bool(self.debug)), precision=int(precision), **opts)
        t = None if tune is None else L.BatchTune(**tune)
        st = np.zeros(self._nmesh, dtype=L.BATCH_STATS_DTYPE)
        tot = L.Stats()
        x = np.empty(self.n) if download else None
        code = L.load().pf_system_batch_solve(self.h, C.byref(o), None if t is None else C.byref(t),
                                              L.ptr(x, C.c_double), st.ctypes.data_as(C.c_void_p), C.byref(tot))
        self.last_stats = tot.as_dict()
        ok = (L.PF_OK, L.PF_NOT_CONVERGED) if raise_on_breakdown else (
            L.PF_OK, L.PF_NOT_CONVERGED, L.PF_BREAKDOWN_VV, L.PF_BREAKDOWN_C2C3)
        L.check(code, "solveBatch", ok=ok)
        return x, st

    def directInfo(self, mesh: int = 0, factor: bool = False) -> dict:
        """Ordering of mesh `mesh` of the last direct solve (0: the whole system) and, with
        factor=True, its Cholesky factor L in LAPACK lower band layout ((b+1) x n, row d of
        column c = L[c+d][c] of the renumbered matrix)."""
        n = C.c_int64(0)
        b = C.c_int32(0)
        kind = C.c_int32(0)
        L.check(L.load().pf_system_direct_info(self.h, int(mesh), C.byref(n), C.byref(b), C.byref(kind), None,
                                               None), "directInfo")
        perm = np.empty(n.value, dtype=np.int32)
        ab = np.empty((n.value, b.value + 1)) if factor else None
        L.check(L.load().pf_system_direct_info(self.h, int(mesh), None, None, None, L.ptr(perm, C.c_int32),
                                               L.ptr(ab, C.c_double)), "directInfo")
        out = {"n": n.value, "bandwidth": b.value, "ordering": ("natural", "rcm", "x", "y", "z")[kind.value],
               "perm": perm}
        if factor:
            out["factor"] = ab.T
        return out

    def setDirectKernel(self, mfma: bool = True):
        """Tile products of the direct factorization: MFMA (default) or vector FMAs."""
        L.check(L.load().pf_system_direct_kernel(self.h, int(bool(mfma))), "setDirectKernel")

    def setKeKernel(self, kind: int = 0):
        """Element-matrix kernel of later setStiffness calls: 0 the reference's operation order of
        StiffnessMatrixFEMDomain (FEMDomainClass.f90:11457-11515, bitwise CRS values), 1 the
        isotropic split Ke = lam M + mu M^T + mu tr(M) I per node pair (M = sum detJ G G^T) on
        the vector units, 2 the same contraction on the FP64 matrix cores (1, 2: to rounding)."""
        L.check(L.load().pf_system_ke_kernel(self.h, int(kind)), "setKeKernel")


def bicgstab_crs(row_ptr, col_idx, val, rhs, x0=None, ctx: L.Context | None = None, **opts):
    """Plug-in LinearSolver(row_ptr, col_idx, val, rhs, x) (FEMSolverClass.f90:1638-1648) on the GPU.

    Returns (x, stats dict).  No Dirichlet elimination, like the reference hook."""
    ctx = ctx or L.default_context()
    rp = np.ascontiguousarray(row_ptr, dtype=np.int64)
    ci = np.ascontiguousarray(col_idx, dtype=np.int32)
    va = _f64(val)
    b = _f64(rhs)
    x = np.zeros(b.size) if x0 is None else np.array(x0, dtype=np.float64)
    pc = opts.pop("preconditioning", None)
    if isinstance(pc, str):
        pc = {"PointJacobi": 1, "ILU": 2, "incompleteLU": 2}[pc]
    if pc is not None:
        opts["preconditioning"] = int(pc)
    o = L.default_opts(**opts)
    st = L.Stats()
    code = L.load().pf_bicgstab_crs(ctx.h, b.size, L.ptr(rp, C.c_int64), L.ptr(ci, C.c_int32), L.ptr(va, C.c_double),
                                    L.ptr(b, C.c_double), L.ptr(x, C.c_double), C.byref(o), C.byref(st))
    L.check(code, "bicgstab_crs", ok=(L.PF_OK, L.PF_NOT_CONVERGED, L.PF_BREAKDOWN_VV, L.PF_BREAKDOWN_C2C3))
    return x, st.as_dict()


def jacobi_preconditioner(row_ptr, col_idx, val, rhs=None, ctx: L.Context | None = None):
    """JacobiPreconditionerCRS (FEMSolverClass.f90:2508-2529) on copies: returns (val, rhs)."""
    ctx = ctx or L.default_context()
    rp = np.ascontiguousarray(row_ptr, dtype=np.int64)
    va = np.array(val, dtype=np.float64)
    b = None if rhs is None else np.array(rhs, dtype=np.float64)
    L.check(L.load().pf_jacobi_crs(ctx.h, rp.size - 1, L.ptr(rp, C.c_int64), L.ptr(_i32(col_idx), C.c_int32),
                                   L.ptr(va, C.c_double), L.ptr(b, C.c_double)), "jacobi_preconditioner")
    return va, b


def ilu0(row_ptr, col_idx, val, ctx: L.Context | None = None) -> np.ndarray:
    """ILUCRS(fill_in_order=0) (SparseClass.f90:1629-1700): the factor values on the same pattern."""
    ctx = ctx or L.default_context()
    rp = np.ascontiguousarray(row_ptr, dtype=np.int64)
    f = np.empty(int(rp[-1] - 1))
    L.check(L.load().pf_ilu0_crs(ctx.h, rp.size - 1, L.ptr(rp, C.c_int64), L.ptr(_i32(col_idx), C.c_int32),
                                 L.ptr(_f64(val), C.c_double), L.ptr(f, C.c_double)), "ilu0")
    return f


def ilu_apply(row_ptr, col_idx, fval, v, ctx: L.Context | None = None) -> np.ndarray:
    """ILU_matvecCRS (SparseClass.f90:1795-1845) of the factor fval: D^-1 L^-1 v."""
    ctx = ctx or L.default_context()
    rp = np.ascontiguousarray(row_ptr, dtype=np.int64)
    y = np.empty(rp.size - 1)
    L.check(L.load().pf_ilu_apply_crs(ctx.h, rp.size - 1, L.ptr(rp, C.c_int64), L.ptr(_i32(col_idx), C.c_int32),
                                      L.ptr(_f64(fval), C.c_double), L.ptr(_f64(v), C.c_double),
                                      L.ptr(y, C.c_double)), "ilu_apply")
    return y


def spmv_crs(row_ptr, col_idx, val, x, ctx: L.Context | None = None) -> np.ndarray:
    """sub_crs_matvec (LinearSolverClass.f90:3728-3776) on host arrays."""
    ctx = ctx or L.default_context()
    rp = np.ascontiguousarray(row_ptr, dtype=np.int64)
    y = np.empty(rp.size - 1)
    L.check(L.load().pf_spmv_crs(ctx.h, rp.size - 1, L.ptr(rp, C.c_int64), L.ptr(_i32(col_idx), C.c_int32),
                                 L.ptr(_f64(val), C.c_double), L.ptr(_f64(x), C.c_double), L.ptr(y, C.c_double)),
            "spmv_crs")
    return y
