"""Synthetic static-elasticity workloads on Cube3D hex meshes (seeded, no data files).

box(n, sigma, nu, seed): unit cube with n^3 hex8 elements, E = 1e7 * lognormal(0, sigma)
per element, bottom face (z = 0) fully clamped and a uniform downward traction of 1e4 Pa
on the top face, as the static-analysis callers set it up (select(z_max=z_min()) for the
fixed nodes, TractionVector_by_elem over the top facets).
"""
from __future__ import annotations

from dataclasses import dataclass, field

import numpy as np

from .fem import PF_GLYCINE_MAX, cube3d, cylinder3d, leaf3d

TOP_TRACTION = 1.0e4  # Pa, downward


@dataclass
class Workload:
    name: str
    nodcoord: np.ndarray   # (nn,3) Fortran order
    elemnod: np.ndarray    # (ne,8) Fortran order, 1-based
    E: np.ndarray          # (ne,)
    nu: np.ndarray         # (ne,)
    fixed_dofs: np.ndarray  # 1-based DOF ids
    traction_box: np.ndarray  # open box (xmin,xmax,ymin,ymax,zmin,zmax)
    pressure: float
    meta: dict = field(default_factory=dict)

    @property
    def nn(self) -> int:
        return int(self.nodcoord.shape[0])

    @property
    def ne(self) -> int:
        return int(self.elemnod.shape[0])

    @property
    def ndof(self) -> int:
        return 3 * self.nn


def lognormal_E(ne: int, sigma: float, rng: np.random.Generator, base: float = 1.0e7,
                clip_ratio: float | None = None) -> np.ndarray:
    f = np.exp(sigma * rng.standard_normal(ne))
    if clip_ratio is not None:
        half = np.sqrt(clip_ratio)
        f = np.clip(f, 1.0 / half, half)
    return base * f


def box(n: int, sigma: float = 0.5, nu=0.3, seed: int = 0, clip_ratio: float | None = None,
        nu_range: tuple | None = None, nx=None, ny=None, nz=None) -> Workload:
    nx = nx or n
    ny = ny or n
    nz = nz or n
    nod, elem = cube3d(nx, ny, nz)
    rng = np.random.default_rng(seed)
    ne = elem.shape[0]
    E = lognormal_E(ne, sigma, rng, clip_ratio=clip_ratio)
    if nu_range is not None:
        nuv = rng.uniform(nu_range[0], nu_range[1], ne)
    else:
        nuv = np.full(ne, float(nu))
    bottom = np.nonzero(nod[:, 2] <= nod[:, 2].min())[0] + 1
    fixed = np.sort(np.concatenate([3 * bottom - 2, 3 * bottom - 1, 3 * bottom])).astype(np.int32)
    zt = nod[:, 2].max()
    eps = 1e-9 * max(1.0, abs(zt))
    tbox = np.array([-1e300, 1e300, -1e300, 1e300, zt - eps, zt + eps])
    return Workload(f"box{nx}x{ny}x{nz}", nod, elem, E, nuv, fixed, tbox, -TOP_TRACTION,
                    {"sigma": sigma, "seed": seed, "clip_ratio": clip_ratio, "nu_range": nu_range})


def config(k: int, seed: int = 0) -> Workload:
    """The four single-mesh configurations of the measurement plan (1-based like the list)."""
    if k == 1:   # 20^3, sigma 0.5, nu 0.3 -- the CPU-runnable case
        return box(20, sigma=0.5, nu=0.3, seed=seed)
    if k == 2:   # 100^3, sigma 1.0, nu ~ U[0.25, 0.35]
        return box(100, sigma=1.0, seed=seed, nu_range=(0.25, 0.35))
    if k == 3:   # 100^3, sigma 2.0 clipped to a 1e4 contrast, nu ~ U[0.25, 0.35]
        return box(100, sigma=2.0, seed=seed, nu_range=(0.25, 0.35), clip_ratio=1.0e4)
    if k == 5:   # 60^3 (adjoint case)
        return box(60, sigma=0.5, nu=0.3, seed=seed)
    raise ValueError(k)


def crs_sizes(nx: int, ny: int, nz: int) -> tuple[int, int]:
    """(n, nnz) of ZeroMatrix(DOF=3) on an nx x ny x nz Cube3D: nnz = 9 prod(3 m_i - 2), m_i = n_i + 1."""
    nn = (nx + 1) * (ny + 1) * (nz + 1)
    return 3 * nn, 9 * (3 * nx + 1) * (3 * ny + 1) * (3 * nz + 1)


@dataclass
class InverseProblem:
    """Recover per-element Young's moduli from displacements observed on the surface."""
    nodcoord: np.ndarray
    elemnod: np.ndarray
    nu: np.ndarray
    fixed_dofs: np.ndarray   # 1-based DOF ids (clamped base)
    E_hidden: np.ndarray     # the field the observations come from
    F: np.ndarray            # (n, ncases) load cases: random top-face traction fields
    obs_nodes: np.ndarray    # 0-based observed surface nodes
    noise: float             # relative standard deviation of the observation noise
    seed: int

    @property
    def obs_dofs(self) -> np.ndarray:
        return np.sort(np.concatenate([3 * self.obs_nodes, 3 * self.obs_nodes + 1, 3 * self.obs_nodes + 2]))


def inverse_problem(n: int = 60, ncases: int = 64, nobs: int = 2000, seed: int = 0, sigma: float = 0.5,
                    noise: float = 0.01) -> InverseProblem:
    """n^3 unit-cube hex mesh clamped at z = 0; hidden E = 1e7 lognormal(sigma); ncases random
    top-face traction fields; nobs random surface nodes (not on the clamped face)."""
    from .autograd import surface_nodes, top_traction_fields  # numpy-only helpers
    nod, el = cube3d(n, n, n)
    rng = np.random.default_rng(seed)
    ne = el.shape[0]
    E = lognormal_E(ne, sigma, rng)
    bottom = np.nonzero(nod[:, 2] <= nod[:, 2].min())[0] + 1
    fixed = np.sort(np.concatenate([3 * bottom - 2, 3 * bottom - 1, 3 * bottom])).astype(np.int32)
    F = top_traction_fields(nod, el, ncases, rng)
    surf = surface_nodes(nod)
    obs = np.sort(rng.choice(surf, size=min(nobs, surf.size), replace=False))
    return InverseProblem(nod, el, np.full(ne, 0.3), fixed, E, F, obs, noise, seed)


@dataclass
class Plant:
    """Organ meshes of a synthetic plant and the overset pairs between them, in the form
    Soybean%getDisplacement consumes (SoybeanClass.f90:8340-8533): domain 1 is the stem,
    then petioles (overset onto the stem) and leaf blades (overset onto their petiole)."""
    meshes: list            # [(nodcoord (nn,3), elemnod (ne,8))], domain order
    pairs: list             # [(MyDomainID, DomainID)] in overset call order
    meta: dict = field(default_factory=dict)


def _shifted(xn, yn, zn, lx, ly, lz, origin):
    nod, el = cube3d(xn, yn, zn, lx, ly, lz)
    return np.asfortranarray(nod + np.asarray(origin, dtype=np.float64)), el


def plant(stem=(6, 6, 30), h=0.01, petioles: int = 8, petiole_len: int = 8, petiole_cells: int = 6,
          petiole_width: float = 0.03, leaf_cells: int = 8, leaf_layers: int = 3,
          leaf_thickness: float = 0.02) -> Plant:
    """A stem of stem[0] x stem[1] x stem[2] cells of size h standing on z = 0; petioles of
    petiole_len cells of length h and a square section petiole_width wide (petiole_cells cells
    across) leave it at rising heights towards +x, -x, +y, -y in turn, each ending under a
    square leaf blade (leaf_cells x leaf_cells cells of size h, leaf_layers through the
    thickness).  Small offsets keep every mesh's Gauss points off the faces of the mesh it
    overlaps."""
    sx, sy, sz = stem
    meshes = [_shifted(sx, sy, sz, sx * h, sy * h, sz * h, (0.0, 0.0, 0.0))]
    pairs = []
    cx, cy = 0.5 * sx * h, 0.5 * sy * h
    L, t, lw = petiole_len * h, petiole_width, leaf_cells * h
    for k in range(petioles):
        z = (0.35 + 0.55 * k / max(petioles, 1)) * sz * h
        d = k % 4
        if d == 0:
            org, dims = (cx + 0.0013, cy - t / 2 + 0.0007, z), (petiole_len, petiole_cells, petiole_cells, L, t, t)
            tip = np.array([org[0] + L, cy, z + t / 2])
        elif d == 1:
            org, dims = (cx - L - 0.0011, cy - t / 2 + 0.0009, z), (petiole_len, petiole_cells, petiole_cells, L, t, t)
            tip = np.array([org[0], cy, z + t / 2])
        elif d == 2:
            org, dims = (cx - t / 2 + 0.0006, cy + 0.0017, z), (petiole_cells, petiole_len, petiole_cells, t, L, t)
            tip = np.array([cx, org[1] + L, z + t / 2])
        else:
            org, dims = (cx - t / 2 + 0.0004, cy - L - 0.0012, z), (petiole_cells, petiole_len, petiole_cells, t, L, t)
            tip = np.array([cx, org[1], z + t / 2])
        meshes.append(_shifted(*dims, org))
        pairs.append((len(meshes), 1))
        lorg = tip - np.array([lw / 2 + 0.0008, lw / 2 + 0.0011, 0.0013])
        meshes.append(_shifted(leaf_cells, leaf_cells, leaf_layers, lw, lw, leaf_thickness, tuple(lorg)))
        pairs.append((len(meshes), len(meshes) - 1))
    nn = sum(m[0].shape[0] for m in meshes)
    ne = sum(m[1].shape[0] for m in meshes)
    return Plant(meshes, pairs, {"domains": len(meshes), "nodes": nn, "elements": ne, "dof": 3 * nn})


def plant_scaled(scale: int = 1) -> Plant:
    """plant() refined scale times in every direction (same geometry, scale^3 more cells)."""
    s = int(scale)
    return plant(stem=(6 * s, 6 * s, 30 * s), h=0.01 / s, petiole_len=8 * s, petiole_cells=6 * s,
                 leaf_cells=8 * s, leaf_layers=3 * s)


# leaf axes (width a, thickness b, length c) -> global axes for a blade pointing along +x, -x,
# +y, -y: proper rotations (determinant +1), so element orientation is kept
_LEAF_ROT = [np.array([[0, 0, 1], [1, 0, 0], [0, 1, 0]], float), np.array([[0, 0, -1], [-1, 0, 0], [0, 1, 0]], float),
             np.array([[-1, 0, 0], [0, 0, 1], [0, 1, 0]], float), np.array([[1, 0, 0], [0, 0, -1], [0, 1, 0]], float)]


def plant_organs(scale: int = 1, petioles: int = 8) -> Plant:
    """A plant of the reference's organ meshes: a Cylinder stem (StemClass), box petioles and
    soybean Leaf3D blades (LeafClass) laid flat at the petiole tips, pointing outwards; the
    petioles are overset onto the stem and the blades onto their petiole."""
    s = int(scale)
    sd, sl = 0.08, 0.6
    stem = cylinder3d(8 * s, 8 * s, 30 * s, sd, sd, sl)
    meshes = [(np.asfortranarray(stem[0]), stem[1])]
    pairs = []
    L, t = 0.10, 0.03
    for k in range(petioles):
        z = (0.35 + 0.55 * k / max(petioles, 1)) * sl
        d = k % 4
        off = 0.2 * sd  # petiole base inside the stem
        if d == 0:
            org, dims = (off + 0.0013, -t / 2 + 0.0007, z), (10 * s, 3 * s, 3 * s, L, t, t)
            tip = np.array([off + L, 0.0, z + t / 2])
        elif d == 1:
            org, dims = (-off - L - 0.0011, -t / 2 + 0.0009, z), (10 * s, 3 * s, 3 * s, L, t, t)
            tip = np.array([-off - L, 0.0, z + t / 2])
        elif d == 2:
            org, dims = (-t / 2 + 0.0006, off + 0.0017, z), (3 * s, 10 * s, 3 * s, t, L, t)
            tip = np.array([0.0, off + L, z + t / 2])
        else:
            org, dims = (-t / 2 + 0.0004, -off - L - 0.0012, z), (3 * s, 10 * s, 3 * s, t, L, t)
            tip = np.array([0.0, -off - L, z + t / 2])
        meshes.append(_shifted(*dims, org))
        pairs.append((len(meshes), 1))
        nod, el = leaf3d(8 * s, 2 * s, 16 * s, 0.12, 0.01, 0.2, species=PF_GLYCINE_MAX)
        # the blade's base (length 0 .. 0.02) reaches back over the petiole's last 2 cm
        g = nod @ _LEAF_ROT[d].T + tip - 0.02 * _LEAF_ROT[d][:, 2] + np.array([0.0011, 0.0007, 0.0003])
        meshes.append((np.asfortranarray(g), el))
        pairs.append((len(meshes), len(meshes) - 1))
    nn = sum(m[0].shape[0] for m in meshes)
    ne = sum(m[1].shape[0] for m in meshes)
    return Plant(meshes, pairs, {"domains": len(meshes), "nodes": nn, "elements": ne, "dof": 3 * nn,
                                 "organs": "Cylinder stem, box petioles, soybean Leaf3D blades"})
