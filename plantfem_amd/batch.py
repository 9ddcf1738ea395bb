"""Batched static solves of many independent organ meshes on one GPU.

Every organ mesh is solved on its own, as one single-domain FEMSolver_ per organ would
(setDomain / setCRS(DOF=3) / StiffnessMatrix / traction / fix / solve, FEMSolverClass.f90:
329-694, :1697-1825).  (The reference's whole-plant Soybean%getDisplacement,
SoybeanClass.f90:8417-8533, instead couples the organs into one multi-domain system with
overset penalty terms, setEbOM; that coupling is FEMDomain.overset + FEMSolver.setEbOM, not
this form.)  Here the meshes are
concatenated into one block-diagonal system: node ids of mesh m are shifted by its first
node, element lists and material arrays are concatenated, so graph build, element assembly,
tractions and fixes run once for the whole batch, and every mesh's CRS block is the
single-domain CRS of that mesh alone.  FEMSolver.solveBatch then solves all meshes in one
persistent-kernel launch.

organ_batch(count, seed) makes the synthetic plant-organ workload: leaf slabs of n x m x 2
hex8 elements (n, m in [20, 120]) clamped along one edge with a pressure on the upper face,
and stems of 4 x 4 x L elements (L in [50, 400]) clamped at the base with a pressure on the
top end; E = 5e6 Pa x lognormal(0.5) per element, nu = 0.3, pressures log-uniform in
[1e2, 1e4] Pa, elements of 1 mm edge.
"""
from __future__ import annotations

from dataclasses import dataclass

import numpy as np

from .fem import FEMDomain, FEMSolver, cube3d

H = 1.0e-3  # element edge (m)


@dataclass
class Organ:
    kind: str              # "leaf" or "stem"
    dims: tuple            # elements along x, y, z
    nodcoord: np.ndarray   # (nn,3) Fortran order
    elemnod: np.ndarray    # (ne,8) Fortran order, 1-based local ids
    E: np.ndarray
    nu: np.ndarray
    fixed_nodes: np.ndarray  # 1-based local node ids, all 3 DOFs fixed
    box: np.ndarray          # traction box (open), 6 values
    pressure: float

    @property
    def nn(self) -> int:
        return int(self.nodcoord.shape[0])

    @property
    def ne(self) -> int:
        return int(self.elemnod.shape[0])

    @property
    def ndof(self) -> int:
        return 3 * self.nn


def _face_box(nod, axis: int, hi: bool) -> np.ndarray:
    v = nod[:, axis].max() if hi else nod[:, axis].min()
    eps = 1e-6 * H
    b = np.array([-1e300, 1e300, -1e300, 1e300, -1e300, 1e300])
    b[2 * axis] = v - eps
    b[2 * axis + 1] = v + eps
    return b


def leaf(n: int, m: int, rng: np.random.Generator, E0: float = 5.0e6, sigma: float = 0.5) -> Organ:
    nod, el = cube3d(n, m, 2, n * H, m * H, 2 * H)
    ne = el.shape[0]
    E = E0 * np.exp(sigma * rng.standard_normal(ne))
    clamp = np.nonzero(nod[:, 0] <= nod[:, 0].min())[0] + 1   # petiole edge x = 0
    p = -float(np.exp(rng.uniform(np.log(1e2), np.log(1e4))))  # pushes the upper face down
    return Organ("leaf", (n, m, 2), nod, el, E, np.full(ne, 0.3), clamp, _face_box(nod, 2, True), p)


def stem(length: int, rng: np.random.Generator, E0: float = 5.0e6, sigma: float = 0.5) -> Organ:
    nod, el = cube3d(4, 4, length, 4 * H, 4 * H, length * H)
    ne = el.shape[0]
    E = E0 * np.exp(sigma * rng.standard_normal(ne))
    clamp = np.nonzero(nod[:, 2] <= nod[:, 2].min())[0] + 1   # base z = 0
    p = -float(np.exp(rng.uniform(np.log(1e2), np.log(1e4))))  # load on the top end, downward
    return Organ("stem", (4, 4, length), nod, el, E, np.full(ne, 0.3), clamp, _face_box(nod, 2, True), p)


def block(n: int, rng: np.random.Generator, E0: float = 1.0e7, sigma: float = 0.5) -> Organ:
    """n x n x n hex block clamped at its base with a pressure on its top face (a compact,
    well-conditioned mesh, e.g. a fruit or tuber section)."""
    nod, el = cube3d(n, n, n, n * H, n * H, n * H)
    ne = el.shape[0]
    E = E0 * np.exp(sigma * rng.standard_normal(ne))
    clamp = np.nonzero(nod[:, 2] <= nod[:, 2].min())[0] + 1
    p = -float(np.exp(rng.uniform(np.log(1e2), np.log(1e4))))
    return Organ("block", (n, n, n), nod, el, E, np.full(ne, 0.3), clamp, _face_box(nod, 2, True), p)


def organ_batch(count: int = 512, seed: int = 0, max_dof: int | None = 60000,
                leaf_range=(20, 120), stem_range=(50, 400)) -> list:
    """`count` organs from `seed`, half leaves and half stems in random order.  Leaf sizes
    are redrawn while the mesh would exceed max_dof (None: no cap)."""
    rng = np.random.default_rng(seed)
    out = []
    for _ in range(count):
        if rng.random() < 0.5:
            while True:
                n, m = (int(v) for v in rng.integers(leaf_range[0], leaf_range[1] + 1, 2))
                if max_dof is None or 9 * (n + 1) * (m + 1) <= max_dof:
                    break
            out.append(leaf(n, m, rng))
        else:
            out.append(stem(int(rng.integers(stem_range[0], stem_range[1] + 1)), rng))
    return out


@dataclass
class Union:
    nodcoord: np.ndarray
    elemnod: np.ndarray
    E: np.ndarray
    nu: np.ndarray
    node_off: np.ndarray   # (nmesh+1,) 0-based
    elem_off: np.ndarray
    fixed_dofs: np.ndarray  # 1-based global DOF ids
    boxes: np.ndarray       # (nmesh,6)
    pressures: np.ndarray

    @property
    def nmesh(self) -> int:
        return int(self.node_off.size - 1)

    def dof_range(self, m: int) -> slice:
        return slice(3 * int(self.node_off[m]), 3 * int(self.node_off[m + 1]))


def union(organs: list) -> Union:
    nn = np.array([o.nn for o in organs], dtype=np.int64)
    ne = np.array([o.ne for o in organs], dtype=np.int64)
    node_off = np.concatenate([[0], np.cumsum(nn)])
    elem_off = np.concatenate([[0], np.cumsum(ne)])
    nod = np.asfortranarray(np.concatenate([o.nodcoord for o in organs], axis=0))
    el = np.asfortranarray(np.concatenate([o.elemnod + np.int32(node_off[i]) for i, o in enumerate(organs)],
                                          axis=0).astype(np.int32))
    E = np.concatenate([o.E for o in organs])
    nu = np.concatenate([o.nu for o in organs])
    fixed_nodes = np.concatenate([o.fixed_nodes + node_off[i] for i, o in enumerate(organs)])
    fixed = np.sort(np.concatenate([3 * fixed_nodes - 2, 3 * fixed_nodes - 1, 3 * fixed_nodes])).astype(np.int32)
    boxes = np.stack([o.box for o in organs])
    pressures = np.array([o.pressure for o in organs])
    return Union(nod, el, E, nu, node_off, elem_off, fixed, boxes, pressures)


def build_solver(u: Union, ctx=None, cost=None) -> FEMSolver:
    """FEMSolver over the union: ZeroMatrix(DOF=3), element stiffness, tractions, fixes."""
    s = FEMSolver(ctx)
    s.setDomain(FEMDomain(u.nodcoord, u.elemnod, ctx))
    s.setStiffness(u.E, u.nu)
    s.addTractions(u.node_off, u.boxes, u.pressures)
    s.fix(1, u.fixed_dofs, FixValue=0.0)
    s.setBatch(u.node_off, cost)
    return s
