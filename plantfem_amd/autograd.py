"""Differentiable static solve: u(E) = K(E)^-1 f for many load cases, as a torch.autograd.Function.

Forward: device assembly of StiffnessMatrix(YoungModulus(:),PoissonRatio(:))
(FEMDomainClass.f90:11123-11232) from E = exp(log_E), Dirichlet elimination and
bicgstab_CRS_2 (FEMSolverClass.f90:1697-1825, :2199-2298) for R load cases at once.
Backward (hand-written adjoint): one more R-case solve K lam = dL/du (the eliminated matrix
P K P + (I - P) is symmetric, so K^T = K), then per element
    dL/d(log E_e) = -sum_l lam_e(l)^T Ke(E_e) u_e(l)
(Ke is linear in E) and dL/df = lam with the fixed rows zeroed.  The reference has no
adjoint; this is the gradient of its solve.

Tensors are float64 on the GPU; load cases are columns: F and U have shape (n, R) and are
stored row-major, i.e. interleaved (dof i of case l at i*R + l) as the library expects.
"""
from __future__ import annotations

import ctypes as C

import numpy as np
import torch

from . import _lib as L
from .fem import FEMDomain, FEMSolver


def _sync():
    torch.cuda.current_stream().synchronize()


class ElasticSystem:
    """Mesh, materials' Poisson ratios and clamped DOFs of a static problem whose Young's
    moduli are the differentiable parameters."""

    def __init__(self, nodcoord, elemnod, fixed_dofs, nu, ctx: L.Context | None = None, jmode: int = 0,
                 device: str = "cuda", **solve_opts):
        self.solver = FEMSolver(ctx)
        self.solver.setDomain(FEMDomain(nodcoord, elemnod, ctx))
        self.solver.setCRS(DOF=3)
        self.solver.fix(1, np.asarray(fixed_dofs, dtype=np.int32), FixValue=0.0)
        self.n = self.solver.n
        self.ne = int(np.asarray(elemnod).shape[0])
        self.device = device
        self.nu = torch.as_tensor(np.asarray(nu, dtype=np.float64), device=device).contiguous()
        self.jmode = jmode
        self.opts = dict(stop_mode=1, itrmax=100000, relative_er=1e-10)
        self.opts.update(solve_opts)
        self._assembled = None  # the E tensor the device matrix was assembled from
        self.last_stats = None
        self.solves = 0

    def assemble(self, E: torch.Tensor):
        E = E.detach().to(dtype=torch.float64).contiguous()
        if E.numel() != self.ne:
            raise ValueError("one Young's modulus per element expected")
        _sync()
        L.check(L.load().pf_system_stiffness_dev(self.solver.h, C.c_void_p(E.data_ptr()),
                                                 C.c_void_p(self.nu.data_ptr()), self.jmode), "stiffness")
        self._assembled = E

    def solve(self, F: torch.Tensor, x0: torch.Tensor | None = None, **opts) -> torch.Tensor:
        """U with K U = F for every column of F (n, R), fixed rows of U = 0.  x0 (n, R): initial
        guess; opts override the system's solver options for this call."""
        F = F.detach().to(dtype=torch.float64).contiguous()
        if F.dim() == 1:
            F = F[:, None]
        n, R = F.shape
        if n != self.n:
            raise ValueError("right-hand sides must have n rows")
        if x0 is None:
            U = torch.empty_like(F)
        else:
            U = x0.detach().to(dtype=torch.float64).reshape(n, R).clone().contiguous()
        kw = dict(self.opts)
        kw.update(opts)
        kw["use_x0"] = 0 if x0 is None else 1
        o = L.default_opts(**kw)
        per = (L.Stats * R)()
        tot = L.Stats()
        _sync()
        code = L.load().pf_system_solve_multi_dev(self.solver.h, R, C.c_void_p(F.data_ptr()),
                                                  C.c_void_p(U.data_ptr()), C.byref(o), per, C.byref(tot))
        self.last_stats = {"total": tot.as_dict(), "cases": [per[i].as_dict() for i in range(R)]}
        self.solves += 1
        L.check(code, "multi-RHS solve")
        return U

    def sensitivity(self, U: torch.Tensor, Lam: torch.Tensor) -> torch.Tensor:
        """s_e = sum_l Lam_e(l)^T Ke(E_e) U_e(l) (fixed DOFs excluded)."""
        U = U.detach().contiguous()
        Lam = Lam.detach().contiguous()
        out = torch.empty(self.ne, dtype=torch.float64, device=U.device)
        _sync()
        L.check(L.load().pf_system_elem_sensitivity_dev(self.solver.h, U.shape[1], C.c_void_p(U.data_ptr()),
                                                        C.c_void_p(Lam.data_ptr()), C.c_void_p(out.data_ptr())),
                "sensitivity")
        return out


class StaticSolve(torch.autograd.Function):
    """U = K(exp(log_E))^-1 F (load cases as columns)."""

    @staticmethod
    def forward(ctx, log_E, F, system: ElasticSystem):
        E = torch.exp(log_E.detach())
        system.assemble(E)
        U = system.solve(F)
        ctx.system = system
        ctx.save_for_backward(E, U)
        return U

    @staticmethod
    def backward(ctx, gU):
        E, U = ctx.saved_tensors
        system = ctx.system
        if system._assembled is None or not torch.equal(system._assembled, E):
            system.assemble(E)  # another forward ran in between
        Lam = system.solve(gU)
        g_logE = -system.sensitivity(U, Lam) if ctx.needs_input_grad[0] else None
        g_F = Lam if ctx.needs_input_grad[1] else None
        return g_logE, g_F, None


def static_solve(log_E: torch.Tensor, F: torch.Tensor, system: ElasticSystem) -> torch.Tensor:
    return StaticSolve.apply(log_E, F, system)


# ---- the inverse problem: recover E from observed surface displacements ----------------
def top_traction_fields(nodcoord, elemnod, ncases: int, rng: np.random.Generator, p0: float = 1.0e4,
                        modes: int = 4) -> np.ndarray:
    """(n, ncases) nodal forces of random smooth pressure fields on the top face z = max:
    for every top facet, p = -p0 (1 + 0.5 sum_k a_k cos(kx pi x) cos(ky pi y)) acting along
    the outward normal, distributed a quarter to each facet node (the traction rule of
    TractionVector_by_elemFEMDomain, FEMDomainClass.f90:17273-17287, for flat facets)."""
    nod = np.asarray(nodcoord)
    el = np.asarray(elemnod) - 1
    zt = nod[:, 2].max()
    top = el[:, 4:8]                                  # local nodes 5..8 form the upper facet
    on_top = np.all(np.abs(nod[top, 2] - zt) < 1e-12 * max(1.0, abs(zt)), axis=1)
    fac = top[on_top]                                 # (nf, 4)
    P = nod[fac]                                      # (nf, 4, 3)
    area = 0.5 * np.linalg.norm(np.cross(P[:, 1] - P[:, 0], P[:, 2] - P[:, 0]), axis=1) + \
        0.5 * np.linalg.norm(np.cross(P[:, 2] - P[:, 1], P[:, 3] - P[:, 1]), axis=1)
    cx, cy = P[:, :, 0].mean(axis=1), P[:, :, 1].mean(axis=1)
    lx = max(nod[:, 0].max() - nod[:, 0].min(), 1e-300)
    ly = max(nod[:, 1].max() - nod[:, 1].min(), 1e-300)
    nn = nod.shape[0]
    F = np.zeros((3 * nn, ncases))
    for c in range(ncases):
        field = np.zeros(fac.shape[0])
        for _ in range(modes):
            kx, ky = rng.integers(0, 4, 2)
            field += rng.uniform(-1, 1) * np.cos(kx * np.pi * cx / lx) * np.cos(ky * np.pi * cy / ly)
        p = -p0 * (1.0 + 0.5 * field / modes)
        fz = 0.25 * p * area                          # upper facet normal is +z
        np.add.at(F[:, c], 3 * fac.ravel() + 2, np.repeat(fz, 4))
    return F


def surface_nodes(nodcoord, exclude_z_min: bool = True) -> np.ndarray:
    nod = np.asarray(nodcoord)
    lo, hi = nod.min(axis=0), nod.max(axis=0)
    tol = 1e-12 * max(1.0, float(np.abs(nod).max()))
    on = np.zeros(nod.shape[0], bool)
    for d in range(3):
        on |= np.abs(nod[:, d] - lo[d]) < tol
        on |= np.abs(nod[:, d] - hi[d]) < tol
    if exclude_z_min:
        on &= ~(np.abs(nod[:, 2] - lo[2]) < tol)
    return np.nonzero(on)[0]
