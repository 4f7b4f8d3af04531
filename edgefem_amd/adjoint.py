"""Differentiable S-parameters: the eigenmode S matrix over a frequency list as a
``torch.autograd.Function`` of the real relative permittivity of a set of design tets.

Forward: one batched solve of every (frequency, excitation) system
(``pyedgefem.frequency_sweep_eigenmode``, the formulation of
src/assemble_maxwell.cpp:637-787) with the solutions kept on the GPU.

Backward: the adjoint.  A is complex-symmetric, so the adjoint system A^T lam_j = conj(g_j)
is the forward system itself.  With g_j = M_s,j e_j real (real port mode), lam_j is the
solved excitation x_j scaled: b_j = 2 c_j g_j with c_j = j beta_j scale, so
lam_j = x_j / (2 c_j).  No extra solve is needed.  Then
    dS_ja / d eps_e = k0^2 lam_j^T M_e x_a
and for a real loss with G = dL/d conj(S) (PyTorch's convention)
    dL/d eps_e = sum_{f,j,a} Re(conj(G_fja) k0_f^2 / (2 c_fj) x_fj^T M_e x_fa),
which is evaluated per design tet by ``efm_eps_gradient``.  That kernel recomputes M_e
from the coordinates rather than storing element blocks.

For a complex port mode (g_j complex) the adjoint right-hand side conj(g_j) differs from the
excitation's g_j, so lam_j = A^-1 conj(g_j) is solved for: a second batched solve of the same
systems with those right-hand sides, kept beside the forward solutions; then
    dL/d eps_e = sum_{f,j,a} Re(conj(G_fja) k0_f^2 lam_fj^T M_e x_fa).

beta_i depends on the permittivity of the first tet with a face on port i, and it is held
fixed here, so such tets may not be design tets (checked).
"""

from __future__ import annotations

import copy
from typing import Optional, Sequence

import numpy as np

from . import _capi
from . import pyedgefem as em


class EpsDesignProblem:
    """Fixed mesh, boundary conditions, ports and frequencies.  The variables are the real
    relative permittivities of ``design`` (tet indices).  The imaginary part of each design
    tet's permittivity (its loss) is taken from ``params`` and held fixed."""

    def __init__(self, mesh: em.Mesh, params: em.MaxwellParams, bc: em.BC, ports: Sequence[em.WavePort],
                 frequencies, design, solver_options: Optional[em.SolveOptions] = None,
                 max_batch: int = 256):
        self.mesh = mesh
        self.params = params
        self.bc = bc
        self.ports = list(ports)
        self.freqs = np.asarray(frequencies, np.float64).ravel()
        self.design = np.asarray(design, np.int64).ravel()
        if self.design.size and (self.design.min() < 0 or self.design.max() >= mesh.tet_nodes.shape[0]):
            raise ValueError("design tet index out of range")
        if np.unique(self.design).size != self.design.size:
            raise ValueError("design tets must be distinct")
        for port in self.ports:
            t = em._port_face_tet(mesh, port.surface_tag)
            if t >= 0 and np.any(self.design == t):
                raise ValueError(f"tet {t} sets beta of port {port.surface_tag}; it cannot be a design tet")
        # complex port modes: separate adjoint solves (A lam = conj(g))
        self.complex_modes = False
        for port in self.ports:
            w = np.asarray(port.weights)
            if np.iscomplexobj(w) and np.abs(w.imag).max(initial=0.0) > 0:
                self.complex_modes = True
        self.solver_options = solver_options
        P = max(1, len(self.ports))
        self.max_batch = max(P, (int(max_batch) // P) * P)   # a frequency's excitations share a batch
        params._check_supported()  # dA/d eps = -k0^2 M_e needs unstretched, non-dispersive media
        self.base_eps, _ = em.tet_materials(mesh, params)

    def permittivity(self, eps_design: np.ndarray) -> np.ndarray:
        eps = self.base_eps.copy()
        eps[self.design] = np.asarray(eps_design, np.float64) + 1j * self.base_eps[self.design].imag
        return eps

    def solve(self, eps_design):
        """S (F, P, P) and the kept batched solutions."""
        p = copy.copy(self.params)
        p.eps_r_tet = self.permittivity(eps_design)
        st = {"keep_solutions": True, "adjoint_conj": self.complex_modes}
        S = em.frequency_sweep_eigenmode(self.mesh, p, self.bc, self.ports, self.freqs,
                                         solver_options=self.solver_options, max_batch=self.max_batch, _state=st)
        self.last_infos = st.get("infos", [])
        return S, st.get("solutions", [])

    def gradient(self, solutions, G: np.ndarray):
        """dL/d eps_design (device float64 tensor) from G = dL/d conj(S), (F, P, P)."""
        import torch
        G = np.asarray(G, np.complex128)
        topo = self.mesh.topology
        grad = None
        for sol in solutions:
            fi, ex = sol["freq_index"], sol["excitation"]
            cols = {(f, a): c for c, (f, a) in enumerate(zip(fi, ex))}
            tl, tr, w = [], [], []
            for (f, a), ca in cols.items():
                k0sq = sol["k0sq"][ca]
                for (f2, j), cj in cols.items():
                    if f2 != f or G[f, j, a] == 0:
                        continue
                    tl.append(cj)
                    tr.append(ca)
                    if sol.get("L") is not None:      # lam_j solved for (complex mode)
                        w.append(np.conj(G[f, j, a]) * k0sq)
                    else:                              # lam_j = x_j / (2 c_j) (real mode)
                        w.append(np.conj(G[f, j, a]) * k0sq / (2.0 * complex(sol["coef"][cj][j])))
            if not w:
                continue
            XL = sol["L"] if sol.get("L") is not None else sol["X"]
            g = _capi.eps_gradient(topo, self.mesh.dev_xyz, self.design, XL, sol["X"], tl, tr, w)
            grad = g if grad is None else grad + g
        if grad is None:
            grad = torch.zeros(self.design.size, dtype=torch.float64, device=topo.row_ptr.device)
        return grad


class SParamsOfPermittivity:
    """``S = SParamsOfPermittivity.apply(eps_design, problem)``: S (F, P, P) complex128 on the
    device of ``eps_design``, differentiable with respect to ``eps_design`` (real, shape
    (D,)).  NaN entries (a frequency below a port cutoff, or an unconverged system) have
    no gradient path, so the loss should not use them."""

    _fn = None

    @staticmethod
    def apply(eps_design, problem: EpsDesignProblem):
        return _function().apply(eps_design, problem)


def _function():
    if SParamsOfPermittivity._fn is None:
        import torch

        class _F(torch.autograd.Function):
            @staticmethod
            def forward(ctx, eps_design, problem):
                e = eps_design.detach().cpu().numpy().astype(np.float64)
                S, sols = problem.solve(e)
                ctx.problem = problem
                ctx.solutions = sols
                ctx.device = eps_design.device
                return torch.from_numpy(S).to(eps_design.device)

            @staticmethod
            def backward(ctx, G):
                grad = ctx.problem.gradient(ctx.solutions, G.detach().cpu().numpy())
                return grad.to(ctx.device), None

        SParamsOfPermittivity._fn = _F
    return SParamsOfPermittivity._fn
