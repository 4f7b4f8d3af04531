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
                 max_batch: int = 256, forward: str = "auto"):
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
        # forward: "batched" (one batched sweep of every (frequency, excitation) system),
        # "per_frequency" (the single-system solver, the excitations of a frequency as one
        # block), or "auto": per frequency when the systems fill less than one 64-lane group
        # of the batched kernels (on the 798,504-edge design guide, 16 frequencies x 2
        # excitations: 12.2 s against 14.9 s batched), and only for real port modes
        if forward not in ("auto", "batched", "per_frequency"):
            raise ValueError(f"unknown forward {forward!r}")
        self.forward = forward
        params._check_supported()  # dA/d eps = -k0^2 M_e needs unstretched, non-dispersive media
        self.base_eps, _ = em.tet_materials(mesh, params)

    def permittivity(self, eps_design: np.ndarray) -> np.ndarray:
        eps = self.base_eps.copy()
        eps[self.design] = np.asarray(eps_design, np.float64) + 1j * self.base_eps[self.design].imag
        return eps

    def _per_frequency(self) -> bool:
        if self.forward == "per_frequency":
            if self.complex_modes:
                raise NotImplementedError("complex port modes need the batched forward (its conjugate solves)")
            return True
        return self.forward == "auto" and not self.complex_modes and self.freqs.size * len(self.ports) < 64

    def solve(self, eps_design):
        """S (F, P, P) and the kept solutions."""
        p = copy.copy(self.params)
        p.eps_r_tet = self.permittivity(eps_design)
        if self._per_frequency():
            P = len(self.ports)
            S = np.full((self.freqs.size, P, P), np.nan + 1j * np.nan, np.complex128)
            sols, infos = [], []
            for i, f in enumerate(self.freqs):
                q = em._params_at(p, 2.0 * np.pi * float(f))
                st = {}
                S[i] = em.calculate_sparams_eigenmode(self.mesh, q, self.bc, self.ports,
                                                      solver_options=self.solver_options, _state=st)
                if "X" not in st:      # below a port cutoff
                    continue
                infos.append(st["info"])
                k0sq = (2.0 * np.pi * float(f) / em.C0) ** 2
                coefs = [1j * q.port_abc_scale * b for b in st["betas"]]
                sols.append(dict(X=st["X"], L=None, freq_index=[i] * P, excitation=list(range(P)),
                                 k0sq=[k0sq] * P, coef=[coefs] * P))
            self.last_infos = infos
            return S, sols
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


class PeriodicEpsDesignProblem:
    """A periodic unit cell (calculate_sparams_periodic, assemble_maxwell.cpp:576-635) as a
    function of the real relative permittivity of ``design`` tets, at the frequency of
    ``params`` and the Floquet phase set on ``pbc``.  Same contract as EpsDesignProblem
    (``solve`` and ``gradient``), so ``SParamsOfPermittivity.apply(eps, problem)`` returns a
    differentiable (P, P) S matrix.

    Forward: the reduced system A' = Z E A F Z + I_S (E adds the slave rows to the master
    rows, F the slave columns to the master columns, Z drops the slaves) with the ports'
    rank-1 loading c w w^H, solved by BiCGSTAB; slaves recovered as x = R x'; S_ji =
    (w_j^H x_i - delta_ji sqrt(Z0_i)) / sqrt(Z0_i).

    Backward: A' is not symmetric (a complex Floquet phase, and the Hermitian rank-1
    loading), so the adjoint needs its transpose.  With lam_j = A'^-T R^T conj(w_j) and
    dA/d eps_t = -k0^2 M_t,
        dS_ji / d eps_t = k0^2 / sqrt(Z0_i) (E^T Z lam_j)^T M_t (F Z x'_i),
    one transposed solve per output port.  A'^T is installed on the same pattern (the
    elimination keeps it structurally symmetric; checked) with the values transposed, the
    port terms' vectors conjugated (c w w^H)^T = c conj(w) conj(w)^H, and the auxiliary
    preconditioner's node phase conjugated (the rows' phase is the columns' of A'^T).
    beta of a port depends on the permittivity of the tets on its face, so those tets may
    not be design tets (checked)."""

    def __init__(self, mesh: em.Mesh, params: em.MaxwellParams, bc: em.BC, pbc, ports: Sequence[em.WavePort],
                 design, mode: str = "reference", solver_options: Optional[em.SolveOptions] = None):
        self.mesh = mesh
        self.params = params
        self.bc = bc
        self.pbc = pbc
        self.ports = list(ports)
        self.mode = mode
        self.design = np.asarray(design, np.int64).ravel()
        if self.design.size and (self.design.min() < 0 or self.design.max() >= mesh.tet_nodes.shape[0]):
            raise ValueError("design tet index out of range")
        if np.unique(self.design).size != self.design.size:
            raise ValueError("design tets must be distinct")
        for port in self.ports:
            t = em._port_face_tet(mesh, port.surface_tag)
            if t >= 0 and np.any(self.design == t):
                raise ValueError(f"tet {t} sets beta of port {port.surface_tag}; it cannot be a design tet")
        self.solver_options = solver_options
        params._check_supported()
        self.base_eps, _ = em.tet_materials(mesh, params)
        self._tperm = None

    def permittivity(self, eps_design: np.ndarray) -> np.ndarray:
        eps = self.base_eps.copy()
        eps[self.design] = np.asarray(eps_design, np.float64) + 1j * self.base_eps[self.design].imag
        return eps

    def solve(self, eps_design):
        """S (P, P) and the kept state of the forward solve."""
        p = copy.copy(self.params)
        p.eps_r_tet = self.permittivity(eps_design)
        st = {}
        S = em.calculate_sparams_periodic(self.mesh, p, self.bc, self.pbc, self.ports, mode=self.mode,
                                          solver_options=self.solver_options, _state=st)
        self.last_info = st.get("info")
        return S, st

    def _transpose_perm(self, A):
        """Position of entry (j, i) for every entry (i, j) of the CSR pattern of A."""
        import torch
        if self._tperm is not None and self._tperm[0] == A.col_idx.numel():
            return self._tperm[1]
        n = A.shape[0]
        rp = A.row_ptr.to(torch.int64)
        ci = A.col_idx.to(torch.int64)
        rows = torch.repeat_interleave(torch.arange(n, device=ci.device), rp[1:] - rp[:-1])
        key = rows * n + ci
        keyT = ci * n + rows
        pos = torch.searchsorted(key, keyT)
        ok = bool((pos < key.numel()).all()) and bool((key[pos.clamp(max=key.numel() - 1)] == keyT).all())
        if not ok:
            raise RuntimeError("the reduced periodic matrix is not structurally symmetric")
        self._tperm = (A.col_idx.numel(), pos)
        return pos

    def gradient(self, st, G: np.ndarray):
        """dL/d eps_design (device float64 tensor) from G = dL/d conj(S), (P, P)."""
        import torch
        G = np.asarray(G, np.complex128).reshape(len(self.ports), len(self.ports))
        A = st["A"]
        n = self.mesh.num_edges()
        P = len(self.ports)
        m, s_, a, c, rec, act = st["pairs"]
        # right-hand sides R^T conj(w_j): the port rows, slave entries moved onto their masters
        rhs = np.zeros((n, P), np.complex128)
        for j, (e, w) in enumerate(st["frees"]):
            np.add.at(rhs[:, j], e, np.conj(w))
        if m.size:
            np.add.at(rhs, m, rec[:, None] * rhs[s_])
            rhs[s_] = 0
        R = st["nrhs"]
        solver = _capi.KrylovSolver(A.row_ptr, A.col_idx, nrhs=R)
        N, en, ec = st["gradient"]
        solver.set_gradient_complex(N, en, np.conj(ec))
        if A.lowrank:
            gi, gv, cf = em._pack_lowrank(A.lowrank)
            solver.set_lowrank(gi, torch.conj_physical(gv), cf)   # a conj view would pass the raw values
        solver.set_matrix(A.values[self._transpose_perm(A)])
        Lam, info = em._block_solve(solver, n, P, em._dense_rhs_block(rhs, R), st["tol"], st["max_iterations"],
                                    _capi.EFM_METHOD_BICGSTAB)
        self.last_adjoint_info = info
        self.last_lam = Lam
        # expansions: E^T Z lam (slave = a lam_master) and F Z x' (slave = c x'_master)
        U = Lam.clone()
        V = st["X_reduced"][:, :P].clone()
        dev = U.device
        if act.any():
            mi = torch.from_numpy(m[act]).to(dev)
            si = torch.from_numpy(s_[act]).to(dev)
            U[si] = torch.from_numpy(a[act]).to(dev)[:, None] * U[mi]
            V[si] = torch.from_numpy(c[act]).to(dev)[:, None] * V[mi]
        # Dirichlet rows and columns of A' do not depend on eps: a Dirichlet master can pick
        # up a slave's port weight in R^T conj(w), which its identity row passes to lam
        dz = torch.from_numpy(np.nonzero(np.asarray(st["dmask"]) != 0)[0]).to(dev)
        U[dz] = 0
        V[dz] = 0
        k0 = float(self.params.omega) / em.C0
        tl, tr, wts = [], [], []
        for i in range(P):
            for j in range(P):
                if G[j, i] != 0:
                    tl.append(j)
                    tr.append(i)
                    wts.append(np.conj(G[j, i]) * k0 * k0 / st["vinc"][i])
        if not wts:
            return torch.zeros(self.design.size, dtype=torch.float64, device=U.device)
        return _capi.eps_gradient(self.mesh.topology, self.mesh.dev_xyz, self.design, U.contiguous(),
                                  V.contiguous(), tl, tr, wts)
