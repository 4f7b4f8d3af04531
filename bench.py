"""Benchmark: eigenmode S-parameter solves of the WR-90 guide, 48x24x192 hexes
(1,327,104 tets, 1,593,480 edges), 100 mm long, lossy dielectric slab z in [30, 70] mm
(per-tet eps' ~ U[1,4], tan d = 0.002), two TE10 wave ports, 10 GHz.

One step = one pass of the hot path at one frequency: K/M element matrices and
deterministic assembly, the two port surface masses, A = K - k0^2 M + sum j beta M_s,
COCG with the auxiliary-gradient preconditioner for both excitations (one block of two
right-hand sides), and the S extraction.  Mesh topology and the port modes (frequency independent setup) are built
before the timed region; every input is resident in HBM when timing starts.

Prints one JSON line (see the `result` dict at the end).
"""

import argparse
import json
import math
import os
import sys
import time

ROOT = os.path.dirname(os.path.abspath(__file__))
sys.path.insert(0, ROOT)

HBM_PEAK = 8.0e12
METRIC = "S-param solves/sec at |ΔS|≤1e-6 vs FP64 LU; grads/sec; SpMV HBM GB/s"


def log(msg):
    print(msg, file=sys.stderr, flush=True)


def build_workload(nx=48, ny=24, nz=192, length=100e-3, freq=10e9):
    import numpy as np
    from edgefem_amd import meshgen as mg
    from edgefem_amd import pyedgefem as em
    g = mg.box_mesh(mg.WR90_A, mg.WR90_B, length, nx, ny, nz, seed=1)
    eps = mg.slab_eps(g, z0=30e-3, z1=70e-3, lo=1.0, hi=4.0, tan_d=0.002, seed=2)
    mesh = em.mesh_from_arrays(g.node_ids, g.xyz, g.tets, g.tet_phys, g.tris, g.tri_phys)
    bc = em.build_edge_pec(mesh, 1)
    mode = em.solve_te10_mode(em.RectWaveguidePort(mg.WR90_A, mg.WR90_B), freq)
    kc2 = (math.pi / mg.WR90_A) ** 2
    ports = [em.build_wave_port_2d(mesh, t, mode, bc.dirichlet_edges, kc2) for t in (2, 3)]
    params = em.MaxwellParams()
    params.omega = 2 * math.pi * freq
    params.eps_r_tet = eps
    return g, eps, mesh, bc, ports, params


class Pipeline:
    """Device-resident inputs and the per-step hot path."""

    def __init__(self, mesh, bc, ports, params, matrix_dtype, tol, max_it, prec="aux"):
        import numpy as np
        import torch
        from edgefem_amd import _capi
        from edgefem_amd import pyedgefem as em
        self.mesh, self.ports, self.params = mesh, ports, params
        self.tol, self.max_it = tol, max_it
        topo = mesh.topology
        eps, mu = em.tet_materials(mesh, params)
        self.eps = _capi.to_dev(eps, torch.complex128)
        self.inv_mu = _capi.to_dev(1.0 / mu, torch.complex128)
        self.dmask = _capi.to_dev(bc.mask(mesh.num_edges()), torch.uint8)
        self.xyz = mesh.dev_xyz
        self.sel = [_capi.to_dev(np.nonzero(mesh.tri_phys == p.surface_tag)[0].astype(np.int32), torch.int32)
                    for p in ports]
        self.ops = em._PortOps(mesh, bc, ports)
        self.betas = em._eigenmode_betas(mesh, params, ports, eps, mu)
        self.B = self.ops.rhs(self.betas, params.port_abc_scale, list(range(len(ports))))
        self.k0 = params.omega / em.C0
        self.dtype = matrix_dtype
        self.solver = _capi.KrylovSolver(topo.row_ptr, topo.col_idx, nrhs=len(ports), dtype=matrix_dtype)
        if prec == "aux":
            self.solver.set_gradient(topo, self.dmask)
        self.X = torch.empty_like(self.B)

    def step(self):
        import numpy as np
        from edgefem_amd import _capi
        topo = self.mesh.topology
        K, M = _capi.assemble_km(topo, self.xyz, self.eps, self.inv_mu, self.dmask)
        Ms = [_capi.port_surface_mass(topo, self.xyz, s, self.dmask) for s in self.sel]
        coefs = [1j * self.params.port_abc_scale * b for b in self.betas]
        A = _capi.combine(topo, K, M, self.k0 * self.k0, Ms, coefs, out_dtype=self.dtype)
        self.solver.set_matrix(A)
        X, info = self.solver.solve(self.B, tol=self.tol, max_iterations=self.max_it, out=self.X)
        ov = _capi.port_overlap(self.ops.gi_dev, self.ops.gv_dev, X, len(self.ports))
        self.last_info = info
        self.last_A = A
        return ov - np.eye(len(self.ports))


def golden_c2():
    """The committed FP64 S of this workload (tests/golden/c2_sparams_10ghz.npz: the oracle's
    CPU assembly and an FP64 COCG solve to a relative residual of 1e-13 plus refinement,
    since a SuperLU factor of this system does not fit in memory), or None."""
    import numpy as np
    path = os.path.join(ROOT, "tests", "golden", "c2_sparams_10ghz.npz")
    if not os.path.exists(path):
        return None
    z = np.load(path)
    return dict(S=z["S"], method=str(z["method"]), residual=[float(v) for v in z["residual"]])


def reference_path_sparams(om, p, pec, oports, time_limit=0.0):
    """The reference's CPU path for eigenmode S (assemble_maxwell.cpp:637-787 with
    solve_linear's defaults, solver.cpp:53-80): per active port the system is assembled
    again (the oracle's triplet assembly plus the j beta M_s port terms), factored by ILUT
    (fill 10, drop 1e-4) and solved by BiCGSTAB to 1e-10, all on one host core (the C
    restatement in oracle/ilut_bicgstab.c).  Returns S, per-excitation info and seconds."""
    import numpy as np
    from oracle import edgefem_oracle as O
    from oracle import ilut
    P = len(oports)
    t0 = time.perf_counter()
    betas, Ms, vecs = O.eigenmode_setup(om, p, pec, oports)
    X, infos = [], []
    for a in range(P):
        A, _ = O.assemble_maxwell(om, p, pec)
        for i in range(P):
            A = A + (1j * p.port_abc_scale * betas[i]) * Ms[i]
        A = A.tocsr()
        b = 2.0 * (1j * p.port_abc_scale * betas[a]) * (Ms[a] @ vecs[a])
        x, info = ilut.solve_linear(A, b, time_limit=time_limit)
        X.append(x)
        infos.append(info)
    S = np.array([[np.vdot(vecs[j], Ms[j] @ X[a]) - (1.0 if j == a else 0.0) for a in range(P)] for j in range(P)])
    return S, infos, time.perf_counter() - t0


def reference_solver_c2_sample(g, eps, freq, budget_s=15.0):
    """A bounded sample of the reference's solve on this workload: the ILUT factorization
    (fill 10, drop 1e-4, RCM order) of the full C2 system for at most budget_s seconds on one
    host core, giving its row rate and the projected factorization time.  BiCGSTAB itself is
    not run here: with this ILUT it does not converge on the lossy-slab systems (measured in
    the build container: on C2 the factored preconditioner gives non-finite iterates at the
    first iteration; on a 16x8x64 slab guide BiCGSTAB diverges (RCM) or stalls at a relative
    residual of 1.0 (natural order) over 2,000+ iterations), and the reference then falls back
    to SparseLU, whose factor of this system does not fit in memory."""
    import numpy as np
    from oracle import edgefem_oracle as O
    from oracle import ilut
    om = O.mesh_from_arrays(g.node_ids, g.xyz, g.tet_ids, g.tet_phys, g.tri_ids, g.tri_phys)
    pec = O.build_edge_pec(om, 1)
    a_ = 22.86e-3
    kc, beta, Z0 = O.solve_te10_mode(a_, freq)
    oports = [O.build_wave_port_2d(om, t, kc, Z0, beta, pec, (math.pi / a_) ** 2) for t in (2, 3)]
    p = O.Params(omega=2 * math.pi * freq, eps_tet=eps)
    betas, Ms, vecs = O.eigenmode_setup(om, p, pec, oports)
    t0 = time.perf_counter()
    A, _ = O.assemble_maxwell(om, p, pec)
    for i in range(2):
        A = A + (1j * betas[i]) * Ms[i]
    A = A.tocsr()
    t_asm = time.perf_counter() - t0
    try:
        M = ilut.ILUT(A, 1e-4, 10, time_limit=budget_s)
        rows, secs = A.shape[0], M.seconds
        M.close()
    except ilut.ILUTTimeLimit as e:
        rows, secs = e.rows_done, e.seconds
    rate = rows / max(secs, 1e-9)
    return dict(status="did not finish", assembly_s=t_asm, ilut_rows_done=int(rows), ilut_rows=int(A.shape[0]),
                ilut_rows_per_s=rate, ilut_projected_s=A.shape[0] / rate, cores=1,
                note="BiCGSTAB + ILUT(fill 10, drop 1e-4), the reference's default solve, restated in C "
                     "(oracle/ilut_bicgstab.c); timed here: assembly and a bounded ILUT sample; BiCGSTAB with "
                     "this preconditioner does not converge on the lossy-slab systems, and the SparseLU "
                     "fallback does not fit in memory")


def cpu_baseline(g, eps, freq, gpu_iters, prec, budget_s=20.0):
    """The oracle (numpy/scipy) on one host core: K/M assembly of the same mesh, then
    iterations of the same preconditioned COCG on two right-hand sides for about
    `budget_s` seconds, extrapolated to the GPU run's iteration count (same algorithm,
    so the same count)."""
    import numpy as np
    import scipy.sparse as sp
    from oracle import edgefem_oracle as O
    t_setup = time.time()
    om = O.mesh_from_arrays(g.node_ids, g.xyz, g.tet_ids, g.tet_phys, g.tri_ids, g.tri_phys)
    pec = O.build_edge_pec(om, 1)
    log(f"  cpu baseline: topology {time.time() - t_setup:.1f}s")
    p = O.Params(omega=2 * math.pi * freq, eps_tet=eps)
    t0 = time.time()
    K, M = O.assemble_maxwell_km(om, p, pec)
    k0 = p.omega / O.C0
    A = (K - (k0 * k0) * M).tocsr()
    t_asm = time.time() - t0
    n = A.shape[0]
    dinv = 1.0 / A.diagonal()
    if prec == "aux":
        pos = {int(v): k for k, v in enumerate(om.node_ids)}
        a_ = np.array([pos[int(v)] for v in om.edges[:, 0]])
        b_ = np.array([pos[int(v)] for v in om.edges[:, 1]])
        G = sp.csr_matrix((np.tile([-1.0, 1.0], n), (np.repeat(np.arange(n), 2), np.stack([a_, b_], 1).ravel())),
                          shape=(n, om.xyz.shape[0]))
        G = (sp.diags((~pec).astype(float)) @ G).tocsr()
        GT = G.T.tocsr()
        dn = (GT @ A @ G).diagonal()
        dninv = np.where(dn != 0, 1.0 / np.where(dn != 0, dn, 1.0), 0.0)

        def P(r):
            return dinv[:, None] * r + G @ (dninv[:, None] * (GT @ r))
    else:
        def P(r):
            return dinv[:, None] * r
    rng = np.random.default_rng(0)
    B = rng.standard_normal((n, 2)) + 1j * rng.standard_normal((n, 2))
    x = np.zeros_like(B)
    r = B.copy()
    z = P(r)
    pp = z.copy()
    rho = np.einsum("ij,ij->j", r, z)
    it = 0
    t1 = time.time()
    while time.time() - t1 < budget_s and it < 100000:
        q = A @ pp
        al = rho / np.einsum("ij,ij->j", pp, q)
        x += al * pp
        r -= al * q
        z = P(r)
        rn = np.einsum("ij,ij->j", r, z)
        pp = z + (rn / rho) * pp
        rho = rn
        it += 1
    t_it = (time.time() - t1) / max(it, 1)
    t_solve = t_asm + gpu_iters * t_it
    return dict(value=2.0 / t_solve, unit="S-solves/s", cores=1, kind="port",
                sample=(f"oracle numpy/scipy on 1 host core: K/M assembly of the full mesh ({t_asm:.1f} s) "
                        f"+ {it} COCG iterations ({prec} preconditioner, 2 RHS, {t_it * 1e3:.1f} ms/iter), "
                        f"extrapolated to the GPU's {gpu_iters} iterations"),
                seconds_per_solve_pair=t_solve)


def sweep_main(args):
    print(json.dumps(run_sweep(args, args.steps, args.warmup, args.sweep_precision)))


def run_sweep(args, steps, warmup, prec, cpu=True):
    """128 frequencies (8.2-12.4 GHz) x 2 excitations = 256 systems on the 16x8x32 WR-90
    guide, one batched launch sequence; compared with the plain path (one frequency at a
    time, both excitations as a 2-RHS block) on a subset of the frequencies."""
    import numpy as np
    import torch
    from edgefem_amd import meshgen as mg
    from edgefem_amd import pyedgefem as em
    torch.cuda.set_device(0)
    g, eps, mesh, bc, ports, params = build_workload(16, 8, 32, 50e-3, 10e9)
    params.eps_r_tet = None
    freqs = np.linspace(8.2e9, 12.4e9, 128)
    gold_path = os.path.join(ROOT, "tests", "golden", "c3_sweep.npz")
    gold = np.load(gold_path) if os.path.exists(gold_path) else None
    for _ in range(warmup):
        em.frequency_sweep_eigenmode(mesh, params, bc, ports, freqs, precision=prec,
                                     _state={"spmv_variant": args.sweep_spmv, "aux_variant": min(args.sweep_spmv, 1),
                                             "iteration": args.sweep_iteration, "graph": args.sweep_graph})
    torch.cuda.synchronize()
    st = {"spmv_variant": args.sweep_spmv, "aux_variant": min(args.sweep_spmv, 1), "iteration": args.sweep_iteration,
          "graph": args.sweep_graph}
    ts = time.perf_counter()
    for _ in range(steps):
        S = em.frequency_sweep_eigenmode(mesh, params, bc, ports, freqs, precision=prec, _state=st)
    torch.cuda.synchronize()
    el = (time.perf_counter() - ts) / steps
    nsol = 2 * freqs.size
    value = nsol / el
    dS = float(np.abs(S - gold["S"]).max()) if gold is not None else None
    # plain path on every 16th frequency
    sub = freqs[::16]
    torch.cuda.synchronize()
    t0 = time.perf_counter()
    Sp = []
    for f in sub:
        p = em.MaxwellParams()
        p.omega = 2 * math.pi * f
        Sp.append(em.calculate_sparams_eigenmode(mesh, p, bc, ports))
    torch.cuda.synchronize()
    plain = 2 * sub.size / (time.perf_counter() - t0)
    dS_plain = float(np.abs(np.array(Sp) - gold["S"][::16]).max()) if gold is not None else None
    its = np.concatenate([i["iters"] for i in st["infos"]])
    res = np.concatenate([i["resid"] for i in st["infos"]])
    want_cpu = cpu
    cpu = None
    if want_cpu and not args.no_cpu:
        # the reference's CPU path (re-assembly, ILUT and BiCGSTAB per excitation, one core)
        # on two of the frequencies, the band edges
        from oracle import edgefem_oracle as O
        om = O.mesh_from_arrays(g.node_ids, g.xyz, g.tet_ids, g.tet_phys, g.tri_ids, g.tri_phys)
        pec = O.build_edge_pec(om, 1)
        a_ = mg.WR90_A
        kc, beta, Z0 = O.solve_te10_mode(a_, 10e9)
        oports = [O.build_wave_port_2d(om, t, kc, Z0, beta, pec, (math.pi / a_) ** 2) for t in (2, 3)]
        t_ref, d_ref, it_ref = 0.0, 0.0, []
        for fi in (0, freqs.size - 1):
            Sr, infos, dt = reference_path_sparams(om, O.Params(omega=2 * math.pi * freqs[fi]), pec, oports)
            t_ref += dt
            it_ref += [i["iters"] for i in infos]
            if gold is not None:
                d_ref = max(d_ref, float(np.abs(Sr - gold["S"][fi]).max()))
        cpu = dict(value=4 / t_ref, unit="S-solves/s", cores=1, kind="port",
                   sample=f"the reference's CPU path restated (per excitation: triplet assembly + port terms, "
                          f"ILUT fill 10 drop 1e-4, BiCGSTAB to 1e-10; oracle/ilut_bicgstab.c) at "
                          f"{freqs[0] / 1e9:.1f} and {freqs[-1] / 1e9:.1f} GHz, 4 S-solves in {t_ref:.1f} s, "
                          f"BiCGSTAB iterations {it_ref}",
                   delta_S_vs_fp64_lu=d_ref if gold is not None else None)
    result = {
        "metric": METRIC, "value": value, "unit": "S-solves/s", "n_gpus": 1, "steps": steps,
        "warmup": warmup, "ms_per_step": 1e3 * el, "higher_is_better": True, "scaling": "weak",
        "vs_baseline": None, "dtype": "c64 vectors + c128 reliable updates" if prec == "fp32" else "c128",
        "data": "synthetic (seeded structured Kuhn mesh)",
        "config": {"workload": "wr90_16x8x32_sweep_128f_x2", "systems": int(nsol), "edges": mesh.num_edges(),
                   "freqs_hz": [float(freqs[0]), float(freqs[-1]), int(freqs.size)],
                   "solver": "batched COCG+Jacobi+nodal-gradient", "tol": 1e-10,
                   "iteration": ("standard", "chronopoulos-gear", "standard-u-rebuilt")[args.sweep_iteration],
                   "precision": prec, "graph_replay": bool(args.sweep_graph)},
        "plain_path_solves_per_s": plain, "speedup_vs_plain": value / plain,
        "delta_S_vs_fp64_lu": dS, "delta_S_plain_vs_fp64_lu": dS_plain,
        "iterations": [int(its.min()), int(its.max())], "max_true_rel_residual": float(res.max()),
        "kernel_variants": {"spmv": int(st["spmv_chosen"]), "aux": int(st["aux_chosen"])},
        "tune_ms": {"spmv": [round(float(t), 4) for t in st["tune_ms"][:3]],
                    "aux": [round(float(t), 4) for t in st["tune_ms"][3:]]},
        "cpu_baseline": cpu,
    }
    return result


def media_main(args):
    print(json.dumps(run_media(args, args.steps, args.warmup)))


def run_media(args, steps, warmup):
    """The sweep where K/M reuse cannot serve: the 16x8x32 WR-90 guide with a two-pole
    Lorentz slab (z 10-20 mm) and a scalar-PML slab (z 30-36 mm, sigma = 0.3 w0), 128
    frequencies (8.2-12.4 GHz) x 2 excitations = 256 systems in one batched solve (K, M at
    the median frequency, a per-system correction on the two slabs' edges), against the
    per-frequency assembly and solve of the reference's sweep (sweep.cpp:330-346) on every
    16th frequency."""
    import numpy as np
    import torch
    from edgefem_amd import materials as MT
    from edgefem_amd import meshgen as mg
    from edgefem_amd import pyedgefem as em
    torch.cuda.set_device(0)
    L = 50e-3
    g = mg.box_mesh(mg.WR90_A, mg.WR90_B, L, 16, 8, 32, seed=1)
    zc = g.centroids()[:, 2]
    phys = g.tet_phys.copy()
    phys[(zc > 10e-3) & (zc < 20e-3)] = 121
    phys[(zc > 30e-3) & (zc < 36e-3)] = 122
    mesh = em.mesh_from_arrays(g.node_ids, g.xyz, g.tets, phys, g.tris, g.tri_phys)
    bc = em.build_edge_pec(mesh, 1)
    mode = em.solve_te10_mode(em.RectWaveguidePort(mg.WR90_A, mg.WR90_B), 10e9)
    kc2 = (math.pi / mg.WR90_A) ** 2
    ports = [em.build_wave_port_2d(mesh, t, mode, bc.dirichlet_edges, kc2) for t in (2, 3)]
    w0 = 2 * math.pi * 10e9
    lor = MT.LorentzMaterial(2.0)
    lor.add_pole(1.5, 1.3 * w0, 0.1 * w0)
    lor.add_pole(0.5, 0.6 * w0, 0.05 * w0)
    params = em.MaxwellParams()
    params.set_eps_model(121, lor)
    params.pml_sigma, params.pml_regions = 0.3 * w0, {122}
    freqs = np.linspace(8.2e9, 12.4e9, 128)
    st = {"graph": bool(args.sweep_graph)}
    for _ in range(warmup):
        em.frequency_sweep_eigenmode(mesh, params, bc, ports, freqs, _state=dict(st))
    torch.cuda.synchronize()
    ts = time.perf_counter()
    for _ in range(steps):
        st = {"graph": bool(args.sweep_graph)}
        S = em.frequency_sweep_eigenmode(mesh, params, bc, ports, freqs, _state=st)
    torch.cuda.synchronize()
    el = (time.perf_counter() - ts) / steps
    sub = freqs[::16]
    torch.cuda.synchronize()
    t0 = time.perf_counter()
    Sp = em.frequency_sweep_eigenmode(mesh, params, bc, ports, sub, _state={"per_frequency": True})
    torch.cuda.synchronize()
    plain = 2 * sub.size / (time.perf_counter() - t0)
    its = np.concatenate([i["iters"] for i in st["infos"]])
    nsol = 2 * freqs.size
    return {
        "metric": METRIC, "value": nsol / el, "unit": "S-solves/s", "n_gpus": 1, "steps": steps,
        "warmup": warmup, "ms_per_step": 1e3 * el, "higher_is_better": True, "scaling": "weak",
        "vs_baseline": None, "dtype": "c128", "data": "synthetic (seeded structured Kuhn mesh)",
        "config": {"workload": "wr90_16x8x32_lorentz_pml_sweep_128f_x2", "systems": int(nsol),
                   "edges": mesh.num_edges(), "correction_entries": int(st["correction_entries"]),
                   "freqs_hz": [float(freqs[0]), float(freqs[-1]), int(freqs.size)], "tol": 1e-10},
        "per_frequency_solves_per_s": plain, "speedup_vs_per_frequency": nsol / el / plain,
        "delta_S_vs_per_frequency": float(np.abs(S[::16] - Sp).max()),
        "iterations": [int(its.min()), int(its.max())],
    }


def fields_main(args):
    """Post-solve field evaluation on the 48x24x192 WR-90 guide (1,327,104 tets, 1,593,480
    edges): E at the centroid and curl E of every tet for 2 solution columns (one step = one
    efm_tet_fields launch over all tets; the solution columns are seeded random values, the
    kernel's work does not depend on them).  Roofline on the algorithmic bytes: per tet the
    node and edge indices and signs (46 B) and the outputs (2 x 2 x 48 B), plus the
    coordinates and the solution gathered once (N x 24 B, n x 2 x 16 B)."""
    import numpy as np
    import torch
    from edgefem_amd import pyedgefem as em
    torch.cuda.set_device(0)
    g, eps, mesh, bc, ports, params = build_workload(*((12, 6, 48, 25e-3) if args.small else (48, 24, 192, 100e-3)))
    T, n, N, R = mesh.num_tets(), mesh.num_edges(), mesh.num_nodes(), 2
    rng = np.random.default_rng(3)
    x = torch.from_numpy(rng.normal(size=(n, R)) + 1j * rng.normal(size=(n, R))).to("cuda")
    for _ in range(max(args.warmup, 1)):
        em.tet_fields(mesh, x)
    torch.cuda.synchronize()
    e0, e1 = torch.cuda.Event(enable_timing=True), torch.cuda.Event(enable_timing=True)
    ts = time.perf_counter()
    e0.record()
    for _ in range(args.steps):
        E, C = em.tet_fields(mesh, x)
    e1.record()
    torch.cuda.synchronize()
    el = time.perf_counter() - ts
    launch_ms = e0.elapsed_time(e1) / args.steps
    bytes_per = T * 46 + T * R * 96 + N * 24 + n * R * 16
    achieved = bytes_per / (launch_ms * 1e-3) / 1e9
    # CPU baseline: the oracle's vectorized evaluation on a bounded sample of tets
    cpu = None
    if not args.no_cpu:
        from oracle import edgefem_oracle as O
        om = O.mesh_from_arrays(g.node_ids, g.xyz, g.tet_ids, g.tet_phys, g.tri_ids, g.tri_phys)
        xs = x.cpu().numpy()
        t0 = time.perf_counter()
        reps = 0
        while time.perf_counter() - t0 < 10.0:
            O.tet_fields(om, xs)
            reps += 1
        cpu_s = (time.perf_counter() - t0) / reps
        cpu = {"value": T / cpu_s, "unit": "tets/s", "cores": 1, "kind": "port",
               "sample": f"oracle numpy tet_fields over all {T} tets, 2 columns, {reps} run(s), 1 host core"}
        Eo, Co = O.tet_fields(om, xs)
        dE = float(np.abs(E.cpu().numpy() - Eo).max() / np.abs(Eo).max())
    else:
        dE = None
    result = {
        "metric": "tet field evaluations/s (E at the centroid and curl E, 2 columns)", "value": T / (launch_ms * 1e-3),
        "unit": "tets/s", "n_gpus": 1, "steps": args.steps, "warmup": args.warmup,
        "ms_per_step": 1e3 * el / args.steps, "higher_is_better": True, "scaling": "weak", "vs_baseline": None,
        "dtype": "c128", "data": "synthetic (seeded structured Kuhn mesh, seeded solution)",
        "config": {"workload": "wr90_48x24x192_tet_fields", "tets": T, "edges": n, "columns": R},
        "roofline": {"bound": "hbm", "achieved": achieved, "peak": 8000.0, "unit": "GB/s", "frac": achieved / 8000.0,
                     "traffic": None, "kernel": "k_tet_fields", "bytes_per_launch": bytes_per,
                     "avg_launch_ms": launch_ms},
        "max_rel_diff_vs_oracle": dE, "cpu_baseline": cpu,
    }
    print(json.dumps(result), flush=True)


def adjoint_main(args):
    print(json.dumps(run_adjoint(args, args.steps, args.warmup)))


def run_adjoint(args, steps, warmup):
    """Adjoint gradient of J = sum_f |S21(f)|^2 over 16 frequencies (9-11 GHz) with respect
    to the permittivity of every tet of a 77-cell design slab: WR-90 guide 48x24x96 hexes,
    50 mm (663,552 tets, 798,504 edges), design cells k = 10..86 (532,224 tets), eps_r ~
    U[1,6] seeded.  One step = forward (32 systems in one batched solve) + backward
    (gradient kernel over the design tets)."""
    import numpy as np
    import torch
    from edgefem_amd import adjoint as adj
    from edgefem_amd import meshgen as mg
    from edgefem_amd import pyedgefem as em
    torch.cuda.set_device(0)
    nx, ny, nz, L = (12, 6, 24, 50e-3) if args.small else (48, 24, 96, 50e-3)
    g = mg.box_mesh(mg.WR90_A, mg.WR90_B, L, nx, ny, nz, seed=1)
    mesh = em.mesh_from_arrays(g.node_ids, g.xyz, g.tets, g.tet_phys, g.tris, g.tri_phys)
    bc = em.build_edge_pec(mesh, 1)
    mode = em.solve_te10_mode(em.RectWaveguidePort(mg.WR90_A, mg.WR90_B), 10e9)
    kc2 = (math.pi / mg.WR90_A) ** 2
    ports = [em.build_wave_port_2d(mesh, t, mode, bc.dirichlet_edges, kc2) for t in (2, 3)]
    kcell = np.floor(g.centroids()[:, 2] / (L / nz)).astype(np.int64)
    k0, k1 = (10, 86) if nz == 96 else (round(10 * nz / 96), round(86 * nz / 96))
    design = np.nonzero((kcell >= k0) & (kcell <= k1))[0]
    eps0 = 1.0 + 5.0 * mg.uniform01(3, design.astype(np.int64))
    freqs = np.linspace(9e9, 11e9, 16)
    opts = em.SolveOptions()
    opts.tolerance = args.tol
    opts.max_iterations = args.max_it
    prob = adj.EpsDesignProblem(mesh, em.MaxwellParams(), bc, ports, freqs, design, solver_options=opts)
    e = torch.tensor(eps0, dtype=torch.float64, device="cuda", requires_grad=True)
    log(f"adjoint: {design.size} design tets, {mesh.num_edges()} edges, {freqs.size} frequencies")

    def step():
        e.grad = None
        S = adj.SParamsOfPermittivity.apply(e, prob)
        J = (S[:, 1, 0].abs() ** 2).sum()
        J.backward()
        return float(J.detach())

    for _ in range(warmup):
        step()
    torch.cuda.synchronize()
    t0 = time.perf_counter()
    for _ in range(steps):
        Jv = step()
    torch.cuda.synchronize()
    el = (time.perf_counter() - t0) / steps
    # forward alone (S only), and the gradient kernel alone
    t1 = time.perf_counter()
    S, sols = prob.solve(eps0)
    torch.cuda.synchronize()
    t_fwd = time.perf_counter() - t1
    G = np.zeros((freqs.size, 2, 2), complex)
    G[:, 1, 0] = 2 * S[:, 1, 0]
    t2 = time.perf_counter()
    grad = prob.gradient(sols, G)
    torch.cuda.synchronize()
    t_grad = time.perf_counter() - t2
    result = {
        "metric": METRIC, "value": 1.0 / el, "unit": "grads/s", "n_gpus": 1, "steps": steps,
        "warmup": warmup, "ms_per_step": 1e3 * el, "higher_is_better": True, "scaling": "weak",
        "vs_baseline": None, "dtype": "c128", "data": "synthetic (seeded structured Kuhn mesh, seeded eps_r)",
        "config": {"workload": f"wr90_{nx}x{ny}x{nz}_adjoint_16f", "design_tets": int(design.size),
                   "edges": mesh.num_edges(), "freqs_hz": [9e9, 11e9, 16], "tol": args.tol},
        "J": Jv, "forward_ms": 1e3 * t_fwd, "gradient_kernel_ms": 1e3 * t_grad,
        "grad_over_forward": el / t_fwd, "max_abs_grad": float(grad.abs().max()),
        "iterations": [int(min(i["iters"].min() for i in prob.last_infos)),
                       int(max(i["iters"].max() for i in prob.last_infos))],
    }
    return result


def cells_main(args):
    """512 periodic unit cells of uneven size in one batched solve: 5 x 5 mm cells, 0.5 mm
    substrate eps 3.5(1 - 0.002j) under 10 mm of air, centred PEC patch w, l ~ U[1, 4.5] mm
    snapped to the grid on the substrate top, hex grid nx = ny ~ U{10..28}, nz ~ U{12..24}
    (seeded splitmix64), Floquet phase along x from theta ~ U[0, 40] degrees at 10 GHz,
    ports top and bottom as the unit-cell caller sets them up (unit_cell.py:700-763): the first
    nodal TE mode of each face (solve_port_eigens, the square face's degenerate TE10/TE01
    pair resolved to E along y) as a wave port (build_wave_port), use_port_abc, eps regions
    substrate and air.  One step = assembly and periodic reduction of every cell on the GPU,
    the batched BiCGSTAB of all 2 x cells systems, and the S extraction."""
    import numpy as np
    import torch
    from edgefem_amd import meshgen as mg
    from edgefem_amd import pyedgefem as em
    torch.cuda.set_device(0)
    ncell = args.cells
    f = 10e9
    k0 = 2 * math.pi * f / em.C0
    u = lambda stream: mg.uniform01(stream, np.arange(ncell, dtype=np.int64))
    nxy = 10 + np.floor(u(11) * 19).astype(int)
    nzs = 12 + np.floor(u(12) * 13).astype(int)
    if args.small:
        nxy = np.minimum(nxy, 8)
        nzs = np.minimum(nzs, 10)
    pw_, pl_ = 1e-3 + 3.5e-3 * u(13), 1e-3 + 3.5e-3 * u(14)
    theta = 40.0 * u(15)
    t0 = time.perf_counter()
    cells, tets = [], 0
    for c in range(ncell):
        g = mg.unit_cell(n=int(nxy[c]), nz=int(nzs[c]), patch=(float(pw_[c]), float(pl_[c])), seed=1000 + c)
        pm = em.mesh_from_arrays(g.node_ids, g.xyz, g.tets, g.tet_phys, g.tris, g.tri_phys)
        tets += g.tets.shape[0]
        T = mg.CELL_TAGS
        bc = em.build_edge_pec(pm, T["patch"])
        pbc = em.build_periodic_pairs(pm, T["x0"], T["x1"], np.array([5e-3, 0.0, 0.0]))
        em.set_floquet_phase(pbc, np.array([k0 * math.sin(math.radians(theta[c])), 0.0]))
        ports = []
        for tag in (T["port_top"], T["port_bottom"]):
            surf = em.extract_surface_mesh(pm, tag)
            mode = em.solve_port_eigens(surf.mesh, 1, 2 * math.pi * f, 1.0, 1.0, em.ModePolarization.TE)[0]
            ports.append(em.build_wave_port(pm, surf, mode))
        params = em.MaxwellParams()
        params.omega = 2 * math.pi * f
        params.eps_r_regions = {T["substrate"]: 3.5 * (1 - 0.002j), T["air"]: 1.0}
        params.use_port_abc = True
        cells.append((pm, params, bc, pbc, ports))
    log(f"cells: {ncell} meshes, {tets} tets, setup {time.perf_counter() - t0:.1f}s")
    opts = em.SolveOptions()
    opts.tolerance = args.tol
    opts.max_iterations = args.max_it
    for _ in range(args.warmup):
        em.calculate_sparams_periodic_batch(cells, solver_options=opts)
    torch.cuda.synchronize()
    st = {}
    ts = time.perf_counter()
    for _ in range(args.steps):
        S = em.calculate_sparams_periodic_batch(cells, solver_options=opts, _state=st)
    torch.cuda.synchronize()
    el = (time.perf_counter() - ts) / args.steps
    nsol = 2 * ncell
    its = st["iters"]
    dS = None
    cpu = None
    if not args.no_check:
        from oracle import edgefem_oracle as O
        pick = np.argsort(nxy * nxy * nzs)[: min(3, ncell)]
        errs = []
        tcpu = 0.0
        for c in (int(i) for i in pick):
            pm, params, bc, pbc, ports = cells[c]
            g = mg.unit_cell(n=int(nxy[c]), nz=int(nzs[c]), patch=(float(pw_[c]), float(pl_[c])), seed=1000 + c)
            om = O.mesh_from_arrays(g.node_ids, g.xyz, g.tet_ids, g.tet_phys, g.tri_ids, g.tri_phys)
            T = mg.CELL_TAGS
            opbc = O.build_periodic_pairs(om, T["x0"], T["x1"], (5e-3, 0, 0))
            O.set_floquet_phase(opbc, (k0 * math.sin(math.radians(theta[c])), 0.0))
            oports = []
            for tag in (T["port_top"], T["port_bottom"]):
                surf = O.extract_surface_mesh(om, tag)
                oports.append(O.build_wave_port(om, surf, O.solve_port_eigens(surf, 1, params.omega, 1.0, 1.0, "TE")[0]))
            t1 = time.perf_counter()
            So = O.calculate_sparams_periodic(
                om, O.Params(omega=params.omega, eps_r_regions={T["substrate"]: 3.5 * (1 - 0.002j), T["air"]: 1.0},
                             use_port_abc=True), O.build_edge_pec(om, T["patch"]), opbc, oports)
            tcpu += time.perf_counter() - t1
            errs.append(float(np.abs(S[c] - So).max()))
        dS = max(errs)
        cpu = dict(value=2 * len(pick) / tcpu, unit="S-solves/s", cores=1, kind="port",
                   sample=f"oracle (assembly, sparse periodic elimination, SuperLU per excitation) on the "
                          f"{len(pick)} smallest cells, 1 host core, {tcpu:.1f} s")
    # the committed FP64 fixture of 8 of these cells (the smallest, the largest and six between)
    dS_gold = None
    gpath = os.path.join(ROOT, "tests", "golden", "c4_cells.npz")
    if os.path.exists(gpath) and ncell == 512 and not args.small:
        gold = np.load(gpath)
        dS_gold = max(float(np.abs(S[int(c)] - Sg).max()) for c, Sg in zip(gold["cells"], gold["S"]))
    result = {
        "metric": METRIC, "value": nsol / el, "unit": "S-solves/s", "n_gpus": 1, "steps": args.steps,
        "warmup": args.warmup, "ms_per_step": 1e3 * el, "higher_is_better": True, "scaling": "weak",
        "vs_baseline": None, "dtype": "c128", "data": "synthetic (seeded unit cells)",
        "delta_S_vs_fp64_golden_8_cells": dS_gold,
        "config": {"workload": f"unit_cells_{ncell}_uneven_bloch_x", "cells": ncell, "tets": int(tets),
                   "rows": int(st["n"]), "nnz": int(st["nnz"]), "freq_hz": f, "tol": args.tol,
                   "solver": "batched BiCGSTAB + periodic auxiliary preconditioner, per-system masks"},
        "solve_ms": 1e3 * st.get("solve_s", float("nan")),
        "iterations": [int(its.min()), int(its.max())], "max_rel_residual": float(st["resid"].max()),
        "delta_S_vs_fp64_lu_sampled": dS, "cpu_baseline": cpu,
    }
    print(json.dumps(result))


def traffic_from_pmc(fetch_csv, write_csv, out_path, variant, commands):
    """Per-kernel HBM bytes per launch from two rocprofv3 --pmc runs (FETCH_SIZE, then
    WRITE_SIZE, KiB per dispatch), each scaled by what the same counter reads for the
    calibration copy (the largest device-to-device copy of the run, of known size)."""
    import csv
    import numpy as np

    def load(path, counter):
        per = {}
        copy_kib = None
        for r in csv.DictReader(open(path)):
            if r["Counter_Name"] != counter:
                continue
            name = r["Kernel_Name"]
            v = float(r["Counter_Value"])
            if name.startswith("__amd_rocclr_copyBuffer"):
                copy_kib = v if copy_kib is None else max(copy_kib, v)   # the calibration copy is the largest
            short = name.replace("void ", "").replace("(anonymous namespace)::", "").replace("efm::", "")
            per.setdefault(short.split("(")[0], []).append(v)
        return {k: float(np.mean(v)) for k, v in per.items()}, copy_kib

    fetch, cf = load(fetch_csv, "FETCH_SIZE")
    write, cw = load(write_csv, "WRITE_SIZE")
    copy_bytes = None
    prev = json.load(open(out_path)) if os.path.exists(out_path) else {}
    copy_bytes = prev.get("calibration", {}).get("copy_bytes", 512 << 20)
    ff = copy_bytes / (cf * 1024.0)
    fw = copy_bytes / (cw * 1024.0)
    kernels = {}
    for k in fetch:
        if k in write and not k.startswith(("__amd", "at::")):
            kernels[k] = dict(fetch_KiB=fetch[k], write_KiB=write[k],
                              hbm_bytes=fetch[k] * 1024 * ff + write[k] * 1024 * fw)
    out = dict(workload="wr90_48x24x192_slab_10ghz", variant=variant, commands=commands,
               calibration=dict(copy_bytes=copy_bytes, copy_fetch_KiB=cf, copy_write_KiB=cw, fetch_factor=ff,
                                write_factor=fw,
                                note="the calibration copy's FETCH_SIZE and WRITE_SIZE against its known size "
                                     "give the factors that every kernel's counters are scaled by"),
               kernels=kernels)
    json.dump(out, open(out_path, "w"), indent=1)
    print(json.dumps(out, indent=1))


def main():
    ap = argparse.ArgumentParser()
    ap.add_argument("--workload", choices=["c2", "sweep", "adjoint", "cells", "fields", "media"], default="c2",
                    help="c2: one large solve (default); sweep: batched 256-system frequency sweep; "
                         "media: the same sweep with a Lorentz slab and a PML slab (per-system correction); "
                         "adjoint: permittivity gradient over 16 frequencies; cells: 512 periodic unit cells; "
                         "fields: E and curl E of every tet of the C2 guide")
    ap.add_argument("--cells", type=int, default=512, help="cells: number of unit cells")
    ap.add_argument("--sweep-precision", choices=["fp32", "fp64"], default="fp64",
                    help="sweep: complex-FP32 inner iterations with FP64 refinement, or plain FP64")
    ap.add_argument("--sweep-spmv", type=int, default=-1, choices=[-1, 0, 1, 2],
                    help="sweep: SpMV and nodal kernels (0 one slice per system group, 1 pipelined, 2 LDS-tiled "
                         "SpMV with the pipelined nodal kernel, -1 autotuned)")
    ap.add_argument("--sweep-graph", type=int, default=1, choices=[0, 1],
                    help="sweep: replay full 16-iteration chunks from a captured HIP graph (1, the default: "
                         "1,036 against 1,021-1,026 S-solves/s with plain launches on the 256-system sweep)")
    ap.add_argument("--sweep-iteration", type=int, default=2, choices=[0, 1, 2],
                    help="sweep: 0 standard COCG (p in its own pass), 1 Chronopoulos-Gear (one reduction), "
                         "2 standard with u rebuilt in the p pass (no u store)")
    ap.add_argument("--gpus", type=int, default=1)
    ap.add_argument("--steps", type=int, default=3)
    ap.add_argument("--warmup", type=int, default=1)
    ap.add_argument("--mode", choices=["fp64", "fp32m"], default="fp64",
                    help="matrix storage: complex FP64, or complex FP32 (FP64 vectors and reductions)")
    ap.add_argument("--tol", type=float, default=1e-10)
    ap.add_argument("--prec", choices=["aux", "jacobi"], default="aux",
                    help="edge Jacobi plus nodal gradient correction, or edge Jacobi only")
    ap.add_argument("--max-it", type=int, default=100000)
    ap.add_argument("--no-cpu", action="store_true")
    ap.add_argument("--no-check", action="store_true")
    ap.add_argument("--no-extra", action="store_true",
                    help="c2: skip the sweep and adjoint summaries added to the line")
    ap.add_argument("--cpu-seconds", type=float, default=20.0)
    ap.add_argument("--small", action="store_true", help="16x8x32 guide (quick check)")
    ap.add_argument("--spmv-variant", type=int, default=-1,
                    help="fixed COCG SpMV variant (-1: autotune at the first solve)")
    ap.add_argument("--calibrate-copy", type=int, default=0, metavar="MB",
                    help="first run a plain device-to-device copy of this many MB (counter calibration)")
    ap.add_argument("--traffic-from", nargs=2, metavar=("FETCH_CSV", "WRITE_CSV"),
                    help="write profiles/traffic_c2.json from the counter CSVs of two rocprofv3 --pmc runs")
    args = ap.parse_args()
    if args.traffic_from:
        cmd = ("rocprofv3 --pmc {c} --output-format csv -d pmc_{d} -o {d} -- python -u bench.py --no-cpu "
               "--no-check --steps 1 --warmup 0 --max-it 20 --spmv-variant {v} --calibrate-copy 512")
        v = args.spmv_variant if args.spmv_variant >= 0 else 9
        traffic_from_pmc(args.traffic_from[0], args.traffic_from[1], os.path.join(ROOT, "profiles", "traffic_c2.json"),
                         v, [cmd.format(c="FETCH_SIZE", d="fetch", v=v), cmd.format(c="WRITE_SIZE", d="write", v=v)])
        return
    if args.gpus != 1:
        print(json.dumps({"error": "edgefem_amd runs on one GPU per process; --gpus must be 1 "
                                   "(multi-GPU sharding is a design note only)"}))
        sys.exit(2)
    if args.workload == "fields":
        fields_main(args)
        return
    if args.workload == "media":
        media_main(args)
        return
    if args.workload == "sweep":
        sweep_main(args)
        return
    if args.workload == "adjoint":
        adjoint_main(args)
        return
    if args.workload == "cells":
        cells_main(args)
        return

    import numpy as np
    import torch
    assert torch.cuda.is_available(), "bench needs a HIP device"
    torch.cuda.set_device(0)
    freq = 10e9
    t0 = time.time()
    if args.small:
        g, eps, mesh, bc, ports, params = build_workload(16, 8, 32, 50e-3, freq)
    else:
        g, eps, mesh, bc, ports, params = build_workload(freq=freq)
    log(f"setup: mesh + topology + ports {time.time() - t0:.1f}s; n={mesh.num_edges()} nnz={mesh.topology.nnz}")
    dtype = torch.complex128 if args.mode == "fp64" else torch.complex64
    pipe = Pipeline(mesh, bc, ports, params, dtype, args.tol, args.max_it, args.prec)
    if args.spmv_variant >= 0:
        pipe.solver.set_spmv_variant(args.spmv_variant)
    if args.calibrate_copy > 0:
        src = torch.ones(args.calibrate_copy * (1 << 20) // 8, dtype=torch.float64, device="cuda")
        dst = torch.empty_like(src)
        dst.copy_(src)
        torch.cuda.synchronize()
        log(f"calibration copy: {src.numel() * 8} bytes")
        del src, dst
    for w in range(args.warmup):
        S = pipe.step()
        log(f"warmup {w}: iters {pipe.last_info['iters']} S21 {S[1, 0]:.6f}")
    torch.cuda.synchronize()
    pipe.solver.profile(True)
    torch.cuda.synchronize()
    ts = time.perf_counter()
    for k in range(args.steps):
        S = pipe.step()
        log(f"step {k}: iters {pipe.last_info['iters']}")
    torch.cuda.synchronize()
    te = time.perf_counter()
    spmv_ms, spmv_n = pipe.solver.profile_read()
    variant, tune_ms = pipe.solver.spmv_variant()
    pipe.solver.profile(False)
    elapsed = te - ts
    ms_per_step = 1e3 * elapsed / args.steps
    value = len(ports) * args.steps / elapsed
    n = mesh.num_edges()
    nnz = mesh.topology.nnz
    R = len(ports)
    s_val = 16 if args.mode == "fp64" else 8
    # algorithmic bytes of the chosen SpMV: the fused form reads z and p_old and writes p and
    # q per row and RHS (4 passes); variants 9-11 (p formed in its own pass, tiled, sliced-ELL
    # or with split values) gather p and write q (2 passes); the sliced-ELL padding slots are
    # not algorithmic bytes, and variant 11's saving (8 B instead of 16 per real value) is
    # not subtracted: the bytes are those of a complex value per entry
    spmv_bytes = nnz * (s_val + 4) + 4 * (n + 1) + R * n * 16 * (2 if variant in (9, 10, 11) else 4)
    avg_ms = spmv_ms / max(spmv_n, 1)
    achieved = spmv_bytes / (avg_ms * 1e-3)
    iters = int(max(pipe.last_info["iters"]))

    traffic = None
    tpath = os.path.join(ROOT, "profiles", "traffic_c2.json")
    if os.path.exists(tpath) and not args.small:
        tj = json.load(open(tpath))
        if tj.get("variant") == variant and args.mode == "fp64":
            for k, v in tj["kernels"].items():
                # the SpMV of the iteration
                if k.startswith(("k_cocg_spmv", "k_cocgz_spmv")):
                    traffic = v["hbm_bytes"] / 1e9
    dS = None
    gold = golden_c2() if not args.small else None
    if gold is not None:
        dS = float(np.abs(S - gold["S"]).max())
        log(f"|dS| vs the FP64 golden = {dS:.3e} ({gold['method']}; residual {max(gold['residual']):.1e})")
    cpu = None
    if not args.no_cpu:
        cpu = cpu_baseline(g, eps, freq, iters, args.prec, budget_s=args.cpu_seconds)
        cpu.pop("seconds_per_solve_pair", None)
        if not args.small:
            log("  cpu baseline: the reference's solve (ILUT sample)")
            cpu["reference_solver"] = reference_solver_c2_sample(g, eps, freq, budget_s=15.0)
    workload = "wr90_48x24x192_slab_10ghz" if not args.small else "wr90_16x8x32_air_10ghz"
    result = {
        "metric": METRIC, "value": value, "unit": "S-solves/s", "n_gpus": 1, "steps": args.steps,
        "warmup": args.warmup, "ms_per_step": ms_per_step, "higher_is_better": True, "scaling": "weak",
        "vs_baseline": None, "dtype": "c128" if args.mode == "fp64" else "c64-matrix/c128-vectors",
        "data": "synthetic (seeded structured Kuhn mesh, seeded per-tet permittivity)",
        "config": {"workload": workload, "tets": mesh.num_tets(), "edges": n, "nnz": nnz,
                   "freq_hz": freq, "excitations": R,
                   "solver": "COCG+" + ("Jacobi+nodal-gradient" if args.prec == "aux" else "Jacobi"),
                   "tol": args.tol,
                   "mode": args.mode},
        "roofline": {"bound": "hbm", "achieved": achieved / 1e9, "peak": HBM_PEAK / 1e9, "unit": "GB/s",
                     "frac": achieved / HBM_PEAK, "traffic": traffic, "traffic_unit": "GB per launch",
                     "kernel": {9: "k_cocgz_spmv", 10: "k_cocgz_sell", 11: "k_cocgz_spmv (split values)"}.get(variant, "k_cocg_spmv"),
                     "bytes_per_launch": spmv_bytes, "avg_launch_ms": avg_ms, "launches": spmv_n,
                     "variant": variant, "autotune_ms": [round(float(t), 4) for t in tune_ms]},
        "cpu_baseline": cpu,
        "iterations": iters,
        "delta_S_vs_fp64_golden": dS,
        "golden": None if gold is None else {"file": "tests/golden/c2_sparams_10ghz.npz", "method": gold["method"],
                                             "true_rel_residual": gold["residual"]},
        "S21": [S[1, 0].real, S[1, 0].imag],
    }
    if not args.no_extra and not args.small:
        # the other two parts of the metric, each on its own workload: the batched sweep
        # (S-solves/s of 256 systems and its speedup over the plain path) and the adjoint
        # gradient (grads/s); compact summaries of their own bench lines
        del pipe
        torch.cuda.empty_cache()

        def extra(key, fn):
            # a failing summary is reported in its own key; the C2 line above still prints
            try:
                result[key] = fn()
            except Exception as e:  # noqa: BLE001 - reported in the line, traceback on stderr
                import traceback
                traceback.print_exc()
                result[key] = {"error": f"{type(e).__name__}: {e}"}
            torch.cuda.empty_cache()

        def sweep_summary(prec):
            sw = run_sweep(args, 2, 1, prec, cpu=(prec == "fp64"))
            out = {k: sw[k] for k in ("value", "unit", "ms_per_step", "speedup_vs_plain", "plain_path_solves_per_s",
                                      "delta_S_vs_fp64_lu", "iterations", "max_true_rel_residual", "dtype",
                                      "cpu_baseline")}
            out["workload"] = sw["config"]["workload"]
            return out

        def adjoint_summary():
            ad = run_adjoint(args, 1, 0)
            out = {k: ad[k] for k in ("value", "unit", "ms_per_step", "forward_ms", "gradient_kernel_ms",
                                      "grad_over_forward", "iterations")}
            out["workload"] = ad["config"]["workload"]
            out["design_tets"] = ad["config"]["design_tets"]
            return out

        def media_summary():
            md = run_media(args, 1, 1)
            out = {k: md[k] for k in ("value", "unit", "ms_per_step", "per_frequency_solves_per_s",
                                      "speedup_vs_per_frequency", "delta_S_vs_per_frequency", "iterations")}
            out["workload"] = md["config"]["workload"]
            return out

        log("extra: batched sweep (fp64 and fp32 vectors)")
        for prec in ("fp64", "fp32"):
            extra(f"sweep_{prec}", lambda prec=prec: sweep_summary(prec))
        log("extra: adjoint gradient")
        extra("adjoint", adjoint_summary)
        log("extra: batched sweep with frequency-dependent media")
        extra("media_sweep", media_summary)
    print(json.dumps(result))


if __name__ == "__main__":
    main()
