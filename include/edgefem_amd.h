/*
 * edgefem_amd.h -- C ABI of the MI355X (gfx950) implementation of EdgeFEM's
 * frequency-domain Nedelec hot path.
 *
 * Conventions
 *   - Every pointer argument is a DEVICE pointer unless its name starts with h_.
 *   - Complex values are interleaved (re, im): "c128" = 2 doubles, "c64" = 2 floats.
 *   - Indices are 0-based int32; gmsh node IDs are int64.
 *   - `stream` is a hipStream_t passed as void* (NULL = the null stream).  All work is
 *     enqueued on it; functions that return host-visible results synchronise it.
 *   - Every function returns an efm_status; on failure efm_last_error() gives a
 *     message (thread-local).  No C++ exception crosses this boundary.
 *   - Objects (efm_topology, efm_solver) are not thread-safe; one thread per object.
 *
 * Multi-RHS layout: a block of nrhs vectors of length n is stored system-minor,
 * element (row r, system s) at [r * nrhs + s].
 *
 * Each entry point names the reference function it replaces (file:line in the
 * EdgeFEM sources).
 */
#ifndef EDGEFEM_AMD_H
#define EDGEFEM_AMD_H

#include <stddef.h>
#include <stdint.h>

#ifdef __cplusplus
extern "C" {
#endif

typedef enum {
  EFM_OK = 0,
  EFM_ERR_ARG = 1,       /* invalid argument / shape */
  EFM_ERR_HIP = 2,       /* HIP runtime error */
  EFM_ERR_NOMEM = 3,     /* device allocation failed */
  EFM_ERR_NOT_CONVERGED = 4,
  EFM_ERR_MESH = 5       /* mesh does not satisfy an assumption (message says which) */
} efm_status;

enum { EFM_C128 = 0, EFM_C64 = 1 };                 /* matrix value storage */
enum { EFM_METHOD_COCG = 0, EFM_METHOD_BICGSTAB = 1 };

const char *efm_last_error(void);
int efm_version(void);

/* ------------------------------------------------------------------------
 * Mesh topology: global edges in first-seen order, orientation signs, the CSR
 * pattern of the assembled matrix and the element->nnz gather maps.
 * Replaces build_edges() (src/mesh_gmsh.cpp:104-146, key rule include/edgefem/mesh.hpp:56-60)
 * and the pattern that setFromTriplets builds (src/assemble_maxwell.cpp:205-207).
 * tet_nodes / tri_nodes index node_ids (0-based positions in $Nodes order).
 * ---------------------------------------------------------------------- */
typedef struct efm_topology efm_topology;

int efm_topology_create(int64_t n_nodes, const int64_t *node_ids, int64_t n_tets,
                        const int32_t *tet_nodes, int64_t n_tris, const int32_t *tri_nodes,
                        void *stream, efm_topology **out);
void efm_topology_destroy(efm_topology *topo);

typedef struct {
  int64_t n_nodes, n_tets, n_tris, n_edges, nnz;
} efm_topology_sizes;
int efm_topology_get_sizes(const efm_topology *topo, efm_topology_sizes *h_out);

/* Copy topology arrays into caller buffers (any may be NULL):
 *   edges[n_edges*2] int64 (min ID, max ID); tet_edges[T*6] int32; tet_sign[T*6] int8;
 *   tri_edges[F*3] int32; tri_sign[F*3] int8; row_ptr[n_edges+1], col_idx[nnz] int32. */
int efm_topology_export(const efm_topology *topo, int64_t *edges, int32_t *tet_edges,
                        int8_t *tet_sign, int32_t *tri_edges, int8_t *tri_sign,
                        int32_t *row_ptr, int32_t *col_idx, void *stream);

/* Node incidence of the edges, for the auxiliary gradient preconditioner:
 *   edge_nodes[2*n_edges] int32 node positions (lower-ID end, higher-ID end);
 *   n2e_ptr[n_nodes+1], n2e[2*n_edges] int32: incident edges of every node, ascending.
 * No reference counterpart (the reference's ILUT needs no mesh data). */
int efm_topology_export_nodes(const efm_topology *topo, int32_t *edge_nodes, int32_t *n2e_ptr,
                              int32_t *n2e, void *stream);

/* ------------------------------------------------------------------------
 * Assembly
 * ---------------------------------------------------------------------- */
/* K = sum_e (K_e / mu_e) s_i s_j and M = sum_e (eps_e M_e) s_i s_j on the CSR pattern,
 * Dirichlet rows/cols zeroed, K diagonal 1 and M diagonal 0 on Dirichlet edges.
 * Replaces assemble_maxwell_km (src/sweep.cpp:90-172), element matrices of
 * src/edge_basis.cpp:14-86.  xyz[n_nodes*3]; eps_tet[T] c128; inv_mu_tet[T] c128 or
 * NULL (= 1); dirichlet[n_edges] uint8 or NULL; K_val[nnz], M_val[nnz] c128. */
int efm_assemble_km(const efm_topology *topo, const double *xyz, const double *eps_tet,
                    const double *inv_mu_tet, const uint8_t *dirichlet, double *K_val,
                    double *M_val, void *stream);

/* A = sum_e (K_e/mu_e - k0^2 eps_e M_e) s_i s_j, + j k0 on the diagonal of edges with
 * abc_mask set, Dirichlet rows/cols zeroed with unit diagonal; k0 = omega / c0.
 * Replaces the volume, ABC and Dirichlet parts of assemble_maxwell
 * (src/assemble_maxwell.cpp:46-207, 244-265, 322-347).  abc_mask may be NULL. */
int efm_assemble_maxwell(const efm_topology *topo, const double *xyz, const double *eps_tet,
                         const double *inv_mu_tet, const uint8_t *dirichlet,
                         const uint8_t *abc_mask, double omega, double *A_val, void *stream);

/* Port surface mass M_s = sum over the selected triangles of s_i s_j M_tri(i,j),
 * Dirichlet rows/cols skipped, written as real values on the volume pattern (zero off
 * the port).  Replaces assemble_port_surface_mass (src/ports/wave_port.cpp:548-585) and
 * triangle_whitney_mass_matrix (src/edge_basis.cpp:88-130).  tri_sel[n_sel] = indices
 * of the port triangles in mesh order. */
int efm_port_surface_mass(const efm_topology *topo, const double *xyz, int64_t n_sel,
                          const int32_t *tri_sel, const uint8_t *dirichlet, double *Ms_val,
                          void *stream);

/* A[r, r] += vals[i] for distinct rows[i] (device int32 [m], c128 [m]) on the volume
 * pattern of topo: the diagonal port ABC terms of src/assemble_maxwell.cpp:274-320. */
int efm_add_diagonal(const efm_topology *topo, int64_t m, const int32_t *rows, const double *vals,
                     double *A_val, void *stream);

/* A_val = K_val - k0sq * M_val + sum_p coef[p] * Ms_val[p]  (c128 out, or c64 when
 * out_type == EFM_C64).  h_coef[n_ports] c128 on the host; Ms_vals[p] device arrays.
 * Replaces KMMatrices::combine (src/sweep.cpp:82-88) plus the jbeta M_s port operator
 * of src/assemble_maxwell.cpp:734-746. */
int efm_combine(const efm_topology *topo, const double *K_val, const double *M_val,
                double k0sq, int n_ports, const double *const *Ms_vals, const double *h_coef,
                int out_type, void *A_out, void *stream);

/* ------------------------------------------------------------------------
 * Krylov solver on a CSR matrix: COCG (complex-symmetric A) or BiCGSTAB, Jacobi or
 * auxiliary-gradient preconditioned, x0 = 0, stop when ||r_s|| <= tol * ||b_s|| per system.  FP64 vectors
 * and accumulation; the matrix may be stored c128 or c64.
 * Replaces solve_linear (src/solver.cpp:35-193, defaults include/edgefem/solver.hpp:16-36).
 * ---------------------------------------------------------------------- */
typedef struct efm_solver efm_solver;

int efm_solver_create(int64_t n, int64_t nnz, const int32_t *row_ptr, const int32_t *col_idx,
                      int nrhs, int value_type, void *stream, efm_solver **out);
void efm_solver_destroy(efm_solver *s);
/* Install matrix values (c128 or c64 per value_type); also computes the Jacobi diagonal
 * and, when a gradient is set, the nodal diagonal diag(G^T A G). */
int efm_solver_set_matrix(efm_solver *s, const void *values, void *stream);

/* Enable the auxiliary-space preconditioner P = D^{-1} + G diag(G^T A G)^{-1} G^T, with G
 * the discrete gradient of the mesh (arrays from efm_topology_export_nodes; they must
 * outlive the solver) and Dirichlet edges removed from G.  n_nodes <= 0 switches it off.
 * Call before efm_solver_set_matrix.  Replaces the preconditioner choice of
 * src/solver.cpp:53-62 (IncompleteLUT) for mesh-backed systems. */
int efm_solver_set_gradient(efm_solver *s, int64_t n_nodes, const int32_t *edge_nodes,
                            const int32_t *n2e_ptr, const int32_t *n2e, const uint8_t *dirichlet,
                            void *stream);

/* Auxiliary preconditioner with a complex gradient incidence, for systems whose edge
 * space has been reduced (periodic elimination): row e of G is
 * h_edge_coefs[2e] at node h_edge_nodes[2e] plus h_edge_coefs[2e+1] at node
 * h_edge_nodes[2e+1] (host arrays [2*n]; c128 coefficients; node -1 = no entry).  The
 * nodal diagonal diag(G^T A G) is formed at efm_solver_set_matrix.  BiCGSTAB only.
 * n_nodes <= 0 switches it off; it replaces efm_solver_set_gradient.  No reference
 * counterpart (the reference uses ILUT). */
int efm_solver_set_gradient_complex(efm_solver *s, int64_t n_nodes, const int32_t *h_edge_nodes,
                                    const double *h_edge_coefs, void *stream);

/* Low-rank operator term: the solver iterates on A_eff = A + sum_p h_coef[p] w_p w_p^H, with
 * w_p given on a compact edge list (gidx[p*len + i] >= 0, -1 = padding; gval the values;
 * device arrays, copied).  This is the rank-1 port loading w w^H / Z0 of
 * src/assemble_maxwell.cpp:209-242, applied inside the operator instead of as dense
 * blocks added to the sparsity pattern; the Jacobi diagonal includes it.  Ports must be
 * disjoint (an edge in two lists is rejected).  n_terms = 0 removes the term.  Call before
 * efm_solver_set_matrix.  At most 32 terms. */
int efm_solver_set_lowrank(efm_solver *s, int n_terms, int64_t len, const int32_t *gidx,
                           const double *gval, const double *h_coef, void *stream);

typedef struct {
  int method;            /* EFM_METHOD_COCG or EFM_METHOD_BICGSTAB */
  double tolerance;      /* relative residual per system */
  int max_iterations;
  int check_every;       /* host convergence poll interval (iterations) */
} efm_solve_opts;

typedef struct {
  int iterations;        /* max over systems */
  int converged;         /* 1 if every system converged */
  double max_rel_residual;
} efm_solve_result;

/* B[n*nrhs] c128 in, X[n*nrhs] c128 out.  h_iters / h_resid (nrhs each, may be NULL)
 * receive per-system iteration counts and final recursive relative residuals. */
int efm_solver_solve(efm_solver *s, const double *B, double *X, const efm_solve_opts *h_opts,
                     efm_solve_result *h_result, int32_t *h_iters, double *h_resid,
                     void *stream);

/* The COCG SpMV has several row-to-lane mappings (0-2: 4, 8, 16 lanes per row; 3-4:
 * LDS-tiled with 4 or 8 lanes per row; 6: LDS-tiled with the next tile's loads issued
 * while the current tile's gathers are in flight; 5, 7 and 8: removed (slower on every
 * matrix measured; set returns EFM_ERR_ARG, the tuning time is -1); 9: the same
 * COCG iteration with p = z + beta p formed in a streaming pass of its own, so the SpMV
 * gathers only p (R values per column) and writes only q, COCG without low-rank terms; 10:
 * the same iteration with a sliced-ELL SpMV: slices of 16 rows, 4 lanes per row, the entries
 * of a slice stored step-major so that every step is one coalesced load, rows ordered by
 * length inside windows of 4096, no LDS; built on the device at first use); 11: variant 9
 * with the values stored as a real part per entry plus the imaginary parts of the tiles
 * that have any (a tile of real values streams 8 B per value instead of 16; complex-double
 * matrices; rebuilt on the device after every set_matrix).  By
 * default (-1) the first solve times every variant on the installed matrix and keeps the
 * fastest; a fixed variant makes results reproducible bit for bit across runs.
 * h_tune_ms (12 entries, may be NULL) receives the
 * measured average ms per launch (-1 = not available).  No reference counterpart. */
int efm_solver_set_spmv_variant(efm_solver *s, int variant);
int efm_solver_get_spmv_variant(const efm_solver *s, int *h_variant, double *h_tune_ms);

/* y = A x (all nrhs systems), FP64 accumulation. */
int efm_solver_spmv(efm_solver *s, const double *x, double *y, void *stream);

/* Per-kernel timing of the fused SpMV launches inside efm_solver_solve, taken with HIP
 * events on the solve's stream: enable, then read the accumulated totals. */
int efm_solver_profile(efm_solver *s, int enable);
int efm_solver_profile_read(efm_solver *s, double *h_spmv_ms_total, int64_t *h_spmv_launches);

/* ------------------------------------------------------------------------
 * Port overlaps for S-parameter extraction: h_out[j * nrhs + a] = g_j^H x_a with g_j
 * given on a compact list of port edges (negative gidx entries are padding).  With
 * g_j = M_s,j e_j this is the modal amplitude of src/assemble_maxwell.cpp:773-783; with
 * g_j = w_j it is the port voltage V_j of 374-391 (the caller applies the -delta /
 * sqrt(Z0) terms).  X[n*nrhs] system-minor; gidx[n_ports*len] int32; gval c128;
 * h_out c128 [n_ports*nrhs]; synchronises the stream.
 * ---------------------------------------------------------------------- */
int efm_port_overlap(int64_t n, int n_ports, int64_t len, const int32_t *gidx, const double *gval,
                     const double *X, int nrhs, double *h_out, void *stream);

/* efm_port_overlap without any device allocation, for callers that cannot allocate (code
 * under HIP graph capture, tight serving loops, harnesses that forbid hipMalloc).  No
 * counterpart in the reference.  Same arguments, kernel and result as efm_port_overlap,
 * but the results are staged in a fixed 4096-entry (64 KiB) device array that the library
 * holds on every GPU it runs on until the process ends, instead of a buffer it allocates:
 * n_ports * nrhs must not exceed 4096 (EFM_ERR_ARG otherwise).  Every call on one GPU
 * shares that array, so calls that overlap on one GPU (from several host threads, or on
 * several streams) race; calls that follow each other are safe, because each call
 * synchronises its stream before it returns. */
int efm_port_overlap_noalloc(int64_t n, int n_ports, int64_t len, const int32_t *gidx,
                             const double *gval, const double *X, int nrhs, double *h_out,
                             void *stream);

/* Direct solve A X = B by a dense LU on the GPU (rocSOLVER zgetrf/zgetrs): the SparseLU
 * path of solve_linear (src/solver.cpp:11-33), taken when use_direct is set and as the
 * fallback after a failed iterative solve (src/solver.cpp:104-115).  A: CSR (device
 * row_ptr, col_idx, c128 values) plus n_lowrank terms c_p w_p w_p^H (lr_idx/lr_val device
 * [n_lowrank][lowrank_len], index -1 = padding; h_lr_coef c128 host), as
 * efm_solver_set_lowrank takes them.  B, X: device c128 [n][nrhs] (system-minor).
 * n <= 40000 (the dense factor: 16 n^2 bytes).  h_info: 0, or k > 0 when U(k, k) is
 * exactly zero (singular A; X is then not a solution). */
int efm_direct_solve(int64_t n, const int32_t *row_ptr, const int32_t *col_idx, const double *values,
                     int n_lowrank, int64_t lowrank_len, const int32_t *lr_idx, const double *lr_val,
                     const double *h_lr_coef, const double *B, double *X, int nrhs, int *h_info,
                     void *stream);

/* ------------------------------------------------------------------------
 * Batched frequency sweep: up to 1024 systems A_s = K - k0sq[s] M + sum_p coef[s,p] M_s,p
 * on one pattern, solved together by COCG with the auxiliary-gradient preconditioner and
 * per-system convergence.  Replaces the per-frequency loop of
 * RectWaveguideDesign.frequency_sweep (python/edgefem/designs/waveguide.py:394-434) over
 * calculate_sparams_eigenmode (src/assemble_maxwell.cpp:637-787), and the K/M reuse of
 * frequency_sweep (src/sweep.cpp:192-329).
 * K_val, M_val: c128 [nnz] from efm_assemble_km; h_Ms_vals[p]: device real [nnz] from
 * efm_port_surface_mass (ports must not share edges); dirichlet[n_edges] or NULL.
 * ---------------------------------------------------------------------- */
typedef struct efm_sweep efm_sweep;

int efm_sweep_create(const efm_topology *topo, const double *K_val, const double *M_val,
                     int n_ports, const double *const *h_Ms_vals, const uint8_t *dirichlet,
                     int max_systems, void *stream, efm_sweep **out);
void efm_sweep_destroy(efm_sweep *sw);
/* Mixed precision by reliable updates (residual replacement): COCG runs with its vectors
 * stored as inner_type (EFM_C64: complex FP32; arithmetic and reductions stay FP64) and
 * carries the solution increment since the last replacement; the FP64 solution x64, the
 * FP64 residual b - A x64 and the convergence test live in FP64.  A system's residual is
 * replaced by b - A x64 (x64 += increment first) once its recursive residual has dropped by
 * replace_factor since the last replacement, and when the recursive residual meets the
 * tolerance; the search direction is kept, so the Krylov recurrence is not restarted.  A
 * system converges only on its true FP64 residual.  Needs the standard COCG form (0 or 2).
 * No reference counterpart (the reference solves in FP64 only). */
typedef struct {
  int inner_type;          /* EFM_C128 or EFM_C64: storage of the Krylov vectors */
  double replace_factor;   /* in (0, 1), e.g. 0.01 */
} efm_refine_opts;

/* h_k0sq[nsys]; h_coef[nsys*n_ports] c128 (host); B, X: [n_edges][nsys] c128 (device);
 * h_refine NULL: one FP64 COCG solve per system to h_opts->tolerance (recursive
 * residual); otherwise reliable updates to h_opts->tolerance in the true FP64 residual.
 * h_iters / h_resid (nsys, may be NULL): iterations and final relative residual per
 * system (the true FP64 residual with h_refine).  h_opts->method must be EFM_METHOD_COCG. */
int efm_sweep_solve(efm_sweep *sw, int nsys, const double *h_k0sq, const double *h_coef,
                    const double *B, double *X, const efm_solve_opts *h_opts,
                    const efm_refine_opts *h_refine, int32_t *h_iters, double *h_resid,
                    void *stream);
/* Kernel variants of the sweep, for the SpMV and for the nodal part of the
 * preconditioner: 0 = one block slice per group of 64 systems; 1 = pipelined (bounds of
 * the next rows/nodes preloaded per wave, fixed-count gathers, groups walked in turn so
 * that each XCD's share of the gathered vectors stays in its L2); 2 (SpMV only) =
 * LDS-tiled: each tile of 8 consecutive rows loads its distinct columns (at most 72)
 * once per group of 64 systems into LDS and its rows read them there; available when
 * every tile fits (set returns an error otherwise); -1 (default) = the fastest, chosen on
 * the first solve: the iterations before its first convergence check cycle through the
 * candidates, each launch timed in place.  get: the variants in use (-1 before they are
 * chosen) and, when h_tune_ms is given (5 entries), the mean times in ms per launch
 * [SpMV 0, 1, 2, nodal 0, 1] (0 for a variant not timed).  Autotuner of this project; no
 * reference counterpart. */
int efm_sweep_set_variants(efm_sweep *sw, int spmv, int aux);
/* Iteration form of the batched COCG: 0 = standard COCG, p = u + beta p in a
 * streaming pass of its own and two reductions per iteration; 1 = Chronopoulos-Gear, one
 * reduction per iteration with A p carried by the recurrence s = w + beta s (fewer
 * passes per iteration, but the recurrence drifts and costs iterations on indefinite
 * Maxwell systems: 623 against 584 on the worst frequency of the WR-90 16x8x32 sweep);
 * 2 (default) = standard COCG with the preconditioned residual u = D^-1 r + G y never
 * stored: the update forms it for its reduction only and the p pass rebuilds it from r
 * and y (one vector pass fewer per iteration; 1,027 against 993 S-solves/s on the WR-90
 * 16x8x32 sweep).  Same solution either way up to the tolerance.  No
 * reference counterpart. */
int efm_sweep_set_iteration(efm_sweep *sw, int form);
int efm_sweep_get_variants(const efm_sweep *sw, int *h_spmv, int *h_aux, double *h_tune_ms);
/* Standard iteration forms (0, 2): enable = 1 captures one full convergence-check chunk of
 * iterations (check_every iterations of kernel launches, no memsets) into a HIP graph once
 * per solve, after the autotuner has chosen the kernels, and replays it for every later full
 * chunk; the convergence polls and residual replacements stay outside the graph.  The
 * results are those of the plain launch sequence.  0 (default) = plain launches.  No
 * reference counterpart. */
int efm_sweep_set_graph(efm_sweep *sw, int enable);
/* Frequency-dependent media (PML stretch, dispersive models) in one batched solve: every
 * system gets a correction C_s on a fixed sub-pattern,
 *   A_s = K - k0sq[s] M + sum_p coef[s,p] M_s,p + C_s,
 * so that K, M (given at efm_sweep_create, typically with the media at a reference
 * frequency) keep serving the preconditioner and the shared SpMV, and C_s carries what
 * differs at each system's frequency.  Replaces the per-frequency assembly and solve of
 * frequency_sweep (src/sweep.cpp:187-190, 330-346) when PML or a dispersive model is present.
 * set_correction: correction row m is matrix row h_rows[m] (each row at most once), with
 * entries h_ptr[m]..h_ptr[m+1]-1 (h_ptr[0] = 0) in columns h_cols (host arrays, copied);
 * n_rows = 0 removes the correction.  set_correction_values: d_vals c128 [entry][ncols]
 * in device memory owned by the caller; system s of the next solves (nsys systems) uses
 * column h_col_of_system[s] (systems of one frequency share a column); read by every later
 * solve until it is set again.  Needs the standard COCG form (0 or 2).  No reference
 * counterpart (the reference re-assembles and solves each frequency). */
int efm_sweep_set_correction(efm_sweep *sw, int64_t n_rows, const int32_t *h_rows,
                             const int32_t *h_ptr, const int32_t *h_cols, void *stream);
int efm_sweep_set_correction_values(efm_sweep *sw, int nsys, int ncols, const double *d_vals,
                                    const int32_t *h_col_of_system, void *stream);
/* The same correction as a few shared matrices with per-system weights,
 *   C_s = sum_q h_alpha[s][q] T_q,   q < nterms <= 8,
 * d_terms c128 [entry][8] in caller device memory (columns past nterms zero), h_alpha c128
 * [nsys][nterms] on the host: where the media vary by region (a dispersive model per tag, a
 * scalar PML stretch per region) every system's correction is a combination of each
 * region's K and M parts, and the values read per entry drop from one per frequency to the
 * terms (four when nterms <= 4, else all eight columns). */
int efm_sweep_set_correction_terms(efm_sweep *sw, int nterms, const double *d_terms, int nsys,
                                   const double *h_alpha, void *stream);

/* ------------------------------------------------------------------------
 * Periodic (Floquet) elimination on the sparse matrix: replaces the dense elimination of
 * assemble_maxwell_periodic (src/assemble_maxwell.cpp:496-574).  For every pair
 * k = (h_master[k], h_slave[k]) with neither edge PEC: row m += h_coef_row[k] * row s,
 * column m += h_coef_col[k] * column s (c128 coefficients: the reference uses
 * phase*orient and conj(phase)*orient); then the rows and columns of the non-PEC
 * slaves become unit rows/columns.  Inputs: CSR (row_ptr[n+1], col[nnz], val c128) on
 * the device; pairs, coefficients and h_dirichlet[n] (may be NULL) on the host.
 * Output: row_ptr_out[n+1] (device, always written) and *h_nnz; when col_out and
 * val_out are given (capacity cap >= *h_nnz) the reduced columns (sorted per row,
 * duplicates summed) and values are written too -- call once to size, once to fill.
 * Exact zeros stay in the pattern (the reference drops them).  Synchronises the stream.
 * ---------------------------------------------------------------------- */
int efm_periodic_reduce(int64_t n, const int32_t *row_ptr, const int32_t *col, const double *val,
                        int64_t n_pairs, const int32_t *h_master, const int32_t *h_slave,
                        const double *h_coef_row, const double *h_coef_col,
                        const uint8_t *h_dirichlet, int32_t *row_ptr_out, int32_t *col_out,
                        double *val_out, int64_t cap, int64_t *h_nnz, void *stream);

/* ------------------------------------------------------------------------
 * Many independent systems of different sizes in one batch (periodic unit cells): the
 * many-cell form of calculate_sparams_periodic (src/assemble_maxwell.cpp:576-635) as
 * UnitCellDesign calls it per cell (python/edgefem/designs/unit_cell.py:700-777).  The
 * reduced matrices of all cells form one block-diagonal CSR (h_row_off [ncells+1] row
 * offsets on the host; row_ptr [n+1], col [nnz] global columns inside the cell, val
 * c128 on the device, copied; a column outside its cell is rejected).  Each cell has
 * nrhs (1 or 2) excitations, and every (cell,
 * excitation) pair is its own BiCGSTAB system (own scalars and convergence).  The
 * preconditioner is the complex-incidence auxiliary one over the concatenated node
 * space (h_node_off [ncells+1], h_edge_nodes [2n] global node or -1, h_edge_coefs [2n]
 * c128).  Port loading: nrhs terms per cell, term q = cell * nrhs + p with entries
 * [h_port_off[q], h_port_off[q+1]) of h_port_idx (global edges, disjoint) / h_port_w
 * (c128) and coefficient h_port_coef[q] (c128): A_cell + sum c_q w_q w_q^H.
 * ---------------------------------------------------------------------- */
typedef struct efm_cells efm_cells;

int efm_cells_create(int64_t ncells, const int64_t *h_row_off, const int32_t *row_ptr,
                     const int32_t *col, const double *val, int64_t nnz, const int64_t *h_node_off,
                     const int32_t *h_edge_nodes, const double *h_edge_coefs, int nrhs,
                     const int64_t *h_port_off, const int32_t *h_port_idx, const double *h_port_w,
                     const double *h_port_coef, void *stream, efm_cells **out);
void efm_cells_destroy(efm_cells *cells);
/* SpMV kernel of the batch: 1 = entry sub-tiles of at most 1024 entries and 64 rows (the
 * default), 0 = rows over 8 lanes (the default when a row holds more than 1024 entries;
 * asking for 1 then fails with EFM_ERR_ARG).  Both give the same products up to the
 * summation order.  No reference counterpart. */
int efm_cells_set_spmv_variant(efm_cells *cells, int variant);
/* B, X: [n][nrhs] c128 (device); h_iters / h_resid: [ncells*nrhs] (may be NULL);
 * h_opts->method must be EFM_METHOD_BICGSTAB. */
int efm_cells_solve(efm_cells *cells, const double *B, double *X, const efm_solve_opts *h_opts,
                    int32_t *h_iters, double *h_resid, void *stream);

/* ------------------------------------------------------------------------
 * Adjoint gradient with respect to the per-tet permittivity (no reference counterpart;
 * it differentiates calculate_sparams_eigenmode, src/assemble_maxwell.cpp:637-787).
 * For every design tet e = design[d] (design NULL: all tets, n_design == n_tets):
 *   grad[d] = sum_t Re( h_w[t] * (s.L[:, l_t])_e^T M_e (s.X[:, r_t])_e )
 * with M_e the Whitney mass block recomputed from xyz and s the edge signs.  L[n][nl],
 * X[n][nx] c128 device (system-minor, e.g. the columns of one batched solve; L may equal
 * X); h_term_l/h_term_r/h_w (host, n_terms <= 2048): column pairs and complex weights.
 * grad: device double [n_design], overwritten; an invalid design index yields NaN.
 * With A complex-symmetric, lam_j = A^{-1} conj(g_j); the caller folds k0^2, the loss
 * weights and the lam/x scaling into h_w.  Synchronises the stream.
 * ---------------------------------------------------------------------- */
int efm_eps_gradient(const efm_topology *topo, const double *xyz, int64_t n_design,
                     const int32_t *design, int nl, const double *Lam, int nx, const double *X,
                     int n_terms, const int32_t *h_term_l, const int32_t *h_term_r,
                     const double *h_w, double *grad, void *stream);

/* ------------------------------------------------------------------------
 * Post-solve field evaluation.  Replaces the per-point evaluate_edge_field
 * (src/edge_basis.cpp:162-192; compute_barycentric 132-151) in the loops over all tets of
 * export_fields_vtk (src/io/vtk_export.cpp:14-40, E at the centroid) and
 * extract_huygens_surface (src/post/huygens_surface.cpp:95-125, E at the centroid and
 * curl E from whitney_edge_curls, src/edge_basis.cpp:34-45).  For every tet t and system
 * s < nrhs (1, 2, 3, 4 or 8):
 *   E[t][s]    = sum_e sign_e x[edge_e][s] (l_a grad l_b - l_b grad l_a) at the point whose
 *                barycentric weights are lam[t][0..3] (lam NULL: the centroid)
 *   curl[t][s] = sum_e sign_e x[edge_e][s] 2 grad l_a x grad l_b  (constant on the tet)
 * xyz: device double [N][3] in the topology's node order; x: device c128 [n][nrhs];
 * lam: device double [T][4] or NULL; E, curl: device c128 [T][nrhs][3] (curl may be NULL).
 * Stream-ordered; no synchronisation.
 * ---------------------------------------------------------------------- */
int efm_tet_fields(const efm_topology *topo, const double *xyz, int nrhs, const double *x,
                   const double *lam, double *E, double *curl, void *stream);

/* ------------------------------------------------------------------------
 * gmsh v2.2 ASCII reader (host code, no device memory).  Replaces the token loop of
 * load_gmsh_v2 (src/mesh_gmsh.cpp:15-101): $Nodes records `id x y z`; $Elements records
 * `id type ntags tags... nodes...` with type 4 (Tet4), 2 (Tri3) and 1 (Line2) kept, the
 * first tag as the physical group, other types skipped.  The sections are split at line
 * boundaries and parsed on n_threads host threads (0: up to 16); every array is in file
 * order.  Fails with "Failed to open mesh file: <path>" like the reference, and on a
 * malformed record or a section count that differs from its records.
 * Element connectivity is returned as gmsh node IDs; edges are numbered by
 * efm_topology_create after the IDs are mapped to node positions.
 * ---------------------------------------------------------------------- */
typedef struct efm_gmsh efm_gmsh;

int efm_gmsh_read(const char *path, int n_threads, efm_gmsh **out);
/* h_sizes[4] = nodes, tets, tris, lines */
int efm_gmsh_sizes(const efm_gmsh *m, int64_t *h_sizes);
/* Host arrays sized from efm_gmsh_sizes; any pointer may be NULL to skip it.
 * node_ids [N], xyz [N][3], tet_nodes [T][4], tet_phys [T], tet_elem [T] (element IDs),
 * tri_nodes [F][3], tri_phys [F], tri_elem [F], line_nodes [L][2], line_phys [L]. */
int efm_gmsh_copy(const efm_gmsh *m, int64_t *node_ids, double *xyz, int64_t *tet_nodes,
                  int32_t *tet_phys, int64_t *tet_elem, int64_t *tri_nodes, int32_t *tri_phys,
                  int64_t *tri_elem, int64_t *line_nodes, int32_t *line_phys);
void efm_gmsh_free(efm_gmsh *m);

/* ------------------------------------------------------------------------
 * Touchstone writers (host code).  efm_touchstone_write replaces write_touchstone_nport
 * (src/io/touchstone.cpp:68-135): nfreq frequencies (Hz), S host c128 [nfreq][P][P]
 * row-major (S(i,j) of frequency f at (f*P + i)*P + j), format EFM_TOUCHSTONE_RI/_MA/_DB,
 * z0 the reference impedance on the option line.  The layout and the number formatting
 * (printf %.12g) are the reference's; more than 2 ports break a record after every 4th
 * pair.  Errors, in the reference's order and words: "Empty frequency or S-parameter
 * data", "Invalid number of ports: P" (P outside 1..99), "Cannot open file for writing:
 * <path>".  (Per-matrix shape and count checks belong to the caller's container.)
 * efm_touchstone_write_2port replaces the legacy write_touchstone (:50-62): S4 host c128
 * [nfreq][4] in the order s11, s21, s12, s22; like the reference it writes nothing and
 * reports nothing when the file cannot be opened.
 * efm_touchstone_extension replaces touchstone_extension (:137-145): ".s<P>p" into out.
 * ---------------------------------------------------------------------- */
enum { EFM_TOUCHSTONE_RI = 0, EFM_TOUCHSTONE_MA = 1, EFM_TOUCHSTONE_DB = 2 };

int efm_touchstone_write(const char *path, int64_t nfreq, const double *freq, int nports,
                         const double *S, int format, double z0);
int efm_touchstone_write_2port(const char *path, int64_t nfreq, const double *freq, const double *S4);
int efm_touchstone_extension(int num_ports, char *out, int cap);

/* ------------------------------------------------------------------------
 * Array-package writers (host code), the reference's file formats byte for byte (numbers
 * at printf %.12g).  Each fails with "Cannot open file for writing: <path>" like the
 * reference.  Host arrays only.
 * efm_array_package_json replaces export_array_package_json (src/io/array_export.cpp:33-123):
 *   positions [n_elements][3] m; freqs [n_freq] Hz; S [n_smat][nports][nports] c128 row-major,
 *   one matrix per frequency (n_smat <= n_freq; the reference reads frequencies[f] for
 *   matrix f); n_pattern_freqs = how many frequencies carry embedded patterns (the JSON holds
 *   only that summary); scan theta, phi (radians, written in degrees) and scan loss (dB).
 * efm_patterns_csv replaces export_patterns_csv (:125-145): n_patterns patterns, pattern p
 *   with port index port_index[p] and npts[p] rows of rows[][5] = (theta, phi0 magnitude,
 *   phi0 phase, phi90 magnitude, phi90 phase), angles in radians (the caller passes
 *   min(len(phi0 cut), len(phi90 cut)) points, the theta of the phi0 cut).
 * efm_coupling_csv replaces export_coupling_csv (:147-165) with extract_all_couplings
 *   (src/array/coupling_extract.cpp:21-57): every (i, j), i != j, of S [nports][nports] c128,
 *   |S| in dB (20 log10 max(|S|, 1e-20)), phase in degrees, element distance (positions
 *   [nports][3], or NULL: 0).
 * efm_scan_csv replaces export_scan_csv (:167-189) with active_vswr
 *   (src/array/active_impedance.cpp:117-123): scan point s at (theta[s], phi[s]) radians has
 *   nelem[s] elements; gamma, Z concatenated c128 per element.
 * ---------------------------------------------------------------------- */
int efm_array_package_json(const char *path, const char *model_name, double reference_impedance,
                           int64_t n_elements, const double *h_positions, int64_t n_freq,
                           const double *h_freqs, int64_t n_smat, int nports, const double *h_S,
                           int64_t n_pattern_freqs, int64_t n_scan, const double *h_scan_theta,
                           const double *h_scan_phi, const double *h_scan_loss_db);
int efm_patterns_csv(const char *path, int64_t n_patterns, const int32_t *h_port_index,
                     const int64_t *h_npts, const double *h_rows);
int efm_coupling_csv(const char *path, int nports, const double *h_S, const double *h_positions);
int efm_scan_csv(const char *path, int64_t n_scan, const double *h_theta, const double *h_phi,
                 const int64_t *h_nelem, const double *h_gamma, const double *h_Z);

#ifdef __cplusplus
}
#endif

#endif /* EDGEFEM_AMD_H */
