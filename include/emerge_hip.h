/*
 * emerge_hip.h -- C-ABI of the MI355X (gfx950) FEM hot path for EMerge.
 *
 * Plain C: opaque handles, plain pointers and sizes, int status codes
 * (0 = OK, < 0 = error; the message is read with emg_last_error).  No
 * exception crosses this boundary.  Host buffers are caller-owned and only
 * read/written during the call; device buffers belong to the context and are
 * released by emg_destroy.  One HIP stream per context; a context is not
 * re-entrant, distinct contexts may be used from distinct threads (the
 * Python binding releases the GIL in every call).
 *
 * Reference paths below are relative to the EMerge repository,
 *   E/   = src/emerge/_emerge/
 *   ASM/ = src/emerge/_emerge/physics/microwave/assembly/
 * The binding style follows the reference's ctypes PARDISO binding
 * (E/solve_interfaces/pardiso_interface.py:188-221, 406-457).
 */
#ifndef EMERGE_HIP_H
#define EMERGE_HIP_H

#include <stddef.h>
#include <stdint.h>

#ifdef __cplusplus
extern "C" {
#endif

typedef struct emg_ctx emg_ctx;

enum {
    EMG_OK = 0,
    EMG_ERR_ARG = -1,      /* invalid argument / inconsistent sizes        */
    EMG_ERR_STATE = -2,    /* call out of order (e.g. assemble before mesh) */
    EMG_ERR_HIP = -3,      /* HIP runtime error                            */
    EMG_ERR_NOMEM = -4,    /* device allocation failed                     */
    EMG_ERR_CONVERGE = -5  /* Krylov solve did not reach the tolerance     */
};

/* ---------------------------------------------------------------- context */

/* Create a context on HIP device `device` (its own stream).
 * Replaces the solver/assembler instance state the reference keeps per
 * object (ASM/assembler.py:105-108 cached_matrices; E/solver.py:541-549). */
int emg_create(int device, emg_ctx** out);
int emg_destroy(emg_ctx* ctx);
/* Copy the last error message of `ctx` (NUL-terminated, truncated to len). */
int emg_last_error(const emg_ctx* ctx, char* buf, size_t len);
/* Library version string "major.minor.patch". */
const char* emg_version(void);

/* Stream-ordered timing on the context stream (hipEvent pair). */
int emg_timer_start(emg_ctx* ctx);
int emg_timer_stop(emg_ctx* ctx, float* elapsed_ms);
int emg_synchronize(emg_ctx* ctx);

/* ------------------------------------------------------------------- mesh */

/* Upload the mesh: xyz is SoA (3 x n_nodes, row-major: all x, then y, then z),
 * tets (4 x n_tets) and tet_to_field (20 x n_tets) are row-major int32, i.e.
 * exactly Mesh3D.nodes / Mesh3D.tets / Nedelec2.tet_to_field
 * (E/mesh3d.py:343-360, E/elements/nedelec2.py:42-46) narrowed to int32.
 * n_field = 2*nE + 2*nF (E/elements/nedelec2.py:35). */
int emg_mesh_set(emg_ctx* ctx, int64_t n_nodes, const double* xyz,
                 int64_t n_tets, const int32_t* tets,
                 const int32_t* tet_to_field, int64_t n_field);

/* ---------------------------------------------------------------- pattern */

/* Build the global CSR sparsity pattern of the volume matrices and the
 * COO->CSR scatter map on the device.  The pattern equals what
 * scipy's coo_matrix((data,(rows,cols))).tocsr() yields for the COO layout of
 * ASM/curlcurl.py:459-465 (duplicates merged, columns sorted, explicit zeros
 * kept, int32 indices) -- the pattern of E and B of
 * tet_mass_stiffness_matrices (ASM/curlcurl.py:145-173).  Returns nnz. */
int emg_pattern_build(emg_ctx* ctx, int64_t* nnz);
/* Copy indptr (n_field + 1) and indices (nnz) to host buffers. */
int emg_pattern_get(emg_ctx* ctx, int32_t* indptr, int32_t* indices);
/* Copy the COO->CSR scatter map (400 * n_tets): entry t*400 + 20*i + j holds
 * the CSR slot of (tet_to_field[i,t], tet_to_field[j,t]). */
int emg_scatter_map_get(emg_ctx* ctx, int32_t* map);
/* Row-chunk statistics of the assembly schedule (for reports). */
int emg_pattern_stats(emg_ctx* ctx, int64_t* n_chunks, int64_t* n_incidences,
                      int64_t* max_rank);

/* --------------------------------------------------------------- assembly */

/* Upload per-tet isotropic material scalars, complex128 interleaved
 * (re, im) of length 2*n_tets: eps_r (after the loss update of
 * ASM/assembler.py:240) and inv_mu_r = 1/mu_r (E/mth/optimized.py:354-357). */
int emg_material_set(emg_ctx* ctx, const double* eps_r, const double* inv_mu_r);

/* General 3x3 complex material tensors per tet, complex128 interleaved in
 * the reference's (3,3,T) layout (component (i,j) of tet t at
 * 2*((3i+j)*T + t)): eps_r (after the loss update of ASM/assembler.py:240)
 * and mu_r; the device inverts mu_r as matinv (E/mth/optimized.py:342-370).
 * emg_assemble then evaluates dot(G_row, eps G_col) and
 * dot(C_row, mu^-1 C_col) as ned2_tet_stiff_mass (ASM/curlcurl.py:180-412)
 * does for anisotropic media and PML.  emg_material_set returns to the
 * isotropic path. */
int emg_material_set_tensor(emg_ctx* ctx, const double* eps_r, const double* mu_r);

/* Assemble E (curl-curl) and B (mass) CSR values on the device.
 * Replaces tet_mass_stiffness_matrices / _matrix_builder /
 * ned2_tet_stiff_mass (ASM/curlcurl.py:145-173, 419-466, 180-412) plus the
 * COO->CSR duplicate sum.  Stream-ordered, asynchronous. */
int emg_assemble(emg_ctx* ctx);
/* Copy the assembled values (complex128 interleaved, 2*nnz doubles each). */
int emg_values_get(emg_ctx* ctx, double* E_vals, double* B_vals);

/* ------------------------------------------------------- reduced system */

/* Keep only the DOFs in keep_ids (distinct; reduced index k = keep_ids[k],
 * so the order of keep_ids is the order of the reduced system): the PEC
 * elimination A[ix_(ids, ids)] of SolveRoutine.solve (E/solver.py:1266-1269)
 * with solve_ids from ASM/assembler.py:274-307, 406-407, optionally
 * permuted.  Builds the reduced CSR pattern (columns sorted per row) on the
 * device and gathers the assembled E and B values into it (FP64 and FP32
 * copies). */
int emg_system_reduce(emg_ctx* ctx, int64_t n_keep, const int32_t* keep_ids);
/* Gather the values of a new emg_assemble (new materials on the same mesh
 * and pattern) into the existing reduced system, keeping its pattern, the
 * surface terms, forcing vectors, port probes and the auxiliary gradient
 * space: the per-frequency re-assembly of frequency-dependent materials
 * (ASM/assembler.py:222-252, where conductivity enters eps as
 * -j cond / (omega eps0), :239-241).  A new mesh or pattern invalidates the
 * reduced system (EMG_ERR_STATE here until emg_system_reduce runs again). */
int emg_system_refresh(emg_ctx* ctx);
/* Copy the reduced pattern (indptr n_keep + 1, indices nnz); either may be
 * NULL.  nnz receives the reduced nonzero count. */
int emg_system_pattern_get(emg_ctx* ctx, int32_t* indptr, int32_t* indices, int64_t* nnz);

/* ------------------------------------------------------ surface / ports */

/* Add the boundary triangles of one Robin/port/absorbing boundary condition:
 * tri_vid (3 x n_tri row-major vertex ids), and for an excited port the
 * incident field sampled at the 6 Dunavant points of gaus_quad_tri(4)
 * (uinc, complex [n_tri][6][3], NULL otherwise).  The device computes per
 * triangle the 8x8 matrix of ned2_tri_stiff / ned2_tri_stiff_force without
 * its gamma factor and the forcing vector (ASM/robinbc.py:126-443).
 * Triangles of all BCs are concatenated; *tri_offset is this BC's first. */
int emg_bc_add(emg_ctx* ctx, int64_t n_tri, const int32_t* tri_vid, const double* uinc,
               int32_t* bc_id, int64_t* tri_offset);
/* Add the boundary triangles of a second-order absorbing boundary as one
 * more surface term: per triangle the 8x8 correction of _abc_order_2_terms
 * without its factor cf (ASM/robin_abc_order2.py:188-311); the solve scales
 * it by the gamma given for this term, 1j*c2/k0 (ASM/assembler.py:342-348).
 * The same triangles also carry the order-2 Robin term (emg_bc_add with
 * gamma = 1j*k0*c1). */
int emg_bc_add_abc2(emg_ctx* ctx, int64_t n_tri, const int32_t* tri_vid, int32_t* bc_id, int64_t* tri_offset);
/* Copy the per-triangle matrices (64 per triangle) and forcing vectors
 * (8 complex per triangle); either may be NULL. */
int emg_tri_terms_get(emg_ctx* ctx, double* mat, double* force);
/* Surface entries on the reduced pattern, CSR over reduced rows: s_ptr
 * (n_red + 1), per entry its column s_col and BC s_bc; entry k is the sum of
 * the triangle matrix values contrib[seg[k] .. seg[k+1]) (index
 * 64*triangle + 8*i + j), summed in the given order -- the Bempty /
 * generate_csr / K + B_p steps of ASM/assembler.py:314-350 with gamma kept
 * separate (K = E - k0^2 B + sum_bc gamma_bc S_bc). */
int emg_surface_set(emg_ctx* ctx, int64_t n_entries, const int32_t* s_ptr, const int32_t* s_col,
                    const int32_t* s_bc, const int32_t* seg, int64_t n_contrib, const int32_t* contrib);
int emg_surface_get(emg_ctx* ctx, double* s_val);
/* Dense reduced forcing vectors (port_vectors of ASM/assembler.py:267, 336):
 * vector p, reduced row r is element rows[k] = p*n_red + r, the sum of the
 * triangle forcing values contrib[seg[k] .. seg[k+1]) (index 8*triangle+i). */
int emg_forcing_set(emg_ctx* ctx, int32_t n_vectors, int64_t n_rows, const int32_t* rows,
                    const int32_t* seg, int64_t n_contrib, const int32_t* contrib);
int emg_forcing_get(emg_ctx* ctx, double* out);

/* ------------------------------------------------------------------ solve */

/* Solve A_b x_b = rhs_b for nbatch systems on the reduced pattern:
 *   A_b = E - k2[b] B + sum_bc gamma[b][bc] S_bc,  rhs_b = sum_p coef[b][p] f_p
 * (gamma complex [nbatch][n_bc], coef complex [nbatch][n_vectors]).
 * Replaces the per-frequency assemble + SolveRoutine.solve of run_sweep
 * (MW/microwave_3d.py:773-794; E/solver.py:1241-1326).  precision 0 = FP64
 * inner solves, 1 = complex64 inner solves inside FP64 iterative refinement
 * (with COCG, systems whose complex64 inner solves stall are finished with
 * complex128 inner solves from their iterate: emg_sweep_escalated), 2 (COCG
 * only) = complex128 inner solves whose SpMV reads the symmetric mass part
 * as complex64 inside the same FP64 refinement; the
 * method is emg_solver_config's; fblock = max systems per batch (1..32, up
 * to 64 with COCG; the matrix is read once per batch).  x_out complex [nbatch][n_red]; relres_out =
 * final FP64 ||b - A x|| / ||b||. */
int emg_sweep_solve(emg_ctx* ctx, int64_t nbatch, const double* k2, const double* gamma,
                    const double* rhs_coef, int precision, int fblock, double tol, int restart,
                    int max_cycles, double* x_out, double* relres_out, int64_t* iters_out,
                    double* overlap_out);

/* --------------------------------------------------------- S-parameters */

/* Register the port faces used for S-parameter overlaps: per probe triangle
 * (ports concatenated, tri_counts[p] each) the tet that owns it, the 6
 * Dunavant points of gaus_quad_tri(4) ([n][6][3]), the port's mode shape at
 * those points (complex [n][6][3], without the k0-dependent factor) and the
 * weight x area products ([n][6]).  With overlap_out non-NULL,
 * emg_sweep_solve then writes per system and port
 *   O = sum_tri sum_q w A (E(p_q) . conj(shape(p_q)))  (complex [nbatch][n_ports])
 * with E interpolated as ned2_tet_interp (E/elements/ned2_interp.py:141-301);
 * the S-matrix follows from O and the norms below as in _compute_s_data /
 * sparam_field_power / sparam_mode_power (MW/microwave_3d.py:1250-1303,
 * MW/sparam.py:71-137). */
int emg_port_probe_set(emg_ctx* ctx, int32_t n_ports, const int64_t* tri_counts,
                       const int32_t* owner_tet, const double* pts, const double* shape,
                       const double* wa);
/* General probes: per port point_counts[p] points (ports concatenated), each
 * with its owner tet, coordinates ([n][3]), complex vector c ([n][3]) and
 * weight w ([n]); emg_sweep_solve then writes O = sum w (E(p) . conj(c)) per
 * system and port.  Voltage-line ports use the points of their integration
 * lines with c = the line step and w = the composite-rule weight / number of
 * lines, so O is the port voltage of Line.line_integral averaged over the
 * lines (MW/microwave_3d.py:1271-1292, E/coord.py:24-61, 149-153).  Replaces
 * any earlier probe set (emg_port_probe_set is its triangle form). */
int emg_port_points_set(emg_ctx* ctx, int32_t n_ports, const int64_t* point_counts, const int32_t* owner_tet,
                        const double* pts, const double* vec, const double* w);
/* k0-independent norms N_p = sum w |c|^2 (n_ports doubles; for mode-overlap
 * ports sum w A |shape|^2). */
int emg_port_norms_get(emg_ctx* ctx, double* norms);

/* Krylov method for emg_sweep_solve / emg_csr_solve: 0 = GMRES(restart)
 * with CGS2 (default), 1 = IDR(idr_s) with bi-orthogonalisation (then
 * `restart` is the matvec budget of one refinement cycle), 2 = COCG on the
 * complex-symmetric part (A + A^T)/2 inside FP64 refinement on A (then
 * `restart` caps the inner iterations of one cycle; Jacobi preconditioning,
 * plus the gradient-space correction of emg_aux_set when set; each
 * correction is scaled to minimise the FP64 residual); inner_rel32 = the
 * residual reduction an inner solve aims for before the FP64 residual is
 * recomputed.  The reference's iterative solvers (scipy gmres/bicgstab/
 * gcrotmk, E/solver.py:458-527) have no counterpart knob; this selects the
 * device algorithm. */
int emg_solver_config(emg_ctx* ctx, int method, int idr_s, double inner_rel32);

/* Number of systems of the last emg_sweep_solve (precision 1, COCG) whose
 * complex64 inner solves stalled and that were finished with complex128
 * inner solves from their current iterate.  No reference counterpart (the
 * reference solves in FP64 throughout). */
int emg_sweep_escalated(emg_ctx* ctx, int64_t* n);
/* on = 0 turns that continuation off (default on); used by the host's
 * block-width autotuner, whose timing probes stop after a few iterations. */
int emg_solver_set_escalation(emg_ctx* ctx, int on);
/* on = 1: the complex64 COCG path keeps a real reduced stiffness in FP64 (8
 * bytes per entry, as complex64) and accumulates its SpMV and the
 * gradient-space correction in FP64; complex64 vectors otherwise.  For
 * meshes with k0^2 h^2 << 1, where E z cancels to O(k0^2 h^2) of its terms
 * on near-gradient z and FP32 stiffness values would swamp the result.  No
 * reference counterpart (the reference solves in FP64). */
int emg_solver_set_safe32(emg_ctx* ctx, int on);
/* Lane layout of the COCG SpMV for the following solves: entries of a row in
 * flight (4, 8, 16 or 32; 0 = the default for the batch width: 8 at 32
 * systems, 16 below).  Fewer entries in flight put more rows in a
 * wavefront.  A value the batch width cannot use (fewer entries than
 * systems per lane, more than lanes per wavefront) runs the default.  Used
 * by the host's autotuner; no reference counterpart. */
int emg_solver_set_spmv_lanes(emg_ctx* ctx, int entries_in_flight);
/* Row segments of the reduced system for the following COCG solves: rows
 * [row_start[s], row_start[s+1]) form segment s (row_start[0] = 0,
 * row_start[nseg] = the reduced size, increasing).  For a block-diagonal
 * system of independent meshes placed side by side: each segment runs its
 * own Krylov recurrence (its own scalars, stop test and refinement on its
 * own right-hand side) inside the same kernel launches, so every mesh
 * converges as it would alone; relres and iterations are reported per system
 * as those of its slowest segment.  nseg <= 1 clears them, and so does
 * emg_system_reduce; they apply to the reduced system only (not to
 * emg_csr_solve).  At most 256 segments; COCG only.  Replaces the reference's one solve per job
 * (MW/microwave_3d.py:773-861, sequential, threaded or process-pool). */
int emg_solver_set_segments(emg_ctx* ctx, int32_t nseg, const int64_t* row_start);

/* Mesh topology on the device (replaces Mesh3D._pre_update,
 * E/mesh3d.py:386-498): unique edges / triangles of the tets (sorted vertex
 * tuples in lexicographic order) and the maps tet_to_edge (6 x T, local
 * pairs (0,1),(0,2),(0,3),(1,2),(3,1),(2,3)), tet_to_tri (4 x T, local
 * triples (0,1,2),(0,2,3),(0,3,1),(1,2,3)), tri_to_edge (3 x nF),
 * tri_to_tet (2 x nF, -1234 for a missing second tet) and edge lengths,
 * bit-exact with the reference.  xyz_soa (3 x V), tets (4 x T) int32
 * row-major; V < 2^21. */
int emg_topology_build(emg_ctx* ctx, int64_t V, const double* xyz_soa, int64_t T, const int32_t* tets,
                       int64_t* n_edges, int64_t* n_tris);
/* Copies the built topology out (row-major int32 arrays as above; any
 * pointer may be null to skip that array). */
int emg_topology_get(emg_ctx* ctx, int32_t* edges, int32_t* tris, int32_t* tet_to_edge, int32_t* tet_to_tri,
                     int32_t* tri_to_edge, int32_t* tri_to_tet, double* edge_lengths);

/* DOF maps of the second-order Nedelec space from the topology of the last
 * emg_topology_build (Nedelec2.__init__, E/elements/nedelec2.py:42-58):
 * tet_to_field [20][T], tri_to_field [8][nF], edge_to_field [2][nE], int32
 * row-major, blocks [edge 1st | face 1st | edge 2nd | face 2nd].  Any
 * output may be NULL. */
int emg_dof_maps(emg_ctx* ctx, int32_t* tet_to_field, int32_t* tri_to_field, int32_t* edge_to_field);
/* DOFs kept by the PEC elimination (ASM/assembler.py:274-307, 406-407) on
 * that topology: every DOF except both DOFs of each edge of the n_pec PEC
 * triangles and the 8 DOFs of each of those triangles, ascending.  *n_keep
 * receives the count; keep_ids (capacity 2 (nE + nF)) may be NULL to ask
 * for the count only. */
int emg_pec_keep(emg_ctx* ctx, int64_t n_pec, const int32_t* pec_tris, int64_t* n_keep, int32_t* keep_ids);

/* Auxiliary gradient space of the COCG preconditioner (method 2 of
 * emg_sweep_solve): the inner solves apply z = D^-1 r + G D_g^-1 G^T r with
 * D the diagonal of the operator, G (n_red x n_nodal, real CSR g_*, its
 * transpose gt_*) the discrete gradient of the second-order nodal functions
 * in the reduced Nedelec-2 rows, and per system
 * D_g = -k0^2 dvol + sum_s gamma_s dsurf[s] (dvol complex n_nodal, dsurf
 * real [n_surf][n_nodal], one row per surface term of emg_surface_set).
 * A Maxwell (Hiptmair-type) gradient correction: Jacobi alone leaves the
 * curl-free near-null space of E - k0^2 B unresolved.  No counterpart in
 * the reference, whose default path is a direct solve (E/solver.py:
 * 535-773; its iterative path uses ILU, :437-527).  n_nodal = 0 turns it
 * off; a new reduced system (emg_system_reduce) clears it. */
int emg_aux_set(emg_ctx* ctx, int64_t n_nodal, const int32_t* g_ptr, const int32_t* g_idx, const double* g_val,
                const int32_t* gt_ptr, const int32_t* gt_idx, const double* gt_val, const double* dvol,
                int n_surf, const double* dsurf);
/* Per-function terms of that auxiliary space (after emg_aux_set): the
 * diagonal becomes D_g = dconst - k0^2 dvol + sum_s gamma_s dsurf[s] and the
 * correction G W D_g^-1 G^T with W = diag(weight) (dconst complex n_nodal,
 * weight real n_nodal; either may be NULL: 0 and 1).  Lets the space hold, next to the nodal
 * gradients, the vector-nodal functions lam_a e_d interpolated into the
 * Nedelec-2 space, whose diagonal has the frequency-independent curl-curl
 * part dconst (auxiliary-space preconditioning of H(curl), Hiptmair-Xu).
 * No counterpart in the reference (E/solver.py:437-527 is its iterative
 * path with ILU); emg_aux_set clears the terms. */
int emg_aux_set_terms(emg_ctx* ctx, const double* dconst, const double* weight);

/* Dense element matrices of every tet of the context's mesh (isotropic
 * materials of emg_material_set): the 20x20 curl-curl K and mass M of
 * ned2_tet_stiff_mass (ASM/curlcurl.py:180-412), batched.  Layout
 * K[e][t], M[e][t] complex128 with e = 20 u + v over the canonical DOFs u, v
 * (tet vertices ranked by global id; DOF order edges (01,02,03,12,13,23),
 * faces (012,013,023,123), then the second kind of each); K and M may be
 * null to keep the result on the device.  use_mfma = 1 contracts the tets'
 * Gram vectors with the dense tables on FP64 matrix cores, 0 with the
 * compile-time sparse tables on the vector units. */
int emg_element_matrices(emg_ctx* ctx, int use_mfma, double* K, double* M);

/* Generic complex CSR system for the Solver plug-in (replaces
 * Solver.solve(A, b, ...), E/solver.py:365-367, next to SuperLU / PARDISO /
 * cuDSS, E/solver.py:535-773): vals complex128 interleaved. */
int emg_csr_set(emg_ctx* ctx, int64_t n, const int32_t* indptr, const int32_t* indices,
                const double* vals);
/* Auxiliary gradient space for the generic matrix (COCG): G (n x n_nodal,
 * real CSR g_*, rows = the matrix rows) and its transpose (gt_*, each row
 * sorted), the discrete gradient of the nodal functions whose gradients lie
 * in the matrix's DOF space.  D_g = diag(G^T A G) is formed on the device
 * from the current matrix before a solve (for E - k0^2 B + sum gamma S it
 * equals the closed form emg_aux_set takes, E g = 0).  Lets the Solver
 * plug-in precondition Maxwell systems as the batched sweep does; no
 * counterpart in the reference (its iterative path uses ILU, E/solver.py:
 * 437-527).  n_nodal = 0 turns it off; a matrix of another size drops it. */
int emg_csr_aux_set(emg_ctx* ctx, int64_t n_nodal, const int32_t* g_ptr, const int32_t* g_idx, const double* g_val,
                    const int32_t* gt_ptr, const int32_t* gt_idx, const double* gt_val);
/* nrhs right-hand sides b (complex [nrhs][n]) -> x (complex [nrhs][n]).
 * Returns EMG_ERR_CONVERGE (x still written) if a system misses tol. */
int emg_csr_solve(emg_ctx* ctx, int32_t nrhs, const double* b, int precision, double tol,
                  int restart, int max_cycles, double* x, double* relres_out, int64_t* iters_out);

/* E and curl E of a solution at points (the field sampling of
 * MWField.interpolate, MW/microwave_data.py:402-430: ned2_tet_interp and
 * ned2_tet_interp_curl, E/elements/ned2_interp.py:141-507) on the context's
 * mesh: field = complex128 [n_field] (all DOFs, eliminated ones 0), pts =
 * real [3][P].  Each point belongs to the last tet that contains it
 * (tolerances 1e-8 on the sum, 1e-6 per barycentric coordinate); E and curlE
 * (complex128 [3][P]) are NaN and owner -1 where no tet does.  H follows as
 * curlE / (-j w mu0 mu_r[owner]).  Host buffers; synchronous. */
int emg_field_points(emg_ctx* ctx, const double* field, int64_t P, const double* pts, double* E, double* curlE,
                     int32_t* owner);

/* Element matrices of the boundary-mode eigenproblem of a port face
 * (replaces generalized_matrix_GQ + _matrix_builder,
 * ASM/generalized_eigen_hb.py:247-458, called from assemble_bma_matrices,
 * ASM/assembler.py:110-186): n triangles with local 2-D vertex coordinates
 * xy (real [n][3][2], vertices in increasing global id), eps_r and mu_r^-1
 * (complex128 [n][3][3], row-major), wavenumber k0.  A, B: complex128
 * [n][14][14] in the mixed Nedelec-2 / Lagrange-2 order of
 * E/elements/nedleg2.py:138-146 (edge x3, face, edge x3, face, vertex x3,
 * edge x3), A = [Att - k0^2 Btt, 0; 0, 0], B = [Dtt, Dzt^T; Dzt, Dzz1 -
 * k0^2 Dzz2] times |area| (the face-face block of Dtt is zero as in the
 * reference).  Host buffers; synchronous. */
int emg_port_mode_elements(emg_ctx* ctx, int64_t n, const double* xy, const double* eps, const double* inv_mu,
                           double k0, double* A, double* B);

/* Stratton-Chu far field (replaces stratton_chu_ff, MW/sc.py:41-132, called
 * by stratton_chu, :135-170): M surface samples with fields E, H (complex128
 * [3][M], interleaved), positions pos and area-weighted normals wn (real
 * [3][M]); N directions theta, phi (radians); wavenumber k0.  Samples with
 * |E| <= 1e-3 max|E| are left out.  Eout, Hout: complex128 [3][N],
 * Eout = -j k0 / (4 pi) r x sum_j [(n x E)_j - Z0 r x (n x H)_j]
 * exp(j k0 r . x_j), Hout = r x Eout / Z0.  Host buffers; synchronous. */
int emg_farfield(emg_ctx* ctx, int64_t M, const double* E, const double* H, const double* pos, const double* wn,
                 int64_t N, const double* theta, const double* phi, double k0, double* Eout, double* Hout);

/* emg_farfield without any allocation (no counterpart in the reference):
 * for callers that cannot allocate device memory (code under HIP graph
 * capture, tight serving loops, harnesses that forbid hipMalloc).  Every
 * pointer is a caller-owned DEVICE buffer with the layout of emg_farfield;
 * the work is enqueued on the context's stream and the call returns without
 * waiting (emg_synchronize waits).  Its scratch (per-sample magnitudes and
 * records, per-chunk partial sums: 17 M + 12 N * chunks doubles, chunks as
 * emg_farfield's launch plan, at most 8 per 256 directions) lives in one
 * 64 MiB device array per GPU, held until the process ends; larger problems
 * are rejected with EMG_ERR_ARG.  Calls that overlap on one GPU (any
 * context) share that array and race.  Results equal emg_farfield's bit for
 * bit (same kernels, same launch plan). */
int emg_farfield_noalloc(emg_ctx* ctx, int64_t M, const double* E, const double* H, const double* pos,
                         const double* wn, int64_t N, const double* theta, const double* phi, double k0,
                         double* Eout, double* Hout);

#ifdef __cplusplus
}
#endif

#endif /* EMERGE_HIP_H */
