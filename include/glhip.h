/*
 * glhip.h -- C ABI of the MI355X (gfx950) Ginzburg-Landau Newton-step library.
 *
 * The reference femgl has no plugin/FFI layer: its hot path is the private
 * step methods of class FemGL<dim> (femgl/inc/femgl.h:121-349), operating on
 * member state (Trilinos CSR system_matrix, system_rhs, residual_vector,
 * local_solution, ... femgl.h:310-316).  This header is the C boundary that
 * replaces those methods; each entry point names the reference method it
 * stands in for.  All functions return 0 on success and a negative status
 * on error (message via glh_last_error); nothing throws across the ABI.
 * Host arrays are caller-owned and copied; device buffers belong to the
 * context.  One host thread and one HIP stream per context.
 *
 * Degrees of freedom are numbered node-major: dof = node * 18 + c with
 * c < 9 the real part u_{c/3, c%3} and c >= 9 the imaginary part
 * v_{(c-9)/3, (c-9)%3} of the 3x3 complex order parameter A = u + i v
 * (femgl/src/s_vector2matrix.cc:159-160).  The Jacobian is a BSR matrix of
 * 18x18 blocks over mesh nodes.
 */
#ifndef GLHIP_H
#define GLHIP_H

#include <stddef.h>
#include <stdint.h>

#ifdef __cplusplus
extern "C" {
#endif

typedef struct glh_ctx glh_ctx;
typedef struct glh_mesh glh_mesh;
typedef struct glh_forest glh_forest;

#define GLH_OK 0
#define GLH_ERR_ARG -1
#define GLH_ERR_HIP -2
#define GLH_ERR_STATE -3
#define GLH_ERR_NOCONV -4     /* SolverControl::NoConvergence (solve.cc:159-174) */
#define GLH_ERR_SINGULAR -5
#define GLH_ERR_RANGE -6      /* pressure outside matep's table (matep.cc:365-406 UB) */

/* assembly flags */
#define GLH_ASM_J 1           /* Jacobian (assemble.cc:177-280) */
#define GLH_ASM_R 2           /* right-hand side -R (assemble.cc:257-276) */
#define GLH_ASM_FP32 4        /* Jacobian stored and solved in FP32 */
#define GLH_ASM_MATRIX_FREE 8 /* Jacobian in matrix-free form (see glh_set_matrix_free); with FP32 its state is FP32 */

/* ---------------------------------------------------------------- context */
/* FemGL<dim> constructor state (femgl.cc:107-203): device + stream. */
int glh_create(int device, glh_ctx **out);
int glh_destroy(glh_ctx *ctx);
const char *glh_last_error(const glh_ctx *ctx);

/* Material and gradient constants for every cell: alpha, beta1..5 from matep
 * (femgl.cc:161-168), K1..K3 (femgl.h:320-322) and the AdGR diffuse length
 * bt (femgl.cc:150, Robin term active iff bt < 1e10, assemble.cc:293). */
int glh_set_material(glh_ctx *ctx, double alpha, const double beta[5], const double K[3], double bt);

/* Per-instance materials of a batch (see glh_mesh_concat): coef[n_inst*10]
 * = (alpha, beta1..5, K1, K2, K3, bt) per instance, the constructor state of
 * one FemGL<dim> per (T, p) point (femgl.cc:145-168).  n_inst must equal the
 * number of instances of the uploaded mesh. */
int glh_set_materials(glh_ctx *ctx, int n_inst, const double *coef);

/* matep restatement (matep/src/matep.cc): out[0]=alpha, [1..5]=beta1..5,
 * [6]=beta_A, [7]=beta_B, [8]=gap_A, [9]=gap_B, [10]=f_A, [11]=f_B.
 * Returns GLH_ERR_RANGE when SCC is on and p is outside [0, 34). */
int glh_matep(double p, double t, int scc, double out[12]);

/* ------------------------------------------------------- host mesh layer */
/* Axis-aligned box of nx*ny*nz hex cells of degree 1 or 2 (the
 * GridGenerator::hyper_rectangle + refine_global of femgl/src/makegrid_*.cc).
 * side_id[6] for faces -x,+x,-y,+y,-z,+z: 0 = homogeneous Dirichlet on all
 * 18 components, 1 = natural, 2/3/4 = AdGR wall (x/y/z-normal column
 * Dirichlet, femgl.h:294-301, plus the Robin term when bt < 1e10). */
int glh_mesh_box(int degree, int nx, int ny, int nz, const double lo[3], const double hi[3],
                 const int side_id[6], glh_mesh **out);
/* Batch of independent instances: the meshes are concatenated with disjoint
 * node and cell ranges (instance i owns nodes [node_off[i], node_off[i+1])).
 * All meshes need the same degree.  The solver treats every instance as its
 * own problem (own Krylov space, line search and stopping test) while each
 * vector operation runs as one launch over the whole batch. */
int glh_mesh_concat(const glh_mesh *const *meshes, int n, glh_mesh **out);
int glh_mesh_destroy(glh_mesh *m);
/* sizes: [0]=degree [1]=n_cells [2]=n_nodes [3]=n_faces [4]=nnzb [5]=n_colors
 *        [6]=n_instances [7]=0 */
int glh_mesh_sizes(const glh_mesh *m, int64_t sizes[8]);
/* Copies the mesh arrays out (any pointer may be NULL):
 * cell_nodes[n_cells*ns] (lexicographic local order), cell_h[n_cells*3],
 * node_xyz[n_nodes*3], faces[n_faces*3] (cell, local face 0..5, id),
 * node_mask[n_nodes] (bit c set = component c Dirichlet-constrained),
 * color_ptr[n_colors+1], color_cells[n_cells], row_ptr[n_nodes+1], col[nnzb]. */
int glh_mesh_get(const glh_mesh *m, int *cell_nodes, double *cell_h, double *node_xyz, int *faces,
                 uint32_t *node_mask, int *color_ptr, int *color_cells, int *row_ptr, int *col);

/* Mesh from the caller's own per-cycle arrays, for a caller that keeps the
 * mesh, the DoF numbering and the constraints in deal.II: the data
 * setup_system() builds (setup_uniform_B-phase.cc:125-198, the
 * AffineConstraints of femgl.h:307-308 that assemble.cc:356-361 applies),
 * passed as plain arrays.  Nodes are the FE_Q support points; node v owns
 * dofs 18 v + c in the layout above, in the caller's node numbering.
 *   degree        1 or 2; ns = (degree+1)^3 nodes per cell
 *   local_order   GLH_ORDER_LEXICOGRAPHIC (x fastest) or GLH_ORDER_HIERARCHIC
 *                 (FE_Q's support-point order: vertices, then the midpoints
 *                 of the lines, faces and cell in GeometryInfo<3> numbering)
 *   cell_nodes    [n_cells*ns] node of each local support point
 *   cell_lo/h     [n_cells*3] lower corner and edge lengths of each box cell
 *   node_xyz      [n_nodes*3] or NULL; when given, every cell must place its
 *                 nodes there (checks the local order)
 *   faces         [n_faces*3] (cell, local face 0..5 = -x +x -y +y -z +z,
 *                 boundary id): 0 = A = 0 wall, 2/3/4 = AdGR wall (normal
 *                 column Dirichlet + Robin term), others natural (periodic
 *                 faces go through constraint lines)
 *   node_mask     [n_nodes] Dirichlet component bits, or NULL to derive them
 *                 from the faces' ids (femgl.h:294-301)
 *   lines         n_lines node constraint lines U[line_node[l]] = sum_k
 *                 line_w[k] U[line_master[k]], k in [line_ptr[l],
 *                 line_ptr[l+1]), shared by the 18 components (hanging nodes,
 *                 periodic images); closed by the library.  A constrained
 *                 node's line takes precedence over its Dirichlet mask, as in
 *                 the reference, which adds the hanging-node lines first.
 * The library derives the colours and the BSR patterns (the condensed one of
 * make_sparsity_pattern with constraints and the cell-loop one).  Upload the
 * result with glh_upload_mesh. */
#define GLH_ORDER_LEXICOGRAPHIC 0
#define GLH_ORDER_HIERARCHIC 1
typedef struct glh_mesh_desc {
  int degree, n_cells, n_nodes, local_order;
  const int *cell_nodes;
  const double *cell_lo, *cell_h, *node_xyz;
  int n_faces;
  const int *faces;
  const uint32_t *node_mask;
  int n_lines;
  const int *line_node, *line_ptr, *line_master;
  const double *line_w;
} glh_mesh_desc;
int glh_mesh_from_arrays(const glh_mesh_desc *desc, glh_mesh **out);
/* Message of the last failed glh_mesh_box / glh_mesh_from_arrays call on the
 * calling thread (the mesh functions have no context to hold it). */
const char *glh_mesh_last_error(void);

/* node_off[n_inst+1], cell_off[n_inst+1] (either may be NULL) */
int glh_mesh_instances(const glh_mesh *m, int *node_off, int *cell_off);

/* Hanging-node data of a mesh (adaptive meshes; zero for conforming ones):
 * sizes[0] = hanging nodes, [1] = constraint-line entries, [2] = blocks of
 * the pattern the cell loop assembles into. */
int glh_mesh_constraint_sizes(const glh_mesh *m, int64_t sizes[3]);
/* Closed hanging-node lines U[h] = sum_k hang_w[k] U[hang_m[k]],
 * k in [hang_ptr[h], hang_ptr[h+1]) (hang_ptr[n_nodes+1]), and the cell-loop
 * pattern arow_ptr[n_nodes+1], acol[blocks].  The system pattern of
 * glh_mesh_get is the condensed one (make_sparsity_pattern with constraints,
 * keep_constrained_dofs = false, setup_uniform_B-phase.cc:190-196). */
int glh_mesh_constraints(const glh_mesh *m, int *hang_ptr, int *hang_m, double *hang_w, int *arow_ptr,
                         int *acol);

/* ------------------------------------------------ adaptive meshes (Q1) */
/* Forest of octrees over nx*ny*nz root cells of the box [lo, hi] with the
 * side ids of glh_mesh_box: the parallel::distributed::Triangulation of the
 * reference's adaptive cycle (femgl/src/refine.cc:128-173). */
int glh_forest_create(int nx, int ny, int nz, const double lo[3], const double hi[3], const int side_id[6],
                      glh_forest **out);
int glh_forest_destroy(glh_forest *f);
int64_t glh_forest_n_cells(const glh_forest *f);
/* New forest (execute_coarsening_and_refinement with 2:1 balance,
 * refine.cc:160-165): flags[e] = 2 marks cell e for coarsening, 1 for
 * refinement, 0 neither.  Complete families of 8 cells all marked 2 become
 * their parent where the 2:1 balance allows; then the cells marked 1 split
 * into 8 and neighbours (faces, edges, vertices) more than one level apart
 * are refined.  f is unchanged. */
int glh_forest_refine(const glh_forest *f, const uint8_t *flags, glh_forest **out);
/* Element degree (1 or 2, default 1) of the meshes glh_forest_mesh builds
 * from f and from the forests refined from it (FE_Q(degree) of
 * femgl/src/femgl.cc:112). */
int glh_forest_set_degree(glh_forest *f, int degree);
/* Q1 or Q2 mesh of the leaves (cell e = leaf e, Morton order) with closed
 * hanging-node constraints (DoFTools::make_hanging_node_constraints,
 * AffineConstraints::close; setup_uniform_B-phase.cc:137-160): a node on a
 * face or edge of a coarser neighbour follows that cell's Q_k field there
 * (Q1: 1/2, 1/4; Q2: products of 3/8, 3/4, -1/8 under 2:1 balance). */
int glh_forest_mesh(const glh_forest *f, glh_mesh **out);
/* SolutionTransfer::interpolate + constraints_solution.distribute
 * (refine.cc:163-170): Un on the mesh mn of fn from Uo on the mesh mo of fo,
 * where fn refines fo. */
int glh_forest_transfer(const glh_forest *fo, const glh_mesh *mo, const double *Uo, const glh_forest *fn,
                        const glh_mesh *mn, double *Un);
/* The same transfer on the device (refine.cc:163-175 without the host round
 * trip of the field): ctx holds mo and its solution; uploads mn
 * (glh_upload_mesh), interpolates on the GPU and distributes the
 * constraints, which leaves the state glh_set_solution would leave with
 * glh_forest_transfer's field, bit for bit.  Single-instance meshes.  On an
 * error after the upload the context holds mn with a zero solution. */
int glh_transfer_solution(glh_ctx *ctx, const glh_forest *fo, const glh_mesh *mo, const glh_forest *fn,
                          const glh_mesh *mn);
/* KellyErrorEstimator<3>::estimate (refine.cc:146-151) of the Q1 field U:
 * eta[n_cells] in float, as the reference's Vector<float>. */
int glh_kelly(const glh_forest *f, const glh_mesh *m, const double *U, float *eta);
/* refine_and_coarsen_fixed_number(eta, frac, 0) (refine.cc:153-157): flags
 * the floor(frac * n) cells of largest eta (ties in cell order). */
int glh_flag_fixed_number(int64_t n, const float *eta, double frac, uint8_t *flags);
/* refine_and_coarsen_fixed_number (refine.cc:153-157) with a coarsening
 * fraction: flags 1 for the floor(refine_frac n) cells of largest indicator,
 * 2 for the floor(coarsen_frac n) of smallest, 0 otherwise (the input of
 * glh_forest_refine); refine_frac + coarsen_frac <= 1. */
int glh_flag_refine_coarsen(int64_t n, const float *eta, double refine_frac, double coarsen_frac, uint8_t *flags);

/* FemGL::run (femgl/src/run.cc:109-260) for adaptive Q1 meshes: cycle 0
 * solves on the mesh of f0 from U0 (n_nodes*18 of that mesh); each of the
 * n_refine further cycles flags refine_frac of the cells by Kelly, refines
 * (2:1 balance), transfers the converged field and solves again.  Each
 * solve is glh_newton_solve with the given controls (fp32: mixed
 * precision).  check_fp64 (with fp32): every cycle is also solved in FP64
 * from the same start.  cycle_stats[(n_refine+1) * GLH_CYCLE_STATS]:
 * [0] cycle [1] cells [2] dofs [3] hanging nodes [4] Newton steps
 * [5] GMRES iterations [6] ||R0|| [7] ||R|| [8] F [9] converged
 * [10] solve seconds [11] cycle seconds (mesh, solve, Kelly, refinement,
 * transfer) [12] F of the FP64 solve [13] |F - F64| / |F64|
 * [14] max|U - U64| / max|U64| [15] FP64 converged (NaN without check).
 * On return ctx holds the last mesh and its solution; *f_out / *m_out (may
 * be NULL) receive the last forest and mesh. */
#define GLH_CYCLE_STATS 16
int glh_run_adaptive(glh_ctx *ctx, const glh_forest *f0, const double *U0, int n_refine, double refine_frac,
                     double rtol, int max_newton, double lin_tol, int lin_maxit, int restart, double ls_step,
                     int damped, int fp32, int check_fp64, double *cycle_stats, glh_forest **f_out,
                     glh_mesh **m_out);

/* FemGL::output_results (femgl/src/io.cc:105-168): dir/solution_NN.0.vtu
 * and dir/solution_NN.pvtu (NN = counter, two digits) with point data
 * du_ij, dv_ij (dU, may be NULL) and u_ij, v_ij (U), cell data "subdomain".
 * Q2 cells are written as 8 sub-hexahedra. */
int glh_write_vtu(const glh_mesh *m, const double *U, const double *dU, const char *dir, int counter);

/* ------------------------------------------------ .prm configuration */
/* The reference's parameter file (deal.II ParameterHandler text: nested
 * "subsection NAME" ... "end", "set KEY = VALUE", '#' comments) with the keys,
 * defaults and patterns of confreader/src/declare.cc:110-300; undeclared keys
 * and pattern violations are errors (message in err).  The subsection
 * "glhip" holds what the reference fixes at compile time or in its driver:
 * geometry (retangle | cube), boundary ids (six ids, faces -x +x -y +y -z +z;
 * 5 5 = periodic axis), finite element degree (1 | 2), initial guess (B | A |
 * BnA | BinA | BnB | PDW), random seed, noise amplitude, mixed precision,
 * GMRES restart, output directory (VTU files per Newton iteration in
 * DIR/refine-cycle_<cycle>/ as run.cc:219-224 writes them; empty: none) and
 * verbose (progress lines of run() on stdout). */
typedef struct glh_prm glh_prm;
int glh_prm_parse(const char *text, glh_prm **out, char *err, int err_len);
int glh_prm_destroy(glh_prm *p);
/* value as text / as a number (double and integer entries) */
int glh_prm_get(const glh_prm *p, const char *subsection, const char *key, char *value, int len);
int glh_prm_get_double(const glh_prm *p, const char *subsection, const char *key, double *v);
/* FemGL::run (femgl/src/run.cc:109-260) driven by a parsed configuration:
 * make_grid (box refined globally "Number of initial global refinments"
 * times), then per cycle up to "Number of interations" + 1 rounds of
 * assemble / solve ("Cycle k linear solver tol") / damped Newton step, leaving
 * the cycle when |R - R_last| < "Cycle k refinement threshold" (refine) or
 * ||R|| <= "converge accuracy" (stop), with adaptive (Kelly, "adaptive
 * refinment ratio") or global refinement between cycles.  rows[max_rows][6]:
 * cycle, cells, dofs, Newton iterations, final ||R||, free energy; *n_rows
 * the number written.  GLH_ERR_NOCONV when a linear solve fails, as the
 * reference's SolverControl::NoConvergence ends its run.  Sets the context's
 * matrix-free switch (glh_set_matrix_free) from the glhip key "matrix free". */
int glh_run_prm(glh_ctx *ctx, const glh_prm *p, int max_rows, double *rows, int *n_rows);

/* ------------------------------------------- coordinate-MLP warm start */
/* A small network that predicts the order parameter from (x, y, z, T/Tc, p,
 * D), trained on converged solutions and used as the Newton initial guess
 * (an addition beside the reference's seeded guesses; it has no
 * counterpart there).  Inputs pass through Fourier features
 * [sin, cos](2 pi x B), B a fixed 6 x n_freq matrix with N(0, sigma^2)
 * entries; `depth` hidden layers of `width` units with SiLU; n_out linear
 * outputs.  Training and inference run in hand-written BF16 kernels with
 * FP32 accumulation (matrix cores); master weights and Adam moments are FP32.
 * Constraints: width a multiple of 64, 2*n_freq a multiple of 64,
 * 1 <= depth <= 7, n_out <= 64.  Inputs X are n x 6, targets and predictions n x n_out,
 * row-major FP32 on the host. */
typedef struct glh_mlp glh_mlp;
int glh_mlp_create(int device, int n_freq, double sigma, int width, int depth, int n_out, uint64_t seed,
                   glh_mlp **out);
int glh_mlp_destroy(glh_mlp *m);
const char *glh_mlp_last_error(const glh_mlp *m);
int64_t glh_mlp_n_params(const glh_mlp *m);
/* flat FP32 parameters: for layer l, W_l (K_l x N_l row-major, y = x W + b)
 * then b_l; fourier: the 6 x n_freq matrix B (either may be NULL) */
int glh_mlp_get_params(glh_mlp *m, float *params, float *fourier);
int glh_mlp_set_params(glh_mlp *m, const float *params);
int glh_mlp_set_data(glh_mlp *m, int64_t n, const float *X, const float *Y);
/* `steps` Adam steps (beta 0.9/0.999, eps 1e-8) on random batches of
 * `batch` points (a multiple of 64) of the data, mean-squared-error loss;
 * loss_out[steps / loss_every] receives the batch loss every loss_every
 * steps (may be NULL). */
int glh_mlp_train(glh_mlp *m, int steps, int batch, double lr, uint64_t seed, int loss_every, float *loss_out);
/* mean-squared error and its gradient (n_params values) at the given points */
int glh_mlp_loss_grad(glh_mlp *m, int n, const float *X, const float *Y, float *loss, float *grad);
int glh_mlp_predict(glh_mlp *m, int64_t n, const float *X, float *Y);

/* ------------------------------------------------- per refinement cycle */
/* Uploads geometry, cell->node map, constraint masks, colour sets and the
 * BSR pattern once per refinement cycle (setup_uniform_B-phase.cc:125-198).
 * Also builds the cell->BSR scatter map used by the deterministic
 * colour-ordered assembly. */
int glh_upload_mesh(glh_ctx *ctx, const glh_mesh *m);

/* --------------------------------------------------------- solution I/O */
/* Seeded initial guess on the host (setup_uniform_B-phase.cc:200-265 for
 * phase 0, setup_uniform_A-phase.cc:225-262 for phase 1): the bulk phase of
 * gap `gap` plus amp*gap*(2u-1), u = (mt19937_64(seed) >> 11) * 2^-53 for
 * every dof in node-major order.  U[n_nodes*18]. */
int glh_initial_guess_host(int n_nodes, double gap, int phase, double amp, uint64_t seed, double *U);
/* A-phase / B-phase flat-wall guess (femgl/inc/BnA.h:130-163, used by
 * setup_uniform_BnA-flatwall-configuration.cc:228-235): bulk B phase where
 * z >= z_a, bulk A phase below, plus amp*gap_B*(2u-1) noise drawn as above.
 * node_xyz[n_nodes*3]. */
int glh_initial_guess_bna(int n_nodes, const double *node_xyz, double z_a, double gap_A, double gap_B, double amp,
                          uint64_t seed, double *U);
/* Analytic initial configurations at the nodes (VectorTools::interpolate
 * at the FE_Q support points), plus amp*scale*(2u-1) noise drawn as above
 * (amp = 0: the reference, which adds none).  kind / par:
 *   GLH_GUESS_BINA  B-in-A ball (femgl/inc/BinA.h:130-161, used by
 *                   setup_uniform_BinA-bubble-like-configuration.cc:229-235):
 *                   {radius, gap_A, gap_B}, scale gap_B;
 *   GLH_GUESS_BNB   B-n-B polka dots (femgl/inc/BinB.h:130-162, used by
 *                   setup_uniform_BnB-polkka-dot-like-configuration.cc:309-313):
 *                   {radius}, scale 1;
 *   GLH_GUESS_PDW   weak-coupling pair-density wave (femgl/inc/pdw_config.h:
 *                   124-299, used by setup_weak-coupling-PDW-configuration.cc:
 *                   311-315): {u, half_D, gap_B}, scale gap_B. */
#define GLH_GUESS_BINA 1
#define GLH_GUESS_BNB 2
#define GLH_GUESS_PDW 3
int glh_initial_guess_field(int kind, int n_nodes, const double *node_xyz, const double *par, double amp,
                            uint64_t seed, double *U);
int glh_set_solution(glh_ctx *ctx, const double *U_host);       /* local_solution */
/* keep a device copy of the current solution / restore it (resets a solve
 * without a host transfer) */
int glh_store_initial(glh_ctx *ctx);
int glh_restore_initial(glh_ctx *ctx);
int glh_get_solution(glh_ctx *ctx, double *U_host);
int glh_get_update(glh_ctx *ctx, double *dU_host);              /* locally_relevant_newton_solution */
int glh_get_rhs(glh_ctx *ctx, double *rhs_host);                /* system_rhs (= -R) */
/* system_matrix values in BSR order [nnzb][18][18] (FP64; an FP32 Jacobian
 * is widened).  For parity dumps of the assembled system. */
int glh_get_jacobian(glh_ctx *ctx, double *val_host);
/* y = J x with the Jacobian of the last assembly or Newton step, assembled
 * or matrix-free (FP64 vectors of n_dofs on the host).  A Jacobian-vector
 * product for parity checks of either form. */
int glh_apply_jacobian(glh_ctx *ctx, const double *x_host, double *y_host);
/* The 18x18 diagonal block of every node (n_nodes * 324 FP64 values on the
 * host, row-major) of the current Jacobian, assembled or matrix-free: the
 * blocks the point-block Jacobi preconditioner inverts (with constraints, the
 * condensed ones; a hanging node's is its diagonal row).  The assembled form's
 * are the diagonal blocks of the BSR values.  No counterpart in the
 * reference, which preconditions with Trilinos AMG (solve.cc:130-152); for
 * parity checks of the preconditioner of either form. */
int glh_diagonal_blocks(glh_ctx *ctx, double *out);
/* y = D^-1 x with the point-block Jacobi preconditioner of the current
 * Jacobian (the inverses of the blocks glh_diagonal_blocks returns, as the
 * GMRES of glh_solve applies them: FP32 for an FP32 Jacobian), FP64 vectors
 * of n_dofs on the host.  Inverts the blocks first if the Jacobian changed
 * (GLH_ERR_SINGULAR for a singular block).  No counterpart in the
 * reference (Trilinos AMG, solve.cc:130-152); for parity checks of the
 * preconditioner. */
int glh_apply_block_jacobi(glh_ctx *ctx, const double *x_host, double *y_host);

/* ---------------------------------------------------------- Newton step */
/* FemGL::assemble_system (assemble.cc:108-372): J and/or system_rhs = -R at
 * the current solution, with constrained rows condensed as
 * AffineConstraints::distribute_local_to_global does. flags: GLH_ASM_*;
 * with GLH_ASM_MATRIX_FREE the Jacobian is kept in the matrix-free form of
 * glh_set_matrix_free instead of BSR values. */
int glh_assemble(glh_ctx *ctx, int flags);
/* ||system_rhs|| after glh_assemble with GLH_ASM_R. */
int glh_rhs_norm(glh_ctx *ctx, double *norm);
/* FemGL::compute_residual (residual.cc:109-297) at the current solution:
 * writes ||R|| (R = residual_vector). */
int glh_residual(glh_ctx *ctx, double *norm);
/* A variant of glh_residual that allocates nothing.  It has no counterpart
 * in the reference: it is for callers that cannot allocate device memory
 * (code under HIP graph capture, tight serving loops).  U (n_dofs), R
 * (n_dofs, receives system_rhs = -R at U) and norm2 (n_inst, receives
 * ||R_s||^2 per instance) are device pointers; the call only enqueues
 * kernels on the context stream and returns (synchronise the device before
 * reading the results).  The partial sums of the norm live in a fixed
 * scratch array of 65536 doubles per device (one per reduction chunk of 1024
 * dofs, 4096 in instances of 2^20 dofs and more), so meshes needing more
 * chunks are refused with GLH_ERR_ARG; every call on one device shares that
 * array, so calls that overlap on a device race.  The cells' contributions
 * go through the context's cell-slot buffer, allocated with the mesh by
 * glh_upload_mesh (work on the same context is ordered on its stream).  The
 * mesh and material come from the context as for glh_residual. */
int glh_residual_into(glh_ctx *ctx, const double *U, double *R, double *norm2);

/* FemGL::solve (solve.cc:108-187) with GMRES(restart) + 18x18 point-block
 * Jacobi instead of FGMRES + ML AMG: J dU = system_rhs to
 * ||r|| <= rtol * ||system_rhs||.  stats[0] = iterations, stats[1] = final
 * relative residual.  GLH_ERR_NOCONV when maxit is reached. */
int glh_solve(glh_ctx *ctx, double rtol, int maxit, int restart, double stats[2]);
/* FemGL::newton_iteration (iteration.cc:109-215): damped back-tracking
 * u <- u + s^i dU for i < 100 (first trial that lowers ||R||), or one full
 * step when damped == 0.  out[0] = previous ||rhs||, out[1] = accepted ||R||,
 * out[2] = step length, out[3] = trials. */
int glh_newton_step(glh_ctx *ctx, double step, int damped, double out[4]);
/* Free energy F = integral of f (bulk + gradient) + Robin wall term; the
 * residual is R = (1/2) dF/dU.  The reference never forms F.  For a batch,
 * glh_free_energy returns the sum and glh_instance_energy one value per
 * instance. */
int glh_free_energy(glh_ctx *ctx, double *F);
int glh_instance_energy(glh_ctx *ctx, double *F);
/* ||R|| of every instance at the current solution (residual.cc:109-297 per
 * instance): R[n_inst]. */
int glh_instance_residuals(glh_ctx *ctx, double *R);

/* ------------------------------------------------------------ FemGL driver */
/* The Newton loop of FemGL::run (run.cc:208-256) on the uploaded mesh:
 * assemble -> solve -> newton_iteration, stopping at ||R|| <= rtol * ||R0||
 * (relative criterion) or after max_newton steps.  A linear solve that reaches
 * lin_maxit is not fatal: its partial update goes to the line search.
 * fp32 != 0 assembles the
 * Jacobian in FP32 and runs FP32 GMRES inside FP64 iterative refinement; the
 * residual, the update and the line search stay FP64.
 * stats: [0] Newton steps, [1] total GMRES iterations, [2] ||R0||,
 * [3] final ||R||, [4] seconds assembling, [5] seconds solving,
 * [6] seconds in line search, [7] free energy, [8] line-search residual
 * evaluations, [9] 1 if every instance converged.  Each instance of a batch
 * stops on its own test and is frozen afterwards. */
int glh_newton_solve(glh_ctx *ctx, double rtol, int max_newton, double lin_tol, int lin_maxit, int restart,
                     double ls_step, int damped, int fp32, double stats[10]);

/* Per-instance results of the last glh_newton_solve, 8 values per instance:
 * [0] Newton steps, [1] converged (0/1), [2] ||R0||, [3] final ||R||,
 * [4] seconds from the start of the solve to convergence (-1 if not),
 * [5] free energy, [6] line-search residual evaluations, [7] GMRES
 * iterations the instance used. */
int glh_instance_stats(glh_ctx *ctx, double *out);

/* timing of the last glh_assemble Jacobian launches on the context stream
 * (HIP events): ms[0] = J+R assembly ms, ms[1] = number of launches. */
int glh_last_assembly_ms(glh_ctx *ctx, double ms[2]);
/* level > 0: one progress line per Newton step on stderr (glh_newton_solve). */
/* Schedule of the solve: GMRES restart used when glh_newton_solve gets
 * restart <= 0 (1..126), Arnoldi steps per replayed graph block of a batch
 * (0: 8), grid caps of the active-set launches in block rows / 4 and in
 * chunks (0: 4096 / 1024), Gram-Schmidt tile (4, 8 or 16 basis vectors per
 * multi-dot workgroup and in flight per thread of the update; 0: 8). */
int glh_set_schedule(glh_ctx *ctx, int restart, int q_block, int row_cap, int chunk_cap, int gs_tile);
/* Second classical Gram-Schmidt pass of GMRES for a column whose squared norm
 * after the projection falls below `ratio` times its squared norm before
 * (default 0.01; 0.5 is the classic 1/sqrt(2) criterion, >= 1 always, 0
 * never).  No counterpart in the reference (deal.II's GMRES orthogonalises
 * inside SolverFGMRES, solve.cc:159-174): a knob of this solver's classical
 * Gram-Schmidt, used by the tests to run the second pass on every column. */
int glh_set_reorth(glh_ctx *ctx, double ratio);
/* Autotuner of the launch schedule and kernel tiles: times `trial_newton`
 * Newton steps (after one untimed step that records the graphs) from the
 * current solution, for a batch over blocks of 4 / 8 / 16 Arnoldi steps per
 * replayed graph and grid caps 2048 / 8192 block rows, then with the best of
 * those over Gram-Schmidt tiles 4 / 8 / 16; keeps the fastest in the context
 * (a later candidate has to be faster by more than 2%) and restores the
 * solution.  These knobs change launch and tile shapes
 * only: the iterates are the same bits.  The restart is the context's
 * (glh_set_schedule, default 100): it changes the iteration, which short
 * trials do not predict.  out: [0] restart [1] block [2] row cap [3] chunk
 * cap [4] best trial seconds [5] trials run [6] Gram-Schmidt tile. */
int glh_autotune(glh_ctx *ctx, int fp32, int trial_newton, double lin_tol, int lin_maxit, double out[7]);
/* Forcing terms of glh_newton_solve: 0 (default) the fixed lin_tol of the
 * reference (solve.cc:119-128); 1 Eisenstat-Walker (choice 2: eta_k =
 * 0.9 (||R_k|| / ||R_k-1||)^2 with the usual safeguard), capped at lin_tol,
 * per instance. */
int glh_set_forcing(glh_ctx *ctx, int eisenstat_walker);
/* Stopping reference of glh_newton_solve: with n = n_inst norms, instance s
 * stops at ||R_s|| <= rtol * r_ref[s] instead of rtol * ||R0_s|| (so solves
 * from different initial guesses reach the same absolute accuracy); n = 0
 * restores the default.  Cleared by glh_upload_mesh. */
int glh_set_residual_reference(glh_ctx *ctx, int n, const double *r_ref);
/* Pseudo-transient continuation in glh_newton_solve (an addition to the
 * reference's damped Newton, iteration.cc:109-215): every Newton matrix
 * becomes J + sigma_k |diag J| on the unconstrained dofs with
 * sigma_k = sigma0 ||R_k|| / ||R_0|| per instance, which follows the
 * gradient flow of F far from a minimum and is plain Newton near it.  The
 * fixed point is unchanged.  0 (default) = off. */
int glh_set_continuation(glh_ctx *ctx, double sigma0);
/* Matrix-free Jacobian in glh_newton_solve (a GPU alternative to the
 * assembled Trilinos matrix behind SolverFGMRES, solve.cc:159-174): on != 0
 * makes each Newton step store the iterate of every cell at its quadrature
 * points and only the 18x18 diagonal blocks, and GMRES applies J x cell by
 * cell from them, the bulk term as the derivative of the residual density at
 * the stored iterate (the same operator as the assembled BSR matrix, to
 * rounding).  With fp32 the iterate is stored in FP32 and the inner GMRES,
 * its Krylov basis and the block Jacobi run in FP32, inside the FP64
 * iterative refinement.  On a mesh with constraint lines (hanging nodes,
 * periodic images) the product is C^T (J_cells (C x)) with the hanging rows
 * diagonal, the condensed matrix of the assembled path, and the block Jacobi
 * uses the condensed diagonal blocks; the continuation shift is added to the
 * diagonal blocks and to the product.  It applies to solves of one instance
 * or, on meshes without constraint lines, of a batch; other solves assemble
 * J as before.  glh_assemble builds the BSR values unless it is given
 * GLH_ASM_MATRIX_FREE; glh_get_jacobian fails after a matrix-free step.
 * 0 (default) = off. */
int glh_set_matrix_free(glh_ctx *ctx, int on);
int glh_set_verbose(glh_ctx *ctx, int level);
/* Live kernel timing with HIP events on the context stream.  on != 0 resets
 * the counters and records an event pair around every Jacobian product
 * (BSR SpMV, or the matrix-free product and its gather) and every assembly
 * from then on; GMRES cycles are then launched directly instead of replayed
 * from captured graphs.  out: [0] total product ms, [1] products, [2] total
 * Jacobian (+ residual) assembly ms, [3] such assemblies, [4] total
 * residual-only ms (the residual's cell pass and node gather, and on a
 * constrained mesh its condensation), [5] such residuals. */
int glh_set_timing(glh_ctx *ctx, int on);
int glh_timing(glh_ctx *ctx, double out[6]);
/* Run k back-to-back SpMV launches y = J x with x = current solution and
 * time them with HIP events on the context stream: ms_avg. */
int glh_bench_spmv(glh_ctx *ctx, int k, double *ms_avg);

/* Average time of a streaming read of nbytes (multiple of 16) over k
 * launches (non-temporal 16-byte loads, 8 in flight per thread, 1024
 * workgroups): the device's sustained HBM read rate, the practical ceiling of
 * the read-dominated SpMV. */
int glh_bench_read(glh_ctx *ctx, size_t nbytes, int k, double *ms_avg);
/* Run k launches of a plain 16-byte-vector device copy of nbytes (a multiple
 * of 16) between two scratch buffers; ms_avg per launch.  A kernel of known
 * traffic, used to calibrate the HBM byte counters. */
int glh_bench_copy(glh_ctx *ctx, size_t nbytes, int k, double *ms_avg);

#ifdef __cplusplus
}
#endif
#endif
