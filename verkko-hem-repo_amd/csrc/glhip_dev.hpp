// Internal interface between the host driver (glhip_api.cpp) and the HIP
// kernels (gl_assemble.hip, gl_linalg.hip).  Plain device pointers only.
#pragma once
#include <hip/hip_runtime.h>

#include <cstdint>

#include "common.hpp"

namespace glh {

// Reference-element tables for FE_Q(k) x QGauss(k+1) on the unit cube
// (assemble.cc:123-132), built once per context on the host.
struct RefTables {
  const double *N;    // [NQ*NS]   N_a(x_q)
  const double *dN;   // [NQ*NS*3] dN_a/dxi_k(x_q)
  const double *w;    // [NQ]      quadrature weights (sum 1)
  const double *Wp;   // [NPAIR*NQ] w_q N_a N_b for a <= b pairs
  const double *Sp;   // [NPAIR*9]  sum_q w_q dN_a/dxi_k dN_b/dxi_l, index k*3+l
  const double *Mf;   // [6*NS*NS]  face mass on each unit face
  const int *pa, *pb; // [NPAIR]    pair list (a <= b)
  const double *L1;   // [N1*N1]   1D Lagrange basis: L1[qx*N1 + bx] = L_bx(x_qx)
  const double *dL1;  // [N1*N1]   its derivative
  const float *L1f, *dL1f;  // the two rounded to FP32
};

struct Material {
  double al, b1, b2, b3, b4, b5;  // alpha, beta1..5
  double K1, K23;                 // K1 and K2+K3 (assemble.cc:229-233)
  double robin;                   // K1/bt when bt < 1e10, else 0 (assemble.cc:293)
};

struct MeshDev {
  const int *cell_mat;       // [n_cells] material index, or null: one material for all cells
  const Material *mats;      // per-instance materials (with cell_mat)
  const int *cell_nodes;     // [n_cells*NS]
  const double *cell_h;      // [n_cells*3]
  const uint32_t *cell_face; // [n_cells] 3 bits per local face: boundary id + 1 (0 = interior)
  const uint32_t *node_mask; // [n_nodes] Dirichlet component bits
  const int *cell_blk;       // [n_cells*NS*NS] BSR block of (a,b); -(blk+1) when this cell writes first
  double *hdiag;             // [n_nodes*18] sum of |local diagonal| of hanging nodes, or null
  const int *cell_inst;      // [n_cells] instance of each cell
  const int *inst_active;    // [n_inst] 0: the instance's cells are skipped (J and R left as they are), or null
};

// Hanging-node constraints on the device (see gl_constraints.hip).
struct ConstraintDev {
  int n_nodes = 0, nnzb = 0, n_hang = 0;
  const int *gptr = nullptr, *gsrc = nullptr;  // Jc gather lists per block
  const double *gw = nullptr;
  const unsigned char *gfl = nullptr;
  const int *brow = nullptr, *col = nullptr;  // block row / column of every Jc block
  const uint32_t *mask = nullptr;             // node masks
  const int *diag = nullptr;                  // diagonal block of every node in Jc
  const int *rptr = nullptr, *rsrc = nullptr; // rhs gather lists per node
  const double *rw = nullptr;
  const unsigned char *rfl = nullptr;
  const int *hnode = nullptr, *hptr = nullptr, *hm = nullptr;  // hanging nodes and their closed lines
  const double *hw = nullptr;
  const int *hidx = nullptr;                  // node -> its index in hnode, -1 if not hanging
};
// The matrix-free Jacobian under constraints.  The condensed diagonal block of
// a master node n is the sum over cells of its own local block (n, n), which
// the setup forms as on an unconstrained mesh, plus the local blocks (a, b) of
// every cell whose nodes A, B both reach n through constraint lines (A or B
// hanging), weighted wA wB.  Per cell e, the setup stores the raw local blocks
// of the pairs xpair[xptr[e] .. xptr[e+1]) (pair index of (a <= b)) into LB
// slots numbered like the list; per target node tnode[t], terms
// tptr[t] .. tptr[t+1] name a slot, the nodes A (row) and B (column) and the
// weight, and tfl: bit 0 A hanging, bit 1 B hanging, bit 2 the slot's block
// transposed (orientation (b, a)), bit 3 a == b.
struct MfCondDev {
  const int *xptr = nullptr, *xpair = nullptr;
  double *LB = nullptr;
  int n_targets = 0;
  const int *tnode = nullptr, *tptr = nullptr, *tslot = nullptr, *tA = nullptr, *tB = nullptr;
  const double *tw = nullptr;
  const unsigned char *tfl = nullptr;
};
// Jc = C^T Jf C over the condensed pattern, hanging rows diagonal (hdiag)
template <typename TJ>
hipError_t launch_condense(const ConstraintDev &C, const TJ *Jf, TJ *Jc, const double *hdiag, hipStream_t s);
// rc = C^T rf, hanging rows 0
hipError_t launch_condense_rhs(const ConstraintDev &C, const double *rf, double *rc, hipStream_t s);
// hanging values of v from their masters
hipError_t launch_distribute(const ConstraintDev &C, double *v, hipStream_t s);
// Matrix-free Jacobian under constraints (with MfCondDev):
// the condensed diagonal blocks in place in Jd ([n_nodes + n_hang][324]):
// the listed terms added to the masters' blocks, and block n_nodes + j of
// hanging node hnode[j] set to diag(hdiag) (its condensed row)
hipError_t launch_mf_cond_diag(const ConstraintDev &C, const MfCondDev &X, int n_nodes, const double *hdiag,
                               double *Jd, hipStream_t s);
// xf = C x: x, with every hanging node's values taken from its masters
// (Dirichlet components of a master contribute 0)
template <typename TV>
hipError_t launch_mf_distribute(const ConstraintDev &C, int n_nodes, const TV *x, TV *xf, const int *row_seg,
                                const int *flag, hipStream_t s);
// y = C^T yf on the unconstrained rows and hdiag x on the hanging ones, or
// b minus that with b != null
template <typename TV>
hipError_t launch_mf_condense(const ConstraintDev &C, int n_nodes, const double *hdiag, const TV *yf, const TV *x,
                              const TV *b, TV *y, const int *row_seg, const int *flag, hipStream_t s);
// continuation shift of the matrix-free Jacobian: sh = sigma |diag| on every
// free component (0 on Dirichlet and hanging rows), added to the diagonal of
// the blocks in Jd
hipError_t launch_mf_shift(int n_nodes, const uint32_t *mask, const int *row_seg, const double *sigma, double *Jd,
                           double *sh, hipStream_t s);
// Un[v*18+c] = Q_deg interpolant of the old cell cell[v] (nodes
// old_cn[cell*(deg+1)^3..], lexicographic) at t[3v..3v+2], from Uo
hipError_t launch_transfer(int n_nodes, int deg, const int *cell, const float *t, const int *old_cn,
                           const double *Uo, double *Un, hipStream_t s);

// ---------------------------------------------------------------- assembly
// The Jacobian, one launch per colour; cells[] lists the cells of that
// colour.  J: double* or float* (jfp32).
hipError_t launch_assemble(int deg, bool jfp32, const RefTables &T, const MeshDev &M, const Material &mat,
                           const double *U, const int *cells, int n, void *J, hipStream_t s);
// Matrix-free Jacobian: per cell the iterate at the quadrature points (Uq:
// [n_cells][NQ][18]; the product forms H(q) X(q) from it) and the diagonal
// blocks (Jd: [n_nodes][324], one more per hanging node on a constrained
// mesh), the setup colour-ordered like the assembly; the product writes
// each listed cell's contribution to its slot of ycell ([n_cells][NS][18]),
// and launch_mf_gather sums them per node (node -> cell*NS + a lists, fixed
// order) and writes the Dirichlet rows.
// u32: the iterate stored in FP32 (for the FP32 inner solve); x32: x in FP32.
// The gather returns y = J x, or y = b - J x with b != null.
// With X.xptr set (constrained meshes) the setup also stores the raw local
// blocks X lists and adds the |local diagonal| of hanging nodes to M.hdiag.
hipError_t launch_mf_setup(int deg, bool u32, const RefTables &T, const MeshDev &M, const Material &mat,
                           const double *U, const int *cells, int n, void *Uq, double *Jd, const MfCondDev &X,
                           hipStream_t s);
// flag (per instance, or null): cells and rows of instances whose flag is 0
// are skipped, as in the batched SpMV.
hipError_t launch_mf_apply(int deg, bool u32, bool x32, const RefTables &T, const MeshDev &M, const Material &mat,
                           const void *Uq, const void *x, const int *cells, int n, const int *flag, void *ycell,
                           hipStream_t s);
// sh (or null): a diagonal term added to every row (the continuation shift).
template <typename TX>
hipError_t launch_mf_gather(int n_nodes, const int *ptr, const int *idx, const void *ycell, const uint32_t *mask,
                            const double *Jd, const double *sh, const TX *x, const TX *b, const int *row_seg,
                            const int *flag, TX *y, hipStream_t s);
// The residual through the product's sum-factorised passes: the listed cells'
// contributions to R at U in their ycell slots (cells of batch instances that
// are not active skipped), then rhs = -R, all rows written: 0 on the
// Dirichlet components and on the rows of instances that are not active.
hipError_t launch_res_apply(int deg, const RefTables &T, const MeshDev &M, const Material &mat, const double *U,
                            const int *cells, int n, double *ycell, hipStream_t s);
hipError_t launch_res_gather(int n_nodes, const int *ptr, const int *idx, const double *ycell, const uint32_t *mask,
                             const int *row_seg, const int *active, double *rhs, hipStream_t s);
// Per-cell free energy into Fcell[cell] for cells [0, n_cells).
hipError_t launch_energy(int deg, const RefTables &T, const MeshDev &M, const Material &mat, const double *U,
                         int n_cells, double *Fcell, hipStream_t s);

// --------------------------------------------------------- linear algebra
// Compacted work lists of the instances that take part in a batched
// operation, built on the device (launch_build_active) so that a replayed
// graph spends workgroups only on them: counts[0] instances, counts[1]
// chunks, counts[2] block rows; seg[] the instance ids, rpref[] the prefix
// of their block rows (counts[0]+1 values), chunk[] their chunk ids.
// counts == nullptr: no list, every instance takes part.
// shift: the grids of the grid-stride launches are divided by 2^shift
// (chosen by the host from the size of the set; any value is correct).
struct ActiveSet {
  const int *counts = nullptr, *seg = nullptr, *rpref = nullptr, *chunk = nullptr, *node_off = nullptr;
  int shift = 0;
  int row_cap = 0, chunk_cap = 0;  // grid caps (0: ACT_ROW_GRID / ACT_CHUNK_GRID)
};
constexpr int ACT_ROW_GRID = 4096;   // workgroups of 4 rows, grid-stride over listed rows
constexpr int ACT_CHUNK_GRID = 1024; // workgroups, grid-stride over listed chunks
constexpr int ACT_SEG_GRID = 256;    // workgroups per reduced value, grid-stride over listed instances

// y = J x over BSR(18x18); TJ in {double,float}, TX in {double,float}.
// row_seg/seg_flag (optional): rows of instances whose flag is 0 are skipped.
template <typename TJ, typename TX>
hipError_t launch_spmv(int n_rows, const int *row_ptr, const int *col, const TJ *val, const TX *x, TX *y,
                       hipStream_t s, const int *row_seg = nullptr, const int *seg_flag = nullptr,
                       ActiveSet act = ActiveSet{});
// r = b - J x in FP64 arithmetic with an FP32 or FP64 matrix.
template <typename TJ>
hipError_t launch_residual_spmv(int n_rows, const int *row_ptr, const int *col, const TJ *val, const double *x,
                                const double *b, double *r, hipStream_t s);
// Dinv[r] = inverse of the (symmetric) diagonal block of row r (Gauss-Jordan
// with partial pivoting in FP64), stored as its upper triangle: 171 values
// row-major over c <= d; status[0] set to 1 if a block is singular.
template <typename TJ>
hipError_t launch_bj_invert(int n_rows, const int *diag_blk, const TJ *val, TJ *Dinv, int *status,
                            hipStream_t s);
// y = Dinv x, block-diagonal apply (Dinv as launch_bj_invert stores it).
template <typename TJ, typename TX>
hipError_t launch_bj_apply(int n_rows, const TJ *Dinv, const TX *x, TX *y, hipStream_t s,
                           const int *row_seg = nullptr, const int *seg_flag = nullptr,
                           ActiveSet act = ActiveSet{});

// Deterministic reductions.  partial must hold MDOT_GRID * k values.
constexpr int MDOT_GRID = 2048;
// out[i] (+)= sum_j V[i*ldv + j] * w[j] for i < k.  accumulate: out += result.
template <typename T>
hipError_t launch_mdot(size_t n, int k, const T *V, size_t ldv, const T *w, T *partial, T *out, bool accumulate,
                       hipStream_t s);
// w[j] -= sum_i h[i] V[i*ldv + j]
template <typename T>
hipError_t launch_sub_lincomb(size_t n, int k, const T *V, size_t ldv, const T *h, T *w, hipStream_t s);
// w[j] -= sum_i h[i] V[i*ldv + j]; *norm2 = ||w||^2 afterwards (partial: MDOT_GRID values)
template <typename T>
hipError_t launch_sub_lincomb_norm(size_t n, int k, const T *V, size_t ldv, const T *h, T *w, T *partial, T *norm2,
                                   hipStream_t s);
// w[j] = sum_i y[i] V[i*ldv + j]
template <typename T>
hipError_t launch_lincomb(size_t n, int k, const T *V, size_t ldv, const T *y, T *w, hipStream_t s);
// dst = src / *scal (scal on device; dst = src when *scal == 0)
template <typename T>
hipError_t launch_scale_by(size_t n, const T *src, const T *scal, T *dst, hipStream_t s);
// y += a x (a on host), converting types
template <typename TX, typename TY>
hipError_t launch_axpy(size_t n, double a, const TX *x, TY *y, hipStream_t s);
// dst = src (type conversion)
template <typename TS, typename TD>
hipError_t launch_convert(size_t n, const TS *src, TD *dst, hipStream_t s);
// dst[r] (324 values) = the block blk[r] of val, converted to FP64
template <typename TJ>
hipError_t launch_gather_blocks(int n_rows, const int *blk, const TJ *val, double *dst, hipStream_t s);
// dst = b - w  (double)
hipError_t launch_sub(size_t n, const double *b, const double *w, double *dst, hipStream_t s);
// X = U + a dU, then Dirichlet components of X set to zero (constraints distribute)
hipError_t launch_trial(int n_nodes, const double *U, const double *dU, double a, const uint32_t *mask, double *X,
                        hipStream_t s);
// zero the Dirichlet components of x
hipError_t launch_mask_zero(int n_nodes, const uint32_t *mask, double *x, hipStream_t s);
// ------------------------------------------------ segmented (batched) vectors
// Instances own contiguous dof ranges.  chunks: [3*n_chunks] (instance, first
// dof, end dof), no chunk straddles an instance; seg_chunk: [n_seg+1] chunk
// ranges per instance.  Outputs are indexed out[s*ldo + off + i].
// longest reduction chunk (values) the segmented kernels accept
constexpr long long GS_MAX_CHUNK = 4096;
struct SegInfo {
  const long long *chunks;
  const int *seg_chunk;
  int n_chunks, n_seg;
  ActiveSet act;  // optional: only the listed instances' chunks are visited
  // Gram-Schmidt tile (4, 8 or 16): basis vectors per multi-dot workgroup,
  // and basis vectors in flight per thread of the update.  The sums run in
  // the same order for every tile, so the results do not depend on it.
  int gs = 8;
  // grid cap of the chunk kernels without an active set (0: one workgroup
  // per chunk): for launches that usually find their flag off (the second
  // Gram-Schmidt pass), where dispatching a workgroup per chunk only to skip
  // it is most of the cost
  int cap = 0;
};
template <typename T>
struct GmresCtl;

// w -= sum_i h[s*ldh + i] V_i per instance; norm2[s*ldo + off] = ||w_s||^2 after.
// With ctl (norm2 == ctl->nrm2, ldo 1, off 0), the workgroup that reduces an
// instance's norm then runs GMRES control op `op` (1 or 2) for column j of
// that instance, in place of a launch_gmres_ctl.  With hpart, the chunk
// partials of a launch_seg_mdot of hk values (reduce = false) and not the
// norm partials' buffer, each workgroup forms its instance's h from them
// (chunk order) and the instance's first chunk stores h[s*ldh + i], i < hk:
// for batches of small instances, where a separate reduction launch costs
// more than the sums.
template <typename T>
hipError_t launch_seg_sub_lincomb_norm(const SegInfo &S, int k, const T *V, size_t ldv, T *h, int ldh, T *w,
                                       T *partial, T *norm2, int ldo, int off, hipStream_t s,
                                       const int *flag = nullptr, const GmresCtl<T> *ctl = nullptr, int op = 0,
                                       int j = 0, const T *hpart = nullptr, int hk = 0, bool reduce = true);
// reduce = false: no norm reduction launch (k_seg_ctl2_scale takes the partials)
// After the second Gram-Schmidt pass: norm2 of each iterating
// instance from the pass's chunk partials (where it ran; else the first
// pass's), control op 2 and w /= sqrt(norm2), in one launch
template <typename T>
hipError_t launch_seg_ctl2_scale(const SegInfo &S, const T *partial, T *norm2, const GmresCtl<T> &C, int j, T *w,
                                 hipStream_t s);
// w = sum_{i < min(k, kseg[s])} y[s*ldy + i] V_i
template <typename T>
hipError_t launch_seg_lincomb(const SegInfo &S, int k, const T *V, size_t ldv, const T *y, int ldy, T *w,
                              hipStream_t s, const int *kseg = nullptr);
// dst = src / scal[s*lds] (sqrt_of_scale: / sqrt(scal)); unchanged where the scale is 0
template <typename T>
hipError_t launch_seg_scale(const SegInfo &S, const T *src, const T *scal, int lds, T *dst, hipStream_t s,
                            const int *flag = nullptr, bool sqrt_of_scale = false);
// r = b - w with norm2[s] = ||r_s||^2
template <typename T>
hipError_t launch_seg_residual(const SegInfo &S, const T *b, const T *w, T *r, T *partial, T *norm2, hipStream_t s);
template <typename T>
hipError_t launch_fill(size_t n, T v, T *d, hipStream_t s);
// J_cc += sigma[row_seg[r]] |J_cc| on unconstrained diagonal entries
template <typename TJ>
hipError_t launch_diag_shift(int n_rows, const int *diag, const uint32_t *mask, const int *row_seg,
                             const double *sigma, TJ *J, hipStream_t s);
// y += x per instance (over the active set when S.act is set)
template <typename T>
hipError_t launch_seg_add(const SegInfo &S, const T *x, T *y, hipStream_t s);
// Active set of the instances whose flag is nonzero (node_off: [n_seg+1]
// first block row of every instance; seg: n_seg, rpref: n_seg+1, chunk:
// n_chunks, counts: 3 ints).
hipError_t launch_build_active(const SegInfo &S, const int *flag, const int *node_off, int *counts, int *seg,
                               int *rpref, int *chunk, hipStream_t s);

// Device state of the batched GMRES: Hessenberg columns, Givens rotations
// and stopping tests live on the device so a restart cycle needs no host
// round trip (and can be captured in a HIP graph).
template <typename T>
struct GmresCtl {
  int S, m, ldh;
  const T *hs, *hq;      // [S*ldh] Gram-Schmidt coefficients of the two passes
  const T *nrm2;         // [S] squared norms
  T *y;                  // [S*ldh] least-squares coefficients
  double *H, *cs, *sn, *g, *beta;
  const double *target;  // [S] absolute tolerances
  int *cyc, *conv, *jseg, *its, *ro;
  const int *frozen;
  int *jlast;  // [S] basis size of the previous cycle (written by op 0)
  double dgks;  // second Gram-Schmidt pass when ||w'||^2 < dgks ||w||^2
};
// Control op 1 taken into a launch_seg_mdot (folded batches, the second
// Gram-Schmidt pass): with npart set, each workgroup forms its instance's
// norm from npart (the first pass's chunk partials) and the pass's flag
// itself, in place of `flag`, and the instance's first chunk writes norm2
// (== C.nrm2) and C.ro (reduce must be false).
template <typename T>
struct Op1In {
  const T *npart = nullptr;
  T *norm2 = nullptr;
  const int *seg_chunk = nullptr;
  GmresCtl<T> C{};
  int j = 0;
};
// flag (optional): instances whose flag is 0 are skipped (outputs untouched).
template <typename T>
hipError_t launch_seg_mdot(const SegInfo &S, int k, const T *V, size_t ldv, const T *w, T *partial, T *out, int ldo,
                           int off, bool accumulate, hipStream_t s, const int *flag = nullptr, bool reduce = true,
                           const Op1In<T> &op1 = Op1In<T>{});
template <typename T>
hipError_t launch_gmres_ctl(const GmresCtl<T> &C, int op, int j, hipStream_t s);
hipError_t launch_seg_trial(const SegInfo &S, const double *U, const double *dU, const double *alpha,
                            const uint32_t *mask, double *X, hipStream_t s);
template <typename T>
hipError_t launch_seg_zero(const SegInfo &S, const int *flag, T *v, hipStream_t s);

// *p = v (stream-ordered scalar write)
template <typename T>
hipError_t launch_set_scalar(T *p, T v, hipStream_t s);
// plain device copy of nbytes (multiple of 16): calibration kernel for HBM counters
hipError_t launch_copy(size_t nbytes, const void *src, void *dst, hipStream_t s);
// streaming read of nbytes (multiple of 16) by grid workgroups of 256 threads;
// out holds grid * 256 floats
hipError_t launch_read(size_t nbytes, const void *src, float *out, int grid, hipStream_t s);
// out = sum_i v[i] in fixed order (deterministic)
hipError_t launch_sum(size_t n, const double *v, double *partial, double *out, hipStream_t s);

// Device scratch of the no-allocation entry point (glh_residual_into): a
// file-scope array of SCRATCH_DOUBLES doubles on every device the library
// runs on, held until the process ends.  scratch_address gives its device
// address on the current device (looked up once per device and kept); host
// code must use that address, never the array's name.
constexpr long long SCRATCH_DOUBLES = 1 << 16;
hipError_t scratch_address(int device, double **out);

}  // namespace glh
