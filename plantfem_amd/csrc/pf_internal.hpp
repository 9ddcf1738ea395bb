// pf_internal.hpp -- shared declarations of the gfx950 kernels behind include/pfgpu.h.
#pragma once
#include <hip/hip_runtime.h>

#include <cstddef>
#include <cstdint>
#include <vector>

#include "../../include/pfgpu.h"

// ---------------------------------------------------------------------------------
// error plumbing: every host wrapper returns a PF_* status and records a message
void pf_set_error(const char *fmt, ...);

#define PF_HIP(call)                                                                 \
  do {                                                                               \
    hipError_t e_ = (call);                                                          \
    if (e_ != hipSuccess) {                                                          \
      pf_set_error("%s:%d %s -> %s", __FILE__, __LINE__, #call, hipGetErrorString(e_)); \
      return PF_ERR_HIP;                                                             \
    }                                                                                \
  } while (0)

#define PF_CHECK(expr)         \
  do {                         \
    int rc_ = (expr);          \
    if (rc_ != PF_OK) return rc_; \
  } while (0)

struct SolverWork;

struct pf_ctx {
  int device = 0;
  hipStream_t stream = nullptr;
  hipEvent_t ev0 = nullptr, ev1 = nullptr;
  int num_cus = 256;
  SolverWork *work = nullptr;  // workspace of the stateless solver entry point
};

// ---------------------------------------------------------------------------------
// device buffer: owns a hipMalloc'd array; grows, never shrinks
template <typename T>
struct DevBuf {
  T *p = nullptr;
  size_t n = 0;
  int ensure(size_t count) {
    if (count <= n && p) return PF_OK;
    release();
    size_t bytes = (count ? count : 1) * sizeof(T);
    hipError_t e = hipMalloc((void **)&p, bytes);
    if (e != hipSuccess) {
      p = nullptr;
      n = 0;
      pf_set_error("hipMalloc(%zu bytes) failed: %s", bytes, hipGetErrorString(e));
      return PF_ERR_NOMEM;
    }
    n = count;
    return PF_OK;
  }
  void release() {
    if (p) (void)hipFree(p);
    p = nullptr;
    n = 0;
  }
  ~DevBuf() { release(); }
  DevBuf() = default;
  DevBuf(const DevBuf &) = delete;
  DevBuf &operator=(const DevBuf &) = delete;
};

// ---------------------------------------------------------------------------------
// mesh + node-level graph (row structure shared by the DOF rows of one node)
struct NodeGraph {
  int64_t nn = 0, ne = 0;
  int32_t nne = 0;
  DevBuf<double> nodcoord;   // nn*3 column-major
  DevBuf<int32_t> elemnod;   // ne*nne column-major, 1-based
  // node -> incident elements, ascending element id
  DevBuf<int64_t> eoff;      // nn+1
  DevBuf<int32_t> elist;     // ne*nne (0-based element ids)
  // node -> distinct neighbour nodes in first-appearance order
  DevBuf<int32_t> deg;       // nn
  DevBuf<int64_t> noff;      // nn+1
  DevBuf<int32_t> ncol;      // noff[nn] (0-based node ids)
  int64_t nblocks = 0;       // noff[nn]
  int32_t maxdeg = 0;
  bool has_coords = false;   // nodcoord uploaded
  bool built = false;        // eoff/elist valid
  bool cols_built = false;   // deg/noff/ncol valid
  mutable DevBuf<double> kebuf;  // element matrices of the two-pass stiffness assembly (ne x 576)
};

// blocked sliced-ELL copy of a DOF=3 CRS: 64 nodes per slice (one wavefront), slot t of
// node l of slice s at block index (soff[s] + t)*64 + l; the 3x3 values of the slot's 64
// blocks (q = 3*row_dof + col_dof) at val[(soff[s] + t)*576 + sell3_value_index<V>(q, l)]:
// 16-byte groups per lane for q < 8, the ninth value after them.
template <typename V>
struct Sell3 {
  int64_t nn = 0, nslices = 0, nslots = 0;  // nslots = soff[nslices]
  DevBuf<int64_t> soff;                    // nslices+1
  DevBuf<int32_t> width;                   // nslices
  DevBuf<int32_t> col;                     // nslots*64 padded node ids (-1 = pad)
  DevBuf<V> val;                           // nslots*64*9
  // slices of a block-diagonal batch never cross a mesh boundary: slice s then covers
  // nodes snode[s] .. send[s]-1 (device arrays owned by the caller); null = s*64 .. s*64+63
  const int64_t *snode = nullptr, *send = nullptr;
};

// slice -> node map of a batched Sell3 (device arrays, nslices entries each)
struct SliceMap {
  int64_t nslices = 0;
  const int64_t *snode = nullptr, *send = nullptr;
};

// ---------------------------------------------------------------------------------
// host wrappers implemented in the .hip files (all enqueue on ctx->stream)

// scan.hip: exclusive scan of n counts into out[0..n] (out[n] = total)
int pf_scan_i32_to_i64(pf_ctx *ctx, const int32_t *in, int64_t n, int64_t *out, DevBuf<int64_t> &tmp);
int pf_scan_i64(pf_ctx *ctx, const int64_t *in, int64_t n, int64_t *out, DevBuf<int64_t> &tmp);
int pf_read_i64(pf_ctx *ctx, const int64_t *dev, int64_t *host);

// graph.hip
int pf_graph_node_elems(pf_ctx *ctx, NodeGraph &g);                 // eoff/elist (sorted)
int pf_graph_node_cols(pf_ctx *ctx, NodeGraph &g);                  // deg/noff/ncol
int pf_graph_expand(pf_ctx *ctx, const NodeGraph &g, int dof, int64_t *row_ptr, int32_t *col_idx);
int pf_graph_check_mesh(pf_ctx *ctx, const NodeGraph &g);           // node ids in range
// node neighbour lists of the multi-domain setCRS: g's element pairs and the node pairs of
// every overset connect (my/pair: 0-based element ids of g, device; gp: Gauss point ids, 0 =
// P2P, where my is the own node and the connect has 9 nodes; NULL = all GPP), sorted and
// distinct; fills out.deg/noff/ncol/nblocks for pf_graph_expand
int pf_graph_merged_cols(pf_ctx *ctx, NodeGraph &g, int64_t ncon, const int32_t *my, const int32_t *pair,
                         const int32_t *gp, NodeGraph &out);
// ascending radix sort of n 64-bit keys over bits [0, end_bit): the result lands in tmp
// (*in_tmp = true) unless n <= 1
int pf_sort_u64(pf_ctx *ctx, uint64_t *keys, uint64_t *tmp, int64_t n, int end_bit, bool *in_tmp);

// assemble.hip
// ke_kernel: element matrices in the reference's operation order (0, bitwise), or through the
// isotropic split with the contraction on the VALU (1) or the FP64 matrix cores (2)
int pf_assemble_hex8(pf_ctx *ctx, const NodeGraph &g, const double *E, const double *nu, int jmode,
                     const int64_t *row_ptr, double *val, int ke_kernel = 0);
int pf_hex8_ke_batch(pf_ctx *ctx, const NodeGraph &g, const double *E, const double *nu, int jmode,
                     int64_t first, int64_t count, double *ke);
// batched setMatrix / setVector: node gathers over a batch in call order
int pf_add_element_mats(pf_ctx *ctx, const NodeGraph &g, int dof, int64_t count, const int32_t *ids,
                        const double *ke, const int64_t *row_ptr, const int32_t *col_idx, double *val,
                        int32_t *err);
int pf_add_element_vecs(pf_ctx *ctx, const NodeGraph &g, int dof, int64_t count, const int32_t *ids,
                        const double *vec, double *rhs);
int pf_traction_box(pf_ctx *ctx, const NodeGraph &g, const double *box, double pressure, double *rhs);
// overset GPP: the Gauss points of A's elements located in elements of B, in (element,
// Gauss point) order: 1-based own element, Gauss point, element of B, and the position.
// B needs its node->element lists (pf_graph_node_elems).
// p2p: the points are A's nodes instead (elem = own node, gp = 0), as overset(P2P)
int pf_overset_search(pf_ctx *ctx, const NodeGraph &A, const NodeGraph &B, bool p2p, DevBuf<int32_t> &elem,
                      DevBuf<int32_t> &gp, DevBuf<int32_t> &pair_elem, DevBuf<double> &position, int64_t *count);
// setEbOM: penalty * ConnectMatrix of ncon connects (0-based own element, or own node when the
// Gauss point id is 0 (P2P); 0-based paired element; positions) added row by row in connect
// order; err set for a missing slot
int pf_ebom_assemble(pf_ctx *ctx, const NodeGraph &g, int64_t ncon, const int32_t *my, const int32_t *gp,
                     const int32_t *pr, const double *pos, double penalty, const int64_t *row_ptr,
                     const int32_t *col_idx, bool sorted, double *val, int32_t *err);
// hex8 stiffness added into a CRS of any column order (sorted: binary search of the rows)
int pf_assemble_hex8_search(pf_ctx *ctx, const NodeGraph &g, const double *E, const double *nu, int jmode,
                            const int64_t *row_ptr, const int32_t *col_idx, bool sorted, double *val,
                            int32_t *err);
// MassVectorFEMDomain of elements first..first+count-1 (0-based first): out[q*24 + 3a + k]
int pf_hex8_mass_batch(pf_ctx *ctx, const NodeGraph &g, double rho, const double *accel, int64_t first,
                       int64_t count, double *out);
// rhs += sum of element mass vectors (rho per element from rho_dev, else rho_all), ascending element order
int pf_body_force(pf_ctx *ctx, const NodeGraph &g, const double *rho_dev, double rho_all, const double *accel,
                  double *rhs);
// MassMatrixFEMDomain (DOF=3) of elements first..first+count-1 (24x24 column-major each) and
// MassMatrix_as_CRS into the DOF=3 CRS values; rho per element (srho = 1) or one value
int pf_hex8_mass_mat(pf_ctx *ctx, const NodeGraph &g, const double *rho, int srho, int64_t first, int64_t count,
                     double *out);
int pf_assemble_mass(pf_ctx *ctx, const NodeGraph &g, const double *rho, int srho, const int64_t *row_ptr,
                     double *val);
// cell-averaged Cauchy stress per element, out[e*9 + 3j + i]; sE / snu: 1 per element, 0 one value
int pf_cell_stress(pf_ctx *ctx, const NodeGraph &g, const double *disp, const double *E, int sE, const double *nu,
                   int snu, int jmode, double *out);
// out[e] = sum_l lam_e(l)^T Ke(E_e, nu_e) u_e(l) over R interleaved cases (device arrays;
// fixed DOFs, if given, excluded)
int pf_elem_bilinear(pf_ctx *ctx, const NodeGraph &g, const double *E, const double *nu, int jmode, int R,
                     const double *U, const double *Lam, const uint8_t *fixed, double *out);
// one box and pressure per node range node_off[m] .. node_off[m+1]-1 (device arrays)
int pf_traction_boxes(pf_ctx *ctx, const NodeGraph &g, int32_t count, const int64_t *node_off_dev,
                      const double *boxes_dev, const double *pressures_dev, double *rhs);

// bc.hip
int pf_bc_apply(pf_ctx *ctx, int64_t n, const int64_t *row_ptr, const int32_t *col_idx, double *val,
                double *rhs, const uint8_t *fixed, const double *ubar, double avg_row);
int pf_bc_fix(pf_ctx *ctx, int64_t n, int64_t nids, const int32_t *ids, const double *values,
              double fixvalue, uint8_t *fixed, double *ubar);
int pf_bc_set_fixed_x(pf_ctx *ctx, int64_t n, const uint8_t *fixed, const double *ubar, double *x);

// spmv.hip
// y = A x (accumulate: y += A x) with a subgroup of lanes per CRS row
template <typename V>
int pf_spmv_crs_launch(pf_ctx *ctx, int64_t n, const int64_t *row_ptr, const int32_t *col_idx,
                       const V *val, const V *x, V *y, double avg_row, bool accumulate = false);
// Y = A X for cols column-major vectors (leading dimension n), every column's sums those of
// pf_spmv_crs_launch bitwise, the matrix streamed once
int pf_spmm_crs_launch(pf_ctx *ctx, int64_t n, int64_t cols, const int64_t *row_ptr, const int32_t *col_idx,
                       const double *val, const double *x, double *y, double avg_row);
// CRS product with x staged in LDS per block of rpb rows: for block b the sorted distinct
// 0-based columns of its entries (ucol[ubase[b] ..], ucnt[b] of them) and for every entry the
// position of its column in that list (lidx, 16-bit).  ucnt[b] = -1 marks a block whose
// entries or distinct columns exceed the staging limits; it gathers x through col_idx.
// Depends on the pattern only (row_ptr, col_idx).
struct CrsStage {
  int64_t n = 0, nnz = 0, nb = 0;
  int sub = 0, rpb = 0;
  int64_t staged = 0, unique = 0;  // staged blocks; distinct columns summed over the blocks
  DevBuf<int32_t> ucnt, usz, ucol;
  DevBuf<int64_t> ubase, tmp;
  DevBuf<uint16_t> lidx;
  bool ready = false;
};
int pf_crs_stage_build(pf_ctx *ctx, int64_t n, const int64_t *row_ptr, const int32_t *col_idx, int64_t nnz,
                       CrsStage &st);
// y = A x (accumulate: y += A x), the sums of pf_spmv_crs_launch bitwise (same subgroup width,
// same per-lane order), x read from the staged copies
template <typename V>
int pf_spmv_crs_stage_launch(pf_ctx *ctx, const CrsStage &st, const int64_t *row_ptr, const int32_t *col_idx,
                             const V *val, const V *x, V *y, bool accumulate = false);
// lanes per CRS row for an average row length (both CRS products use it)
int pf_crs_sub(double avg_row);
int pf_block_structure_check(pf_ctx *ctx, int64_t n, const int64_t *row_ptr, const int32_t *col_idx,
                             int64_t nn, int32_t *ok_dev, int *ok_host);
// scale (may be NULL): values stored as s_i a_ij c_j, c = cscale (NULL: c = s, the symmetric form)
template <typename V>
int pf_sell3_build(pf_ctx *ctx, int64_t n, const int64_t *row_ptr, const int32_t *col_idx,
                   const double *val, Sell3<V> &s, const double *scale = nullptr,
                   const SliceMap *map = nullptr, const double *cscale = nullptr);
template <typename V>
int pf_sell3_values(pf_ctx *ctx, int64_t n, const int64_t *row_ptr, const double *val, Sell3<V> &s,
                    const double *scale = nullptr, const double *cscale = nullptr);
int pf_jacobi_scale(pf_ctx *ctx, int64_t n, const int64_t *row_ptr, const int32_t *col_idx, const double *val,
                    double *scale);
int pf_jacobi_row_scale(pf_ctx *ctx, int64_t n, const int64_t *row_ptr, const int32_t *col_idx, const double *val,
                        double *scale);  // 1/a_ii (1 where 0)
int pf_crs_to_f32(pf_ctx *ctx, int64_t n, const int64_t *row_ptr, const int32_t *col_idx, const double *val,
                  const double *scale, float *out, const double *cscale = nullptr);
template <typename V>
int pf_spmv_sell3_launch(pf_ctx *ctx, const Sell3<V> &s, const V *x, V *y);

// ilu.hip: ILU(0) factor of ILUCRS (SparseClass.f90:1629-1700) and its application
// ILU_matvecCRS (SparseClass.f90:1795-1845), level-scheduled
struct IluFactor {
  int64_t n = 0, nnz = 0, ngroups = 0, lnnz = 0;
  int G = 1;                  // rows per group: 3 = one node's rows (DOF=3 node blocks), else 1
  std::vector<int64_t> lev;   // host: level l holds level-order positions lev[l] .. lev[l+1]-1
  DevBuf<double> f;           // factor values on the CRS pattern (the reference's Pmat%val)
  DevBuf<double> dv;          // diagonal of the factor per row (0 when not stored)
  DevBuf<int32_t> grp;        // group ids in level order
  DevBuf<int64_t> lptr;       // per row in level order: its entries in columns of earlier groups
  DevBuf<int32_t> lcol;       //   0-based columns
  DevBuf<double> lval;        //   factor values
  DevBuf<double> wv;          // G = 3: within-node lower entries (1,0), (2,0), (2,1) per position
  DevBuf<double> dd;          // per row in level order: the divisor of the apply (0 -> 1)
  DevBuf<double> t;           // n: L^-1 src before the division
  DevBuf<int32_t> err;
  bool ready = false;
};
// inputs of one sweep y = D^-1 L^-1 src (mode 0: src = v; 1: src = p = r + beta (p - omega y),
// stored into pe; 2: src = e = r - alpha y, stored into pe); scalars are device pointers
struct IluSweepArgs {
  const double *v = nullptr, *r = nullptr, *y = nullptr;
  double *pe = nullptr;
  const double *alpha = nullptr, *beta = nullptr, *omega = nullptr;
  const int32_t *done = nullptr;  // nonzero: the sweep returns at entry
  double *out = nullptr;
};
// factor = true: F.f = ILU(0) of val; false: val already is a factor (schedule and sweep
// structures only).  G = 3 only for a CRS with the DOF=3 node-block structure.
int pf_ilu_factor(pf_ctx *ctx, int64_t n, const int64_t *row_ptr, const int32_t *col_idx, const double *val,
                  int64_t nnz, int G, IluFactor &F, bool factor = true);
int pf_ilu_sweep_enqueue(pf_ctx *ctx, IluFactor &F, int mode, const IluSweepArgs &a);
// JacobiPreconditionerCRS in place (rhs may be NULL)
int pf_jacobi_rows(pf_ctx *ctx, int64_t n, const int64_t *row_ptr, const int32_t *col_idx, double *val,
                   double *rhs);

// bicgstab.hip
struct SolveResult {
  int status = 0, iterations = 0, outer = 0, kind = 0;
  double rr = 0, rr1 = 0, true_relres = 0, solve_ms = 0, setup_ms = 0;
  double t_ms[4] = {0, 0, 0, 0};  // kernel_timing averages: spmv_p, spmv_e, update, pupdate
  int64_t timed = 0;
};
// SolverWork: device workspace, defined in bicgstab.hip
int pf_solver_work_create(SolverWork **w);
void pf_solver_work_destroy(SolverWork *w);
// Solve A x = b on the device; A given as CRS (always) and optionally as blocked layout
// built inside.  x holds x0 on entry.
int pf_bicgstab_device(pf_ctx *ctx, SolverWork *w, int64_t n, const int64_t *row_ptr,
                       const int32_t *col_idx, const double *val, int64_t nnz, const double *b,
                       double *x, const pf_solve_opts &o, bool matrix_changed, SolveResult &res);
// bicgstab_CRS_2 (or the stop rules of stop_mode 1 and 2) in FP64 on the CRS layout with every
// work vector, partial sum and scalar in the caller's device workspace ws (ws_doubles doubles,
// at least pf_bicgstab_workspace_doubles(ctx, n)): no device allocation, no hipGraph; host polls
// go to a stack copy.  Every vector the kernels read is written first within the call.
size_t pf_bicgstab_workspace_doubles(pf_ctx *ctx, int64_t n);
// noalloc.hip: look up the device address of the allocation-free solver's scratch array on
// `device` (current), once per device
int pf_noalloc_setup(int device);
int pf_bicgstab_crs_workspace(pf_ctx *ctx, int64_t n, const int64_t *row_ptr, const int32_t *col_idx,
                              const double *val, int64_t nnz, const double *b, double *x, const pf_solve_opts &o,
                              double *ws, size_t ws_doubles, SolveResult &res);
// y = A x on device vectors with layout kind (1 CRS FP64, 2 blocked FP64, 3 blocked FP32);
// `repeat` timed launches, *ms = average device ms per launch
int pf_work_spmv(pf_ctx *ctx, SolverWork *w, int64_t n, const int64_t *row_ptr, const int32_t *col_idx,
                 const double *val, int64_t nnz, bool matrix_changed, int kind, const double *x, double *y,
                 int repeat, double *ms);

// the ILU(0) factor kept in w (computed from val when none is kept or the sizes changed)
int pf_work_ilu(pf_ctx *ctx, SolverWork *w, int64_t n, const int64_t *row_ptr, const int32_t *col_idx,
                const double *val, int64_t nnz, bool matrix_changed, IluFactor **out);
void pf_work_ilu_reset(SolverWork *w);

// eig.hip: eigFEMSolver; fixed_ids 1-based (removed DOFs), nev <= 0: every pair (dense),
// else the nev lowest (LOBPCG).  values (count), vectors (n x count, column-major) on the host.
struct EigResult {
  int status = 0, iterations = 0, nconv = 0;
  double max_res = 0.0, ms = 0.0;
};
int pf_eig_device(pf_ctx *ctx, int64_t n, const int64_t *row_ptr, const int32_t *col_idx, const double *A,
                  const double *B, int64_t nnz, const std::vector<int32_t> &fixed_ids, int32_t nev, double tol,
                  int32_t maxit, double *values, double *vectors, int32_t *count, EigResult &res);

// batch.hip: many independent meshes of one block-diagonal system, one persistent kernel
struct BatchWork;
void pf_batch_work_destroy(BatchWork *b);
// DOF offsets (3 x node offsets, nmesh + 1 entries) of a configured batch
int pf_batch_dof_offsets(const BatchWork *b, std::vector<int64_t> &off);

// direct.hip: banded Cholesky direct solve per mesh (solve(algorithm="GaussJordan"))
struct DirectWork;
void pf_direct_work_destroy(DirectWork *d);
void pf_direct_invalidate(DirectWork *d);
// ordering and band layout of the meshes dof_off[m] .. dof_off[m+1]-1 (host offsets); node
// orderings when the CRS has DOF=3 node blocks (coords: nn x 3 device coordinates or NULL),
// DOF orderings otherwise; kept while pattern_key, n and the offsets are unchanged
int pf_direct_analyse(pf_ctx *ctx, DirectWork **pd, int64_t n, const int64_t *row_ptr, const int32_t *col_idx,
                      int32_t nmesh, const int64_t *dof_off, const double *coords_dev, int64_t ncoord,
                      uint64_t pattern_key);
// band assembly (when mat_version differs from the held factor), factorization, substitution,
// `refine` refinement steps and per-mesh FP64 residuals; x receives the solution
int pf_direct_solve_device(pf_ctx *ctx, DirectWork *D, int64_t n, const int64_t *row_ptr, const int32_t *col_idx,
                           const double *val, int64_t nnz, const double *rhs, double *x, uint64_t mat_version,
                           int32_t refine, int32_t mfma, pf_batch_mesh_stats *mesh_stats, pf_stats *total);
int pf_direct_info(DirectWork *D, int32_t mesh, int64_t *n, int32_t *bandwidth, int32_t *kind, int32_t *perm);
int pf_direct_factor_get(pf_ctx *ctx, DirectWork *D, int32_t mesh, double *ab);
// node_off (host, nmesh+1): mesh m owns nodes node_off[m] .. node_off[m+1]-1; cost (host,
// may be NULL) orders the queue, largest first (default: the mesh's nonzeros)
int pf_batch_configure(pf_ctx *ctx, BatchWork **b, int64_t n, const int64_t *row_ptr, const int32_t *col_idx,
                       int32_t nmesh, const int64_t *node_off, const double *cost);
int pf_batch_solve_device(pf_ctx *ctx, BatchWork *b, const int64_t *row_ptr, const int32_t *col_idx,
                          const double *val, int64_t nnz, const double *rhs, double *x, const pf_solve_opts &o,
                          const pf_batch_tune *tune, uint64_t mat_version, pf_batch_mesh_stats *mesh_stats,
                          pf_stats *total);
int pf_batch_autotune_device(pf_ctx *ctx, BatchWork *b, const int64_t *row_ptr, const int32_t *col_idx,
                             const double *val, int64_t nnz, const double *rhs, double *x, const pf_solve_opts &o,
                             int32_t iters, uint64_t mat_version, pf_batch_tune *best, int32_t max_results,
                             pf_batch_tune *tried, double *tried_ms, int32_t *ntried);

// multi.hip: one matrix, many right-hand sides (interleaved, dof i / case l at i*R + l)
struct MultiWork;
void pf_multi_work_destroy(MultiWork *m);
int pf_multi_solve_device(pf_ctx *ctx, MultiWork **w, int64_t n, const int64_t *row_ptr, const int32_t *col_idx,
                          const double *val, uint64_t mat_version, const uint8_t *fixed, int32_t nrhs,
                          const double *B_dev, double *X_dev, const pf_solve_opts &o, pf_stats *per_rhs,
                          pf_stats *total);

// hipGraphs: pf_graphs_allowed() is false when the process runs under a profiler that
// intercepts the HSA queues (rocprofv3 sets ROCPROFILER_LIBRARY_CTOR and preloads its tool
// library) or when PF_NO_GRAPH is set; callers then launch the same kernels directly.
bool pf_graphs_allowed();

// Capture enqueue() (which returns a PF_* status) on ctx->stream into *exec.  The captured
// graph is destroyed on every path.  Returns the enqueue's error, or PF_ERR_HIP when the
// capture itself failed; *exec stays null when instantiation fails (the caller then
// launches directly).
template <typename F>
int pf_capture_graph(pf_ctx *ctx, F &&enqueue, hipGraphExec_t *exec) {
  *exec = nullptr;
  hipGraph_t graph = nullptr;
  PF_HIP(hipStreamBeginCapture(ctx->stream, hipStreamCaptureModeThreadLocal));
  const int rc = enqueue();
  const hipError_t ce = hipStreamEndCapture(ctx->stream, &graph);
  if (rc != PF_OK || ce != hipSuccess) {
    if (graph) (void)hipGraphDestroy(graph);
    if (rc != PF_OK) return rc;
    pf_set_error("hipStreamEndCapture: %s", hipGetErrorString(ce));
    return PF_ERR_HIP;
  }
  const hipError_t ie = hipGraphInstantiate(exec, graph, nullptr, nullptr, 0);
  (void)hipGraphDestroy(graph);
  if (ie != hipSuccess) *exec = nullptr;
  return PF_OK;
}

// common launch geometry
static inline unsigned pf_blocks(int64_t work, int threads, int64_t cap = 1 << 20) {
  int64_t b = (work + threads - 1) / threads;
  if (b < 1) b = 1;
  if (b > cap) b = cap;
  return (unsigned)b;
}
