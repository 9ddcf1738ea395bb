// batch.hip -- many independent static solves in one persistent kernel.
//
// A batch is a block-diagonal system: the union of several independent meshes (every
// element of mesh m references nodes of mesh m only), assembled by the ordinary
// ZeroMatrix / StiffnessMatrix / traction / fix kernels, so the CRS block of each mesh is
// exactly the one the reference builds for that mesh alone (column ids shifted by the
// mesh's first DOF).  It stands for solving such meshes one after another with one
// single-domain FEMSolver_ each; the reference's whole-plant Soybean%getDisplacement
// (SoybeanClass.f90:8417-8533) couples organs through overset penalties instead, which
// this form does not do.
//
// Here every mesh runs bicgstab_CRS_2 (FEMSolverClass.f90:2199-2298) inside ONE launch of a
// persistent kernel: a workgroup takes the next mesh from a queue (largest cost first),
// runs that mesh's whole BiCGSTAB with workgroup-level reductions (wavefront shuffles + LDS,
// bitwise the same scalars in every thread), records the result and takes the next mesh
// until the queue is empty.  No workgroup ever waits for another, so the kernel needs no
// fences, no grid barrier and no co-residency.  The queue reset and the kernel are captured
// in one hipGraph.  Vectors of mesh m live in the DOF range of the mesh inside the union
// vectors, so the gathers of a mesh stay inside its own rows.
//
// Launch shapes (chosen by the autotuner): the blocked Sell3 layout with 64-node slices, or
// CRS rows reduced by SUB-lane subgroups ("subwarp width"); BTB threads per workgroup, i.e.
// BTB/64 slices or BTB/SUB rows per pass of the workgroup ("rows per block").
//
// precision 32: the FP32 recurrence on D^-1 A restarted at every FP64 replacement (opts.jacobi
// 2, the default), or on S A S (S = diag(A)^-1/2, jacobi 1) or A (jacobi 0) with reliable
// updates (FP64 solution accumulation, FP64 true-residual replacement, FP64 dot products):
// the per-mesh form of the mixed solver in bicgstab.hip.
#include <algorithm>
#include <cmath>
#include <cstring>
#include <numeric>
#include <vector>

#include "spmv_dev.hpp"

namespace {

struct BMesh {
  int64_t d0, d1;  // DOF rows d0 .. d1-1 of the union system
  int64_t s0, s1;  // Sell3 slices s0 .. s1-1
};

template <typename V>
struct BMat {
  const int64_t *row_ptr;
  const int32_t *col_idx;
  const V *val;  // CRS values (FP64 system or FP32 scaled copy)
  const int64_t *soff;
  const int32_t *width;
  const int32_t *col;
  const V *sval;
  const int64_t *snode, *send;
};

struct BArgs {
  int32_t nmesh, itrmax, stop_mode, inner_max;
  double er0, rel, target, delta;
  const BMesh *mesh;
  const int32_t *order;
  int32_t *queue;
  BMat<double> A;
  BMat<float> A32;
  const double *b;
  double *x;
  double *r, *r0, *p, *y, *v;
  float *r32, *r032, *p32, *y32, *v32, *zlo;
  double *zhi;
  const double *scale;   // FP32 system row scale
  const double *cscale;  // FP32 system column scale (the unknown is x = C z)
  int64_t restart;       // replacements restart the FP32 recurrence (r0 = p = r)
  pf_batch_mesh_stats *stat;
  int64_t *t_raw;  // per mesh: start, end wall-clock ticks
};

template <int BTB, int NV>
__device__ inline void wg_sum(double (&v)[NV], double (*sh)[BTB / 64]) {
#pragma unroll
  for (int k = 0; k < NV; ++k)
#pragma unroll
    for (int off = 32; off > 0; off >>= 1) v[k] += __shfl_xor(v[k], off, 64);
  const int w = threadIdx.x >> 6;
  if ((threadIdx.x & 63) == 0)
#pragma unroll
    for (int k = 0; k < NV; ++k) sh[k][w] = v[k];
  __syncthreads();
#pragma unroll
  for (int k = 0; k < NV; ++k) {
    double t = 0.0;
#pragma unroll
    for (int j = 0; j < BTB / 64; ++j) t += sh[k][j];
    v[k] = t;
  }
  __syncthreads();
}

// every row of the mesh's block of y = A*xf, one call op(row, sum) per row
template <typename V, int L, int SUB, int BTB, typename XF, typename OP>
__device__ inline void mesh_rows(const BMat<V> &A, const BMesh &bm, XF xf, OP op) {
  if (L == 1) {
    const int wid = threadIdx.x >> 6, lane = threadIdx.x & 63;
    for (int64_t s = bm.s0 + wid; s < bm.s1; s += BTB / 64) {
      V y0, y1, y2;
      pfdev::sell3_node<V, V>(A.soff, A.width, A.col, A.sval, s, lane, xf, y0, y1, y2);
      const int64_t i = A.snode[s] + lane;
      if (i < A.send[s]) {
        op(3 * i, y0);
        op(3 * i + 1, y1);
        op(3 * i + 2, y2);
      }
    }
  } else {
    const int sub = threadIdx.x % SUB;
    for (int64_t row = bm.d0 + threadIdx.x / SUB; row < bm.d1; row += BTB / SUB) {
      const V s = pfdev::crs_row<V, V, SUB>(A.row_ptr, A.col_idx, A.val, row, sub, xf);
      if (sub == 0) op(row, s);
    }
  }
}

template <typename V>
struct EFromRY {
  const V *__restrict__ r;
  const V *__restrict__ y;
  V alpha;
  __device__ V operator()(int64_t j) const { return r[j] - alpha * y[j]; }
};

struct ScaledX {
  const double *__restrict__ scale;
  const double *__restrict__ z;
  __device__ double operator()(int64_t j) const { return scale[j] * z[j]; }
};

struct MeshResult {
  int status, iterations, replacements;
  double rr, rr1;
};

// FP64 bicgstab_CRS_2 on one mesh, x0 = 0
template <int L, int SUB, int BTB>
__device__ MeshResult solve64(const BArgs &a, const BMesh &bm, double (*sh)[BTB / 64]) {
  const int64_t d0 = bm.d0, d1 = bm.d1;
  double t1[1] = {0.0};
  for (int64_t i = d0 + threadIdx.x; i < d1; i += BTB) {
    const double bv = a.b[i];
    a.x[i] = 0.0;
    a.r[i] = bv;
    a.r0[i] = bv;
    a.p[i] = bv;
    t1[0] += bv * bv;
  }
  wg_sum<BTB, 1>(t1, sh);
  const double bnorm2 = t1[0];
  double c1 = t1[0], rr = t1[0], rr1 = 0.0;
  MeshResult res{PF_NOT_CONVERGED, 0, 0, rr, 0.0};
  if (!(bnorm2 > 0.0)) {  // b = 0: x = 0 solves it (the reference's recurrence would divide 0 by 0)
    res.status = PF_OK;
    return res;
  }
  if (a.itrmax <= 0) return res;
  int itr = 0;
  for (;;) {
    // y = A p ; c2 = r0.y
    double q1[1] = {0.0};
    mesh_rows<double, L, SUB, BTB>(a.A, bm, pfdev::PlainX<double>{a.p}, [&](int64_t row, double yv) {
      a.y[row] = yv;
      q1[0] += a.r0[row] * yv;
    });
    wg_sum<BTB, 1>(q1, sh);
    const double c2 = q1[0];
    const double alpha = c1 / c2;
    // v = A e, e = r - alpha y ; e.v, v.v
    const EFromRY<double> ef{a.r, a.y, alpha};
    double q2[2] = {0.0, 0.0};
    mesh_rows<double, L, SUB, BTB>(a.A, bm, ef, [&](int64_t row, double vv) {
      a.v[row] = vv;
      const double e = ef(row);
      q2[0] += e * vv;
      q2[1] += vv * vv;
    });
    wg_sum<BTB, 2>(q2, sh);
    if (q2[1] == 0.0) {  // FEMSolverClass.f90:2266, inside iteration itr+1
      res.status = PF_BREAKDOWN_VV;
      itr += 1;
      break;
    }
    const double omega = q2[0] / q2[1];
    // x += alpha p + omega e ; r = e - omega v ; r.r, r0.r
    double q3[2] = {0.0, 0.0};
    for (int64_t i = d0 + threadIdx.x; i < d1; i += BTB) {
      const double e = a.r[i] - alpha * a.y[i];
      a.x[i] = a.x[i] + alpha * a.p[i] + omega * e;
      const double rn = e - omega * a.v[i];
      a.r[i] = rn;
      q3[0] += rn * rn;
      q3[1] += a.r0[i] * rn;
    }
    wg_sum<BTB, 2>(q3, sh);
    rr = q3[0];
    itr += 1;
    if (itr == 1) rr1 = rr;
    const bool stop = a.stop_mode == 0 ? (sqrt(rr / rr1) < a.rel || sqrt(rr) < a.er0)
                                       : sqrt(rr) < a.rel * sqrt(bnorm2);
    if (stop) {
      res.status = PF_OK;
      break;
    }
    const double c2w = c2 * omega;
    const double beta = q3[1] / c2w;
    if (c2w == 0.0 || !isfinite(beta) || !isfinite(rr)) {  // FEMSolverClass.f90:2295
      res.status = PF_BREAKDOWN_C2C3;
      break;
    }
    if (itr >= a.itrmax) break;
    c1 = q3[1];
    for (int64_t i = d0 + threadIdx.x; i < d1; i += BTB) a.p[i] = a.r[i] + beta * (a.p[i] - omega * a.y[i]);
    __syncthreads();
  }
  res.iterations = itr;
  res.rr = rr;
  res.rr1 = rr1;
  return res;
}

// FP32 recurrence on S A S with FP64 reliable updates on one mesh, x0 = 0
template <int L, int SUB, int BTB>
__device__ MeshResult solve32(const BArgs &a, const BMesh &bm, double (*sh)[BTB / 64]) {
  const int64_t d0 = bm.d0, d1 = bm.d1;
  // z_hi = z_lo = 0: the first true residual S (b - A S z_hi) is S b
  double t2[2] = {0.0, 0.0};
  for (int64_t i = d0 + threadIdx.x; i < d1; i += BTB) {
    const double bv = a.b[i];
    const float rf = (float)(a.scale[i] * bv);
    a.zhi[i] = 0.0;
    a.zlo[i] = 0.0f;
    a.r32[i] = rf;
    a.r032[i] = rf;
    a.p32[i] = rf;
    t2[0] += (double)rf * (double)rf;
    t2[1] += bv * bv;
  }
  wg_sum<BTB, 2>(t2, sh);
  const double bn = sqrt(t2[1]);
  double c1 = t2[0], rr = t2[0];
  double maxr = sqrt(t2[0]);
  MeshResult res{PF_NOT_CONVERGED, 0, 0, rr, 0.0};
  if (!(bn > 0.0)) {  // b = 0: x = 0 solves it
    res.status = PF_OK;
    return res;
  }
  if (a.itrmax <= 0) return res;
  int itr = 0, itr_rep = 0, nupd = 0;
  for (;;) {
    double q1[1] = {0.0};
    mesh_rows<float, L, SUB, BTB>(a.A32, bm, pfdev::PlainX<float>{a.p32}, [&](int64_t row, float yv) {
      a.y32[row] = yv;
      q1[0] += (double)a.r032[row] * (double)yv;
    });
    wg_sum<BTB, 1>(q1, sh);
    const double c2 = q1[0];
    const double alpha_d = c1 / c2;
    if (!isfinite(alpha_d)) {
      res.status = PF_BREAKDOWN_C2C3;
      break;
    }
    const float alpha = (float)alpha_d;
    const EFromRY<float> ef{a.r32, a.y32, alpha};
    double q2[2] = {0.0, 0.0};
    mesh_rows<float, L, SUB, BTB>(a.A32, bm, ef, [&](int64_t row, float vv) {
      a.v32[row] = vv;
      const float e = ef(row);
      q2[0] += (double)e * (double)vv;
      q2[1] += (double)vv * (double)vv;
    });
    wg_sum<BTB, 2>(q2, sh);
    if (q2[1] == 0.0 || !isfinite(q2[0] / q2[1])) {
      res.status = PF_BREAKDOWN_VV;
      itr += 1;
      break;
    }
    const double omega_d = q2[0] / q2[1];
    const float omega = (float)omega_d;
    double q3[2] = {0.0, 0.0};
    for (int64_t i = d0 + threadIdx.x; i < d1; i += BTB) {
      const float e = a.r32[i] - alpha * a.y32[i];
      a.zlo[i] = a.zlo[i] + alpha * a.p32[i] + omega * e;
      const float rn = e - omega * a.v32[i];
      a.r32[i] = rn;
      q3[0] += (double)rn * (double)rn;
      q3[1] += (double)a.r032[i] * (double)rn;
    }
    wg_sum<BTB, 2>(q3, sh);
    rr = q3[0];
    itr += 1;
    const double nr = sqrt(rr);
    maxr = fmax(maxr, nr);
    const bool replace = nr < a.delta * maxr || (a.inner_max > 0 && itr - itr_rep >= a.inner_max);
    const double c2w = c2 * omega_d;
    const double beta_d = q3[1] / c2w;
    if (c2w == 0.0 || !isfinite(beta_d) || !isfinite(rr)) {
      res.status = PF_BREAKDOWN_C2C3;
      break;
    }
    if (itr >= a.itrmax) break;
    c1 = q3[1];
    const float beta = (float)beta_d;
    if (replace) {
      for (int64_t i = d0 + threadIdx.x; i < d1; i += BTB) {
        a.p32[i] = a.r32[i] + beta * (a.p32[i] - omega * a.y32[i]);
        a.zhi[i] += (double)a.zlo[i];
        a.zlo[i] = 0.0f;
      }
      __syncthreads();
      // r = S (b - A C z_hi) in FP64 -> FP32 ; |r|^2, |b - A x|^2, r0.r (restart: r0 = p = r)
      double q4[3] = {0.0, 0.0, 0.0};
      mesh_rows<double, L, SUB, BTB>(a.A, bm, ScaledX{a.cscale, a.zhi}, [&](int64_t row, double ax) {
        const double bu = a.b[row] - ax;
        const float rf = (float)(a.scale[row] * bu);
        a.r32[row] = rf;
        if (a.restart) {
          a.r032[row] = rf;
          a.p32[row] = rf;
        }
        q4[0] += (double)rf * (double)rf;
        q4[1] += bu * bu;
        q4[2] += (double)a.r032[row] * (double)rf;
      });
      wg_sum<BTB, 3>(q4, sh);
      nupd += 1;
      itr_rep = itr;
      c1 = q4[2];
      rr = q4[0];
      maxr = sqrt(q4[0]);
      const double relres = sqrt(q4[1]) / bn;
      if (!(relres >= a.target)) {
        res.status = isfinite(relres) ? PF_OK : PF_BREAKDOWN_C2C3;
        break;
      }
    } else {
      for (int64_t i = d0 + threadIdx.x; i < d1; i += BTB) a.p32[i] = a.r32[i] + beta * (a.p32[i] - omega * a.y32[i]);
      __syncthreads();
    }
  }
  // x = C (z_hi + z_lo)
  for (int64_t i = d0 + threadIdx.x; i < d1; i += BTB) a.x[i] = a.cscale[i] * (a.zhi[i] + (double)a.zlo[i]);
  res.iterations = itr;
  res.replacements = nupd;
  res.rr = rr;
  return res;
}

__global__ void k_queue_reset(int32_t *q) {
  if (threadIdx.x == 0 && blockIdx.x == 0) q[0] = 0;
}

template <int PREC, int L, int SUB, int BTB>
__global__ void __launch_bounds__(BTB) k_batch(BArgs a) {
  __shared__ double sh[4][BTB / 64];
  __shared__ int32_t next;
  for (;;) {
    __syncthreads();
    if (threadIdx.x == 0) next = atomicAdd(a.queue, 1);
    __syncthreads();
    const int32_t q = next;
    if (q >= a.nmesh) break;
    const int32_t m = a.order[q];
    const BMesh bm = a.mesh[m];
    const int64_t t0 = wall_clock64();
    MeshResult res = PREC == 64 ? solve64<L, SUB, BTB>(a, bm, sh) : solve32<L, SUB, BTB>(a, bm, sh);
    __syncthreads();
    // FP64 true residual |b - A x| / |b| of the mesh
    double t[2] = {0.0, 0.0};
    mesh_rows<double, L, SUB, BTB>(a.A, bm, pfdev::PlainX<double>{a.x}, [&](int64_t row, double ax) {
      const double rv = a.b[row] - ax;
      t[0] += rv * rv;
      t[1] += a.b[row] * a.b[row];
    });
    wg_sum<BTB, 2>(t, sh);
    if (threadIdx.x == 0) {
      pf_batch_mesh_stats st;
      st.status = res.status;
      st.iterations = res.iterations;
      st.replacements = res.replacements;
      st.workgroup = (int32_t)blockIdx.x;
      st.rr = res.rr;
      st.rr1 = res.rr1;
      st.true_relres = t[1] > 0.0 ? sqrt(t[0] / t[1]) : sqrt(t[0]);
      st.start_ms = 0.0;
      st.end_ms = 0.0;
      a.stat[m] = st;
      a.t_raw[2 * m] = t0;
      a.t_raw[2 * m + 1] = wall_clock64();
    }
  }
}

// rows of mesh m must reference columns of mesh m only (block-diagonal union)
__global__ void k_block_diag_check(int64_t n, const int64_t *__restrict__ row_ptr,
                                   const int32_t *__restrict__ col_idx, int32_t nmesh,
                                   const BMesh *__restrict__ mesh, int32_t *__restrict__ bad) {
  const int64_t row = blockIdx.x * (int64_t)blockDim.x + threadIdx.x;
  if (row >= n) return;
  int32_t lo = 0, hi = nmesh;  // largest m with mesh[m].d0 <= row
  while (hi - lo > 1) {
    const int32_t mid = (lo + hi) >> 1;
    if (mesh[mid].d0 <= row) lo = mid; else hi = mid;
  }
  const int64_t c0 = mesh[lo].d0 + 1, c1 = mesh[lo].d1;  // 1-based column range
  bool ok = true;
  for (int64_t k = row_ptr[row] - 1; k < row_ptr[row + 1] - 1 && ok; ++k) ok = col_idx[k] >= c0 && col_idx[k] <= c1;
  if (!ok) atomicOr(bad, 1);
}

// mapped Sell3 slices pad with the slice's own first node, so padded products read a
// finite entry of the same mesh and never another mesh's vector
__global__ void k_bsell_pad(int64_t nslices, const int64_t *__restrict__ soff, const int64_t *__restrict__ snode,
                            int32_t *__restrict__ col, int64_t ncells) {
  for (int64_t g = blockIdx.x * (int64_t)blockDim.x + threadIdx.x; g < ncells; g += (int64_t)gridDim.x * blockDim.x) {
    if (col[g] >= 0) continue;
    const int64_t slot = g >> 6;
    int64_t lo = 0, hi = nslices;
    while (hi - lo > 1) {
      const int64_t mid = (lo + hi) >> 1;
      if (soff[mid] <= slot) lo = mid; else hi = mid;
    }
    col[g] = (int32_t)snode[lo];
  }
}

}  // namespace

// ========================================================================================
struct BatchWork {
  int32_t nmesh = 0;
  int64_t n = 0;
  std::vector<int64_t> node_off;
  std::vector<double> cost;
  std::vector<BMesh> hmesh;
  DevBuf<BMesh> mesh;
  DevBuf<int32_t> order, queue, flag;
  DevBuf<int64_t> snode, send, t_raw;
  int64_t nslices = 0;
  Sell3<double> s64;
  Sell3<float> s32;
  DevBuf<float> val32;
  DevBuf<double> scale, ones;
  DevBuf<double> r, r0, p, y, v, zhi;
  DevBuf<float> r32, r032, p32, y32, v32, zlo;
  DevBuf<pf_batch_mesh_stats> stat;
  bool block_ok = false, have_s64 = false, have_s32 = false, have_v32 = false, have_scale = false;
  int scaled = 0;  // the FP32 copies hold 0: A, 1: S A S (S = diag(A)^-1/2), 2: D^-1 A
  uint64_t mat_version = ~0ull;
  hipGraphExec_t gexec = nullptr;
  // what the captured launch bakes in: the kernel argument block (no padding: every member
  // after the four int32 is 8 bytes wide), the launch shape and the grid
  struct GraphKey {
    BArgs a;
    pf_batch_tune t;
    uint32_t grid;
    int32_t mixed;
  } gkey;
  bool gkey_valid = false;
  pf_batch_tune tuned[2] = {};  // [0] FP64, [1] mixed
  bool have_tuned[2] = {false, false};
  ~BatchWork() {
    if (gexec) (void)hipGraphExecDestroy(gexec);
  }
};

void pf_batch_work_destroy(BatchWork *b) { delete b; }

int pf_batch_dof_offsets(const BatchWork *b, std::vector<int64_t> &off) {
  if (!b || b->nmesh < 1) {
    pf_set_error("batch: call pf_system_batch_set first");
    return PF_INVALID;
  }
  off.resize(b->node_off.size());
  for (size_t i = 0; i < off.size(); ++i) off[i] = 3 * b->node_off[i];
  return PF_OK;
}

int pf_batch_configure(pf_ctx *ctx, BatchWork **pb, int64_t n, const int64_t *row_ptr, const int32_t *col_idx,
                       int32_t nmesh, const int64_t *node_off, const double *cost) {
  if (nmesh < 1 || !node_off || n % 3 != 0) {
    pf_set_error("batch: nmesh >= 1, node offsets and a DOF=3 system required");
    return PF_INVALID;
  }
  const int64_t nn = n / 3;
  if (node_off[0] != 0 || node_off[nmesh] != nn) {
    pf_set_error("batch: node offsets must start at 0 and end at the node count %lld", (long long)nn);
    return PF_INVALID;
  }
  for (int32_t m = 0; m < nmesh; ++m)
    if (node_off[m + 1] <= node_off[m]) {
      pf_set_error("batch: mesh %d has no nodes (offsets must increase)", m);
      return PF_INVALID;
    }
  if (cost)
    for (int32_t m = 0; m < nmesh; ++m)
      if (!(cost[m] >= 0.0)) {
        pf_set_error("batch: cost of mesh %d is not a non-negative number", m);
        return PF_INVALID;
      }
  if (!*pb) *pb = new BatchWork();
  BatchWork *b = *pb;
  b->nmesh = 0;  // unusable until every check below has passed
  b->n = n;
  b->node_off.assign(node_off, node_off + nmesh + 1);
  b->have_s64 = b->have_s32 = b->have_v32 = b->have_scale = false;
  b->mat_version = ~0ull;
  b->gkey_valid = false;
  // slices of 64 nodes that never cross a mesh
  std::vector<int64_t> sn, se;
  b->hmesh.resize(nmesh);
  for (int32_t m = 0; m < nmesh; ++m) {
    BMesh &bm = b->hmesh[m];
    bm.d0 = 3 * node_off[m];
    bm.d1 = 3 * node_off[m + 1];
    bm.s0 = (int64_t)sn.size();
    for (int64_t a = node_off[m]; a < node_off[m + 1]; a += 64) {
      sn.push_back(a);
      se.push_back(std::min(a + 64, node_off[m + 1]));
    }
    bm.s1 = (int64_t)sn.size();
  }
  b->nslices = (int64_t)sn.size();
  // queue order: largest cost first (default cost: the mesh's nonzeros)
  std::vector<int64_t> hrp(nmesh + 1);
  b->cost.assign(nmesh, 0.0);
  for (int32_t m = 0; m < nmesh; ++m) {
    if (cost) {
      b->cost[m] = cost[m];
    } else {
      int64_t a = 0, e = 0;
      PF_CHECK(pf_read_i64(ctx, row_ptr + b->hmesh[m].d0, &a));
      PF_CHECK(pf_read_i64(ctx, row_ptr + b->hmesh[m].d1, &e));
      b->cost[m] = (double)(e - a);
    }
  }
  std::vector<int32_t> ord(nmesh);
  std::iota(ord.begin(), ord.end(), 0);
  std::stable_sort(ord.begin(), ord.end(), [&](int32_t x, int32_t y) { return b->cost[x] > b->cost[y]; });
  PF_CHECK(b->mesh.ensure(nmesh));
  PF_CHECK(b->order.ensure(nmesh));
  PF_CHECK(b->queue.ensure(1));
  PF_CHECK(b->flag.ensure(1));
  PF_CHECK(b->snode.ensure(b->nslices));
  PF_CHECK(b->send.ensure(b->nslices));
  PF_CHECK(b->t_raw.ensure(2 * (size_t)nmesh));
  PF_CHECK(b->stat.ensure(nmesh));
  PF_HIP(hipMemcpyAsync(b->mesh.p, b->hmesh.data(), sizeof(BMesh) * nmesh, hipMemcpyHostToDevice, ctx->stream));
  PF_HIP(hipMemcpyAsync(b->order.p, ord.data(), sizeof(int32_t) * nmesh, hipMemcpyHostToDevice, ctx->stream));
  PF_HIP(hipMemcpyAsync(b->snode.p, sn.data(), sizeof(int64_t) * b->nslices, hipMemcpyHostToDevice, ctx->stream));
  PF_HIP(hipMemcpyAsync(b->send.p, se.data(), sizeof(int64_t) * b->nslices, hipMemcpyHostToDevice, ctx->stream));
  // block-diagonal check and DOF=3 node blocks (needed by the blocked layout)
  PF_HIP(hipMemsetAsync(b->flag.p, 0, sizeof(int32_t), ctx->stream));
  k_block_diag_check<<<pf_blocks(n, 256, INT32_MAX), 256, 0, ctx->stream>>>(n, row_ptr, col_idx, nmesh, b->mesh.p,
                                                                           b->flag.p);
  PF_HIP(hipGetLastError());
  int32_t bad = 1;
  PF_HIP(hipMemcpyAsync(&bad, b->flag.p, sizeof(int32_t), hipMemcpyDeviceToHost, ctx->stream));
  PF_HIP(hipStreamSynchronize(ctx->stream));
  if (bad) {
    pf_set_error("batch: the system is not block-diagonal over the given node offsets");
    return PF_INVALID;
  }
  int ok = 0;
  PF_CHECK(pf_block_structure_check(ctx, n, row_ptr, col_idx, nn, b->flag.p, &ok));
  b->block_ok = ok != 0;
  b->have_tuned[0] = b->have_tuned[1] = false;
  b->nmesh = nmesh;
  return PF_OK;
}

namespace {

template <int PREC, int L, int SUB, int BTB>
int launch_one(pf_ctx *ctx, const BArgs &a, unsigned grid) {
  k_batch<PREC, L, SUB, BTB><<<grid, BTB, 0, ctx->stream>>>(a);
  PF_HIP(hipGetLastError());
  return PF_OK;
}

template <int PREC, int L, int SUB, int BTB>
int occupancy(int *blocks) {
  PF_HIP(hipOccupancyMaxActiveBlocksPerMultiprocessor(blocks, k_batch<PREC, L, SUB, BTB>, BTB, 0));
  return PF_OK;
}

// dispatch over the tuned launch shapes: (layout 1, any SUB) or (layout 0, SUB 8/16/32),
// BTB 256/512/1024
template <int PREC, typename F>
int dispatch(const pf_batch_tune &t, F f) {
#define PF_B_BTB(L_, S_)                                              \
  switch (t.block) {                                                  \
    case 256: return f.template operator()<PREC, L_, S_, 256>();      \
    case 512: return f.template operator()<PREC, L_, S_, 512>();      \
    case 1024: return f.template operator()<PREC, L_, S_, 1024>();    \
    default: break;                                                   \
  }
  if (t.layout == 1) {
    PF_B_BTB(1, 64)
  } else if (t.layout == 0) {
    switch (t.sub) {
      case 8: PF_B_BTB(0, 8) break;
      case 16: PF_B_BTB(0, 16) break;
      case 32: PF_B_BTB(0, 32) break;
      default: break;
    }
  }
#undef PF_B_BTB
  pf_set_error("batch: unsupported launch shape layout %d sub %d block %d", t.layout, t.sub, t.block);
  return PF_INVALID;
}

struct Launcher {
  pf_ctx *ctx;
  const BArgs *a;
  unsigned grid;
  template <int PREC, int L, int SUB, int BTB>
  int operator()() const {
    return launch_one<PREC, L, SUB, BTB>(ctx, *a, grid);
  }
};

struct Occ {
  int *blocks;
  template <int PREC, int L, int SUB, int BTB>
  int operator()() const {
    return occupancy<PREC, L, SUB, BTB>(blocks);
  }
};

int valid_tune(const pf_batch_tune &t, bool block_ok) {
  const bool shape_ok = (t.block == 256 || t.block == 512 || t.block == 1024) &&
                        ((t.layout == 1 && block_ok) || (t.layout == 0 && (t.sub == 8 || t.sub == 16 || t.sub == 32)));
  if (!shape_ok || t.blocks_per_cu < 0) {
    pf_set_error("batch: invalid launch shape (layout %d, sub %d, block %d, blocks_per_cu %d)%s", t.layout, t.sub,
                 t.block, t.blocks_per_cu, (t.layout == 1 && !block_ok) ? ": no DOF=3 node blocks" : "");
    return PF_INVALID;
  }
  return PF_OK;
}

// layouts and vectors the launch shape needs, built once per matrix
__global__ void k_fill_ones(int64_t n, double *__restrict__ a) {
  for (int64_t i = blockIdx.x * (int64_t)blockDim.x + threadIdx.x; i < n; i += (int64_t)gridDim.x * blockDim.x)
    a[i] = 1.0;
}

int prepare(pf_ctx *ctx, BatchWork *b, const int64_t *row_ptr, const int32_t *col_idx, const double *val,
            int64_t nnz, bool mixed, int layout, int jacobi) {
  const int64_t n = b->n;
  SliceMap map;
  map.nslices = b->nslices;
  map.snode = b->snode.p;
  map.send = b->send.p;
  if (mixed && (!b->have_scale || b->scaled != jacobi)) {
    PF_CHECK(b->scale.ensure(n));
    PF_CHECK(b->ones.ensure(n));
    k_fill_ones<<<pf_blocks(n, 256, 65536), 256, 0, ctx->stream>>>(n, b->ones.p);
    PF_HIP(hipGetLastError());
    if (jacobi == 1) {
      PF_CHECK(pf_jacobi_scale(ctx, n, row_ptr, col_idx, val, b->scale.p));
    } else if (jacobi == 2) {
      PF_CHECK(pf_jacobi_row_scale(ctx, n, row_ptr, col_idx, val, b->scale.p));
    } else {
      k_fill_ones<<<pf_blocks(n, 256, 65536), 256, 0, ctx->stream>>>(n, b->scale.p);
      PF_HIP(hipGetLastError());
    }
    b->have_scale = true;
    b->scaled = jacobi;
    b->have_s32 = b->have_v32 = false;  // FP32 copies follow the scaling
  }
  if (layout == 1 && !b->have_s64) {
    PF_CHECK(pf_sell3_build<double>(ctx, n, row_ptr, col_idx, val, b->s64, nullptr, &map));
    const int64_t nc = b->s64.nslots * 64;
    if (nc > 0)
      k_bsell_pad<<<pf_blocks(nc, 256, 65536), 256, 0, ctx->stream>>>(b->nslices, b->s64.soff.p, b->snode.p,
                                                                    b->s64.col.p, nc);
    PF_HIP(hipGetLastError());
    b->have_s64 = true;
  }
  if (mixed && layout == 1 && !b->have_s32) {
    PF_CHECK(pf_sell3_build<float>(ctx, n, row_ptr, col_idx, val, b->s32, b->scale.p, &map,
                                   b->scaled == 2 ? b->ones.p : b->scale.p));
    const int64_t nc = b->s32.nslots * 64;
    if (nc > 0)
      k_bsell_pad<<<pf_blocks(nc, 256, 65536), 256, 0, ctx->stream>>>(b->nslices, b->s32.soff.p, b->snode.p,
                                                                    b->s32.col.p, nc);
    PF_HIP(hipGetLastError());
    b->have_s32 = true;
  }
  if (mixed && layout == 0 && !b->have_v32) {
    PF_CHECK(b->val32.ensure((size_t)nnz + 1));
    PF_CHECK(pf_crs_to_f32(ctx, n, row_ptr, col_idx, val, b->scale.p, b->val32.p,
                           b->scaled == 2 ? b->ones.p : b->scale.p));
    b->have_v32 = true;
  }
  PF_CHECK(b->r.ensure(n));
  PF_CHECK(b->r0.ensure(n));
  PF_CHECK(b->p.ensure(n));
  PF_CHECK(b->y.ensure(n));
  PF_CHECK(b->v.ensure(n));
  if (mixed) {
    PF_CHECK(b->zhi.ensure(n));
    PF_CHECK(b->r32.ensure(n));
    PF_CHECK(b->r032.ensure(n));
    PF_CHECK(b->p32.ensure(n));
    PF_CHECK(b->y32.ensure(n));
    PF_CHECK(b->v32.ensure(n));
    PF_CHECK(b->zlo.ensure(n));
  }
  return PF_OK;
}

BArgs make_args(BatchWork *b, const int64_t *row_ptr, const int32_t *col_idx, const double *val, const double *rhs,
                double *x, const pf_solve_opts &o, bool mixed) {
  BArgs a;
  std::memset(&a, 0, sizeof(a));
  a.nmesh = b->nmesh;
  a.itrmax = o.itrmax;
  a.stop_mode = o.stop_mode == 1 ? 1 : 0;
  a.inner_max = o.inner_max > 0 ? o.inner_max : 0;
  a.er0 = o.er0;
  a.rel = o.relative_er;
  a.target = o.target_relres;
  a.delta = o.inner_tol > 0.0 && o.inner_tol < 1.0 ? o.inner_tol : 0.1;
  a.mesh = b->mesh.p;
  a.order = b->order.p;
  a.queue = b->queue.p;
  a.A = BMat<double>{row_ptr, col_idx, val, b->s64.soff.p, b->s64.width.p, b->s64.col.p, b->s64.val.p,
                     b->snode.p, b->send.p};
  a.A32 = BMat<float>{row_ptr, col_idx, b->val32.p, b->s32.soff.p, b->s32.width.p, b->s32.col.p, b->s32.val.p,
                      b->snode.p, b->send.p};
  a.b = rhs;
  a.x = x;
  a.r = b->r.p;
  a.r0 = b->r0.p;
  a.p = b->p.p;
  a.y = b->y.p;
  a.v = b->v.p;
  if (mixed) {
    a.r32 = b->r32.p;
    a.r032 = b->r032.p;
    a.p32 = b->p32.p;
    a.y32 = b->y32.p;
    a.v32 = b->v32.p;
    a.zlo = b->zlo.p;
    a.zhi = b->zhi.p;
    a.scale = b->scale.p;
    a.cscale = b->scaled == 2 ? b->ones.p : b->scale.p;
    a.restart = b->scaled == 2 ? 1 : 0;
  }
  a.stat = b->stat.p;
  a.t_raw = b->t_raw.p;
  return a;
}

unsigned batch_grid(pf_ctx *ctx, BatchWork *b, const pf_batch_tune &t, bool mixed) {
  int occ = 1;
  const int rc = mixed ? dispatch<32>(t, Occ{&occ}) : dispatch<64>(t, Occ{&occ});
  if (rc != PF_OK || occ < 1) occ = 1;
  int per_cu = t.blocks_per_cu > 0 ? std::min(t.blocks_per_cu, occ) : occ;
  int64_t g = (int64_t)ctx->num_cus * per_cu;
  if (g > b->nmesh) g = b->nmesh;
  return (unsigned)std::max<int64_t>(g, 1);
}

// queue reset + persistent kernel, captured once per (shape, pointers) and replayed
int run_batch(pf_ctx *ctx, BatchWork *b, const BArgs &a, const pf_batch_tune &t, bool mixed, bool use_graph,
              float *ms) {
  const unsigned grid = batch_grid(ctx, b, t, mixed);
  const Launcher ln{ctx, &a, grid};
  BatchWork::GraphKey key;
  std::memset(&key, 0, sizeof(key));
  std::memcpy(&key.a, &a, sizeof(BArgs));
  key.t = t;
  key.grid = grid;
  key.mixed = mixed ? 1 : 0;
  PF_HIP(hipEventRecord(ctx->ev0, ctx->stream));
  bool launched = false;
  if (use_graph && pf_graphs_allowed()) {
    if (!b->gexec || !b->gkey_valid || std::memcmp(&b->gkey, &key, sizeof(key)) != 0) {
      if (b->gexec) {
        PF_HIP(hipGraphExecDestroy(b->gexec));
        b->gexec = nullptr;
      }
      b->gkey_valid = false;
      PF_CHECK(pf_capture_graph(
          ctx,
          [&] {
            k_queue_reset<<<1, 64, 0, ctx->stream>>>(b->queue.p);
            return mixed ? dispatch<32>(t, ln) : dispatch<64>(t, ln);
          },
          &b->gexec));
      if (b->gexec) {
        std::memcpy(&b->gkey, &key, sizeof(key));
        b->gkey_valid = true;
      }
    }
    if (b->gexec) {
      PF_HIP(hipGraphLaunch(b->gexec, ctx->stream));
      launched = true;
    }
  }
  if (!launched) {
    k_queue_reset<<<1, 64, 0, ctx->stream>>>(b->queue.p);
    PF_CHECK(mixed ? dispatch<32>(t, ln) : dispatch<64>(t, ln));
  }
  PF_HIP(hipEventRecord(ctx->ev1, ctx->stream));
  PF_HIP(hipEventSynchronize(ctx->ev1));
  PF_HIP(hipEventElapsedTime(ms, ctx->ev0, ctx->ev1));
  return PF_OK;
}

}  // namespace

int pf_batch_solve_device(pf_ctx *ctx, BatchWork *b, const int64_t *row_ptr, const int32_t *col_idx,
                          const double *val, int64_t nnz, const double *rhs, double *x, const pf_solve_opts &o,
                          const pf_batch_tune *tune, uint64_t mat_version, pf_batch_mesh_stats *mesh_stats,
                          pf_stats *total) {
  if (!b || b->nmesh < 1) {
    pf_set_error("batch: call pf_system_batch_set first");
    return PF_INVALID;
  }
  const bool mixed = o.precision == 32;
  if (o.precision != 32 && o.precision != 64) {
    pf_set_error("batch: precision must be 64 or 32");
    return PF_INVALID;
  }
  if (o.stop_mode != 0 && o.stop_mode != 1) {
    pf_set_error("batch: stop_mode must be 0 (bicgstab_CRS_2) or 1 (|r| < relative_er |b|)");
    return PF_INVALID;
  }
  if (mat_version != b->mat_version) {
    b->have_s64 = b->have_s32 = b->have_v32 = b->have_scale = false;
    b->mat_version = mat_version;
    b->gkey_valid = false;
  }
  pf_batch_tune t;
  if (tune) {
    t = *tune;
  } else if (b->have_tuned[mixed ? 1 : 0]) {
    t = b->tuned[mixed ? 1 : 0];
  } else {
    t.layout = b->block_ok ? 1 : 0;
    t.sub = 16;
    t.block = 512;
    t.blocks_per_cu = 0;
  }
  PF_CHECK(valid_tune(t, b->block_ok));
  PF_CHECK(prepare(ctx, b, row_ptr, col_idx, val, nnz, mixed, t.layout, o.jacobi == 2 ? 2 : (o.jacobi != 0 ? 1 : 0)));
  const BArgs a = make_args(b, row_ptr, col_idx, val, rhs, x, o, mixed);
  float ms = 0.f;
  PF_CHECK(run_batch(ctx, b, a, t, mixed, o.use_graph != 0, &ms));
  // per-mesh results; wall-clock ticks -> ms from the first mesh's start
  std::vector<pf_batch_mesh_stats> st(b->nmesh);
  std::vector<int64_t> raw(2 * (size_t)b->nmesh);
  PF_HIP(hipMemcpyAsync(st.data(), b->stat.p, sizeof(pf_batch_mesh_stats) * b->nmesh, hipMemcpyDeviceToHost,
                        ctx->stream));
  PF_HIP(hipMemcpyAsync(raw.data(), b->t_raw.p, sizeof(int64_t) * 2 * b->nmesh, hipMemcpyDeviceToHost, ctx->stream));
  PF_HIP(hipStreamSynchronize(ctx->stream));
  int rate_khz = 0;
  PF_HIP(hipDeviceGetAttribute(&rate_khz, hipDeviceAttributeWallClockRate, ctx->device));
  const double tick_ms = rate_khz > 0 ? 1.0 / (double)rate_khz : 0.0;
  int64_t tmin = raw[0];
  for (int32_t m = 0; m < b->nmesh; ++m) tmin = std::min(tmin, raw[2 * m]);
  int status = PF_OK;
  int64_t iters = 0, reps = 0;
  double worst = 0.0;
  for (int32_t m = 0; m < b->nmesh; ++m) {
    st[m].start_ms = (double)(raw[2 * m] - tmin) * tick_ms;
    st[m].end_ms = (double)(raw[2 * m + 1] - tmin) * tick_ms;
    iters += st[m].iterations;
    reps += st[m].replacements;
    worst = std::max(worst, st[m].true_relres);
    if (status == PF_OK && st[m].status != PF_OK) status = st[m].status;
  }
  if (mesh_stats) std::memcpy(mesh_stats, st.data(), sizeof(pf_batch_mesh_stats) * b->nmesh);
  if (total) {
    std::memset(total, 0, sizeof(*total));
    total->status = status;
    total->iterations = (int32_t)std::min<int64_t>(iters, INT32_MAX);
    total->outer_iterations = (int32_t)std::min<int64_t>(reps, INT32_MAX);
    total->spmv_kind_used = t.layout == 1 ? 2 : 1;
    total->true_relres = worst;
    total->solve_ms = ms;
    total->timed_iterations = iters;
  }
  return status;
}

// time every launch shape with the iteration count capped at `iters` per mesh and keep the
// one with the highest row-iteration throughput; returns the table in ms (may be NULL)
int pf_batch_autotune_device(pf_ctx *ctx, BatchWork *b, const int64_t *row_ptr, const int32_t *col_idx,
                             const double *val, int64_t nnz, const double *rhs, double *x, const pf_solve_opts &o,
                             int32_t iters, uint64_t mat_version, pf_batch_tune *best, int32_t max_results,
                             pf_batch_tune *tried, double *tried_ms, int32_t *ntried) {
  if (!b || b->nmesh < 1) {
    pf_set_error("batch: call pf_system_batch_set first");
    return PF_INVALID;
  }
  std::vector<pf_batch_tune> cand;
  for (int blk : {256, 512, 1024}) {
    if (b->block_ok) cand.push_back(pf_batch_tune{1, 64, blk, 0});
    for (int sub : {8, 16, 32}) cand.push_back(pf_batch_tune{0, sub, blk, 0});
  }
  pf_solve_opts ot = o;
  ot.itrmax = iters > 0 ? iters : 20;
  ot.relative_er = 0.0;  // run exactly itrmax iterations (no early stop)
  ot.er0 = 0.0;
  ot.target_relres = 0.0;
  ot.use_graph = 0;
  const bool mixed = o.precision == 32;
  double best_ms = 0.0;
  int nt = 0;
  pf_batch_tune pick = cand[0];
  for (const pf_batch_tune &t : cand) {
    pf_stats tot;
    // one warm launch, then the timed one
    int rc = pf_batch_solve_device(ctx, b, row_ptr, col_idx, val, nnz, rhs, x, ot, &t, mat_version, nullptr, &tot);
    if (rc < 0) return rc;
    rc = pf_batch_solve_device(ctx, b, row_ptr, col_idx, val, nnz, rhs, x, ot, &t, mat_version, nullptr, &tot);
    if (rc < 0) return rc;
    const double ms = tot.solve_ms;
    if (tried && tried_ms && nt < max_results) {
      tried[nt] = t;
      tried_ms[nt] = ms;
    }
    ++nt;
    if (best_ms == 0.0 || ms < best_ms) {
      best_ms = ms;
      pick = t;
    }
  }
  if (ntried) *ntried = nt;
  b->tuned[mixed ? 1 : 0] = pick;
  b->have_tuned[mixed ? 1 : 0] = true;
  if (best) *best = pick;
  return PF_OK;
}
