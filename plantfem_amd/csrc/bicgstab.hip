// bicgstab.hip -- bicgstab_CRS_2 (FEMSolverClass.f90:2199-2298) on the device.
//
// One iteration of the reference is
//   c1 = r0.r ; y = A p ; c2 = r0.y ; alp = c1/c2 ; e = r - alp y ; v = A e ;
//   c3 = (e.v)/(v.v) ; x += alp p + c3 e ; r = e - c3 v ; rr = r.r ; stop tests ;
//   c1 = r0.r ; bet = c1/(c2 c3) ; p = r + bet (p - c3 y)
// and is mapped onto four kernels that keep every scalar on the device:
//   K1  y = A p            + partial r0.y          -> last block: c2, alpha
//   K2  v = A e, e formed on the fly as r - alpha y while gathering
//                          + partials e.v, v.v     -> last block: omega (or v.v == 0 stop)
//   K3  x, r update        + partials r.r, r0.r    -> last block: stop tests, beta
//   K4  p = r + beta (p - omega y)
// e is never stored.  Each reduction writes one partial per block; the last block to
// arrive (atomic ticket) sums the partials in a fixed order, so results are bitwise
// reproducible run to run.  After a stop every kernel returns at entry, so the host can
// enqueue iterations in chunks (optionally one hipGraph per chunk) and poll a pinned copy
// of the scalars once per chunk.  The redundant loop-top c1 = r0.r of the reference equals
// the c1 of the previous iteration's end and is not recomputed.
//
// The same kernels, instantiated for float vectors and matrix values with double partial
// sums, are the inner solver of the mixed-precision form: FP32 SpMV with FP32 accumulation,
// FP64 global dot products, FP64 residual and solution update in the refinement loop.
#include <algorithm>
#include <cmath>
#include <cstdio>
#include <cstring>
#include <vector>

#include "spmv_dev.hpp"

namespace {

constexpr int BT = 256;  // threads per block of every fused kernel (4 wavefronts)

struct Scal {
  double c1, c2, alpha, omega, beta;
  double rr, rr1, bnorm2;
  double er0, rel;
  double rnorm2;  // refinement residual norm^2
  int32_t itr, itrmax, done, status;
  int32_t stop_mode, hist_cap;
  uint32_t arrive[4];
};

template <typename V>
struct Mat {
  // CRS
  int64_t n;
  const int64_t *row_ptr;
  const int32_t *col_idx;
  const V *val;
  // Sell3
  int64_t nn, nslices;
  const int64_t *soff;
  const int32_t *width;
  const int32_t *col;
  const V *sval;
};

// Visit every row of y = A*xf with the calling thread's share of the grid; op(row, sum) is
// called by exactly one thread per row.  L = 1: Sell3 (3 rows per lane), L = 0: CRS rows
// reduced by SUB-lane subgroups.
template <typename V, int L, int SUB, typename XF, typename OP>
__device__ inline void for_rows(const Mat<V> &m, XF xf, OP op) {
  if (L == 1) {
    const int64_t wave = (blockIdx.x * (int64_t)BT + threadIdx.x) >> 6;
    const int lane = threadIdx.x & 63;
    const int64_t nwaves = ((int64_t)gridDim.x * BT) >> 6;
    for (int64_t s = wave; s < m.nslices; s += nwaves) {
      V y0, y1, y2;
      pfdev::sell3_node<V, V>(m.soff, m.width, m.col, m.sval, s, lane, xf, y0, y1, y2);
      const int64_t i = s * 64 + lane;
      if (i < m.nn) {
        op(3 * i, y0);
        op(3 * i + 1, y1);
        op(3 * i + 2, y2);
      }
    }
  } else {
    const int64_t g = (blockIdx.x * (int64_t)BT + threadIdx.x) / SUB;
    const int sub = threadIdx.x % SUB;
    const int64_t ng = (int64_t)gridDim.x * BT / SUB;
    for (int64_t row = g; row < m.n; row += ng) {
      V s = pfdev::crs_row<V, V, SUB>(m.row_ptr, m.col_idx, m.val, row, sub, xf);
      if (sub == 0) op(row, s);
    }
  }
}

template <int NV>
__device__ inline void block_partials(double (&v)[NV], double *part) {
  __shared__ double sh[NV][BT / 64];
#pragma unroll
  for (int k = 0; k < NV; ++k) {
#pragma unroll
    for (int off = 32; off > 0; off >>= 1) v[k] += __shfl_xor(v[k], off, 64);
  }
  const int w = threadIdx.x >> 6;
  if ((threadIdx.x & 63) == 0)
#pragma unroll
    for (int k = 0; k < NV; ++k) sh[k][w] = v[k];
  __syncthreads();
  if (threadIdx.x == 0)
#pragma unroll
    for (int k = 0; k < NV; ++k) part[(int64_t)k * gridDim.x + blockIdx.x] = ((sh[k][0] + sh[k][1]) + sh[k][2]) + sh[k][3];
}

// true in every thread of the last block to finish; its partial writes are visible
__device__ inline bool last_block(uint32_t *ticket) {
  __shared__ bool am_last;
  __threadfence();
  __syncthreads();
  if (threadIdx.x == 0) {
    uint32_t prev = atomicAdd(ticket, 1u);
    am_last = (prev == gridDim.x - 1);
  }
  __syncthreads();
  if (am_last) __threadfence();
  return am_last;
}

template <int NV>
__device__ inline void final_sums(const double *part, double (&out)[NV]) {
  __shared__ double sh[NV][BT / 64];
  double v[NV];
#pragma unroll
  for (int k = 0; k < NV; ++k) v[k] = 0.0;
  for (int64_t b = threadIdx.x; b < gridDim.x; b += BT)
#pragma unroll
    for (int k = 0; k < NV; ++k) v[k] += __builtin_nontemporal_load(&part[(int64_t)k * gridDim.x + b]);
#pragma unroll
  for (int k = 0; k < NV; ++k) {
#pragma unroll
    for (int off = 32; off > 0; off >>= 1) v[k] += __shfl_xor(v[k], off, 64);
  }
  const int w = threadIdx.x >> 6;
  if ((threadIdx.x & 63) == 0)
#pragma unroll
    for (int k = 0; k < NV; ++k) sh[k][w] = v[k];
  __syncthreads();
#pragma unroll
  for (int k = 0; k < NV; ++k) out[k] = ((sh[k][0] + sh[k][1]) + sh[k][2]) + sh[k][3];
}

// ---- K0: r = b - A x (or r = b when x0 = 0); r0 = p = r; c1 = r.r; |b|^2 -----------
template <typename V, int L, int SUB, bool ZEROX>
__global__ void __launch_bounds__(BT) k_init(Mat<V> m, const V *__restrict__ b, V *__restrict__ x,
                                             V *__restrict__ r, V *__restrict__ r0, V *__restrict__ p,
                                             Scal *__restrict__ sc, double *__restrict__ part) {
  double acc[2] = {0.0, 0.0};
  if (ZEROX) {
    for (int64_t i = blockIdx.x * (int64_t)BT + threadIdx.x; i < m.n; i += (int64_t)gridDim.x * BT) {
      const V bv = b[i];
      x[i] = (V)0;
      r[i] = bv;
      r0[i] = bv;
      p[i] = bv;
      acc[0] += (double)bv * (double)bv;
    }
    acc[1] = acc[0];
  } else {
    auto op = [&](int64_t row, V ax) {
      const V bv = b[row];
      const V rv = bv - ax;
      r[row] = rv;
      r0[row] = rv;
      p[row] = rv;
      acc[0] += (double)rv * (double)rv;
      acc[1] += (double)bv * (double)bv;
    };
    for_rows<V, L, SUB>(m, pfdev::PlainX<V>{x}, op);
  }
  block_partials<2>(acc, part);
  if (last_block(&sc->arrive[0])) {
    double t[2];
    final_sums<2>(part, t);
    if (threadIdx.x == 0) {
      sc->c1 = t[0];
      sc->rr = t[0];
      sc->bnorm2 = t[1];
      sc->itr = 0;
      sc->done = sc->itrmax <= 0 ? 1 : 0;
      sc->status = PF_NOT_CONVERGED;
      sc->arrive[0] = 0;
    }
  }
}

// ---- K1: y = A p ; c2 = r0.y ; alpha = c1/c2 ------------------------------------------
template <typename V, int L, int SUB>
__global__ void __launch_bounds__(BT) k_spmv_p(Mat<V> m, const V *__restrict__ p, V *__restrict__ y,
                                               const V *__restrict__ r0, Scal *__restrict__ sc,
                                               double *__restrict__ part) {
  if (sc->done) return;
  double acc[1] = {0.0};
  auto op = [&](int64_t row, V yv) {
    y[row] = yv;
    acc[0] += (double)r0[row] * (double)yv;
  };
  for_rows<V, L, SUB>(m, pfdev::PlainX<V>{p}, op);
  block_partials<1>(acc, part);
  if (last_block(&sc->arrive[1])) {
    double t[1];
    final_sums<1>(part, t);
    if (threadIdx.x == 0) {
      sc->c2 = t[0];
      sc->alpha = sc->c1 / t[0];
      sc->arrive[1] = 0;
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

// ---- K2: v = A e with e = r - alpha y ; e.v, v.v ; omega -------------------------------
template <typename V, int L, int SUB>
__global__ void __launch_bounds__(BT) k_spmv_e(Mat<V> m, const V *__restrict__ r, const V *__restrict__ y,
                                               V *__restrict__ v, Scal *__restrict__ sc,
                                               double *__restrict__ part) {
  if (sc->done) return;
  const V alpha = (V)sc->alpha;
  const EFromRY<V> ef{r, y, alpha};
  double acc[2] = {0.0, 0.0};
  auto op = [&](int64_t row, V vv) {
    v[row] = vv;
    const V e = ef(row);
    acc[0] += (double)e * (double)vv;
    acc[1] += (double)vv * (double)vv;
  };
  for_rows<V, L, SUB>(m, ef, op);
  block_partials<2>(acc, part);
  if (last_block(&sc->arrive[2])) {
    double t[2];
    final_sums<2>(part, t);
    if (threadIdx.x == 0) {
      if (t[1] == 0.0) {
        sc->status = PF_BREAKDOWN_VV;
        sc->done = 1;
      } else {
        sc->omega = t[0] / t[1];
      }
      sc->arrive[2] = 0;
    }
  }
}

// ---- K3: x += alpha p + omega e ; r = e - omega v ; stop tests ; beta ----------------
template <typename V>
__global__ void __launch_bounds__(BT) k_update(int64_t n, V *__restrict__ x, V *__restrict__ r,
                                               const V *__restrict__ r0, const V *__restrict__ p,
                                               const V *__restrict__ y, const V *__restrict__ v,
                                               Scal *__restrict__ sc, double *__restrict__ part,
                                               double *__restrict__ hist) {
  if (sc->done) return;
  const V alpha = (V)sc->alpha, omega = (V)sc->omega;
  double acc[2] = {0.0, 0.0};
  for (int64_t i = blockIdx.x * (int64_t)BT + threadIdx.x; i < n; i += (int64_t)gridDim.x * BT) {
    const V e = r[i] - alpha * y[i];
    x[i] = x[i] + alpha * p[i] + omega * e;
    const V rn = e - omega * v[i];
    r[i] = rn;
    acc[0] += (double)rn * (double)rn;
    acc[1] += (double)r0[i] * (double)rn;
  }
  block_partials<2>(acc, part);
  if (last_block(&sc->arrive[3])) {
    double t[2];
    final_sums<2>(part, t);
    if (threadIdx.x == 0) {
      const int itr = sc->itr + 1;
      const double rr = t[0];
      sc->itr = itr;
      sc->rr = rr;
      if (itr == 1) sc->rr1 = rr;
      if (hist && itr <= sc->hist_cap) hist[itr - 1] = sqrt(rr);
      bool stop;
      if (sc->stop_mode == 0)
        stop = (sqrt(rr / sc->rr1) < sc->rel) || (sqrt(rr) < sc->er0);
      else
        stop = sqrt(rr) < sc->rel * sqrt(sc->bnorm2);
      if (stop) {
        sc->status = PF_OK;
        sc->done = 1;
      } else {
        sc->c1 = t[1];
        const double c2w = sc->c2 * sc->omega;
        sc->beta = t[1] / c2w;
        if (c2w == 0.0) {
          sc->status = PF_BREAKDOWN_C2C3;
          sc->done = 1;
        } else if (itr >= sc->itrmax) {
          sc->status = PF_NOT_CONVERGED;
          sc->done = 1;
        }
      }
      sc->arrive[3] = 0;
    }
  }
}

// ---- K4: p = r + beta (p - omega y) ----------------------------------------------------
template <typename V>
__global__ void __launch_bounds__(BT) k_pupdate(int64_t n, V *__restrict__ p, const V *__restrict__ r,
                                                const V *__restrict__ y, const Scal *__restrict__ sc) {
  if (sc->done) return;
  const V beta = (V)sc->beta, omega = (V)sc->omega;
  for (int64_t i = blockIdx.x * (int64_t)BT + threadIdx.x; i < n; i += (int64_t)gridDim.x * BT)
    p[i] = r[i] + beta * (p[i] - omega * y[i]);
}

// ---- refinement (mixed precision) -----------------------------------------------------
// r64 = b - A64 x64 ; rnorm2 = |r64|^2 (bnorm2 too on the first call)
template <int L, int SUB>
__global__ void __launch_bounds__(BT) k_residual64(Mat<double> m, const double *__restrict__ b,
                                                   const double *__restrict__ x, double *__restrict__ r,
                                                   Scal *__restrict__ sc, double *__restrict__ part) {
  double acc[2] = {0.0, 0.0};
  auto op = [&](int64_t row, double ax) {
    const double rv = b[row] - ax;
    r[row] = rv;
    acc[0] += rv * rv;
    acc[1] += b[row] * b[row];
  };
  for_rows<double, L, SUB>(m, pfdev::PlainX<double>{x}, op);
  block_partials<2>(acc, part);
  if (last_block(&sc->arrive[0])) {
    double t[2];
    final_sums<2>(part, t);
    if (threadIdx.x == 0) {
      sc->rnorm2 = t[0];
      sc->bnorm2 = t[1];
      sc->arrive[0] = 0;
    }
  }
}

// inner right-hand side: (S r) / |r| in FP32, S = diag(scale) or identity
__global__ void k_scale_to_f32(int64_t n, const double *__restrict__ r, const Scal *__restrict__ sc,
                               const double *__restrict__ scale, float *__restrict__ r32) {
  const double s = 1.0 / sqrt(sc->rnorm2);
  for (int64_t i = blockIdx.x * (int64_t)BT + threadIdx.x; i < n; i += (int64_t)gridDim.x * BT)
    r32[i] = (float)(r[i] * s * (scale ? scale[i] : 1.0));
}

// x += |r| S d  (FP64 solution update)
__global__ void k_correct64(int64_t n, double *__restrict__ x, const float *__restrict__ d,
                            const Scal *__restrict__ sc, const double *__restrict__ scale) {
  const double s = sqrt(sc->rnorm2);
  for (int64_t i = blockIdx.x * (int64_t)BT + threadIdx.x; i < n; i += (int64_t)gridDim.x * BT)
    x[i] = x[i] + s * (double)d[i] * (scale ? scale[i] : 1.0);
}

__global__ void k_f64_to_f32(int64_t n, const double *__restrict__ a, float *__restrict__ b) {
  for (int64_t i = blockIdx.x * (int64_t)blockDim.x + threadIdx.x; i < n; i += (int64_t)gridDim.x * blockDim.x)
    b[i] = (float)a[i];
}

}  // namespace

// ========================================================================================
struct SolverWork {
  int64_t n = 0;
  DevBuf<double> r, r0, p, y, v, rref;  // FP64 recurrence vectors (+ refinement residual)
  DevBuf<float> x32, b32, r32, r032, p32, y32, v32, val32;
  DevBuf<double> part, hist;
  DevBuf<Scal> sc;  // [0] FP64 / outer, [1] FP32 inner
  Scal *h_sc = nullptr;
  Sell3<double> s64;
  Sell3<float> s32;
  DevBuf<int32_t> flag;
  DevBuf<double> jscale;  // 1/sqrt|a_ii|
  bool have_s64 = false, have_s32 = false, have_v32 = false, have_scale = false;
  bool s32_scaled = false, v32_scaled = false;
  int block_ok = -1;  // -1 unknown
  const void *key_val = nullptr;
  hipGraphExec_t gexec[2] = {nullptr, nullptr};
  uint64_t gkey[2][16] = {};  // everything a captured chunk bakes in
  std::vector<hipEvent_t> evs;  // kernel_timing: 5 per iteration of a chunk
  double t_sum[4] = {0, 0, 0, 0};
  int64_t t_count[4] = {0, 0, 0, 0};
  ~SolverWork() {
    for (int k = 0; k < 2; ++k)
      if (gexec[k]) (void)hipGraphExecDestroy(gexec[k]);
    for (hipEvent_t e : evs) (void)hipEventDestroy(e);
    if (h_sc) (void)hipHostFree(h_sc);
  }
};

int pf_solver_work_create(SolverWork **w) {
  *w = new SolverWork();
  hipError_t e = hipHostMalloc((void **)&(*w)->h_sc, sizeof(Scal) * 2, hipHostMallocDefault);
  if (e != hipSuccess) {
    delete *w;
    *w = nullptr;
    pf_set_error("hipHostMalloc failed: %s", hipGetErrorString(e));
    return PF_ERR_NOMEM;
  }
  std::memset((*w)->h_sc, 0, sizeof(Scal) * 2);
  return PF_OK;
}

void pf_solver_work_destroy(SolverWork *w) { delete w; }

namespace {

struct Launch {
  unsigned grid;
  int sub;
  int layout;  // 0 CRS, 1 Sell3
};

unsigned fused_grid(pf_ctx *ctx, int64_t work_threads) {
  // enough resident waves to cover HBM latency on every CU, capped so the last-block
  // reduction reads at most a few thousand partials
  int64_t cap = (int64_t)ctx->num_cus * 8;
  int64_t b = (work_threads + BT - 1) / BT;
  if (b > cap) b = cap;
  if (b < 1) b = 1;
  return (unsigned)b;
}

int choose_sub(double avg_row) {
  int sub = 4;
  while (sub < 64 && sub * 2 <= avg_row) sub *= 2;
  return sub;
}

#define PF_DISPATCH_SUB(SUBV, ...)                       \
  switch (SUBV) {                                        \
    case 4: { constexpr int S_ = 4; __VA_ARGS__; } break;   \
    case 8: { constexpr int S_ = 8; __VA_ARGS__; } break;   \
    case 16: { constexpr int S_ = 16; __VA_ARGS__; } break; \
    case 32: { constexpr int S_ = 32; __VA_ARGS__; } break; \
    default: { constexpr int S_ = 64; __VA_ARGS__; } break; \
  }

template <typename V>
int enqueue_init(pf_ctx *ctx, const Launch &L, const Mat<V> &m, const V *b, V *x, V *r, V *r0, V *p, Scal *sc,
                 double *part, bool zerox) {
  hipStream_t st = ctx->stream;
  if (zerox) {
    k_init<V, 0, 4, true><<<L.grid, BT, 0, st>>>(m, b, x, r, r0, p, sc, part);
  } else if (L.layout == 1) {
    k_init<V, 1, 64, false><<<L.grid, BT, 0, st>>>(m, b, x, r, r0, p, sc, part);
  } else {
    PF_DISPATCH_SUB(L.sub, (k_init<V, 0, S_, false><<<L.grid, BT, 0, st>>>(m, b, x, r, r0, p, sc, part)));
  }
  PF_HIP(hipGetLastError());
  return PF_OK;
}

template <typename V>
int enqueue_iters(pf_ctx *ctx, const Launch &L, const Mat<V> &m, V *x, V *r, const V *r0, V *p, V *y, V *v,
                  Scal *sc, double *part, double *hist, int count, hipEvent_t *ev = nullptr) {
  hipStream_t st = ctx->stream;
  const unsigned ge = fused_grid(ctx, m.n);
  for (int it = 0; it < count; ++it) {
    if (ev) PF_HIP(hipEventRecord(ev[5 * it + 0], st));
    if (L.layout == 1) {
      k_spmv_p<V, 1, 64><<<L.grid, BT, 0, st>>>(m, p, y, r0, sc, part);
      if (ev) PF_HIP(hipEventRecord(ev[5 * it + 1], st));
      k_spmv_e<V, 1, 64><<<L.grid, BT, 0, st>>>(m, r, y, v, sc, part);
    } else {
      PF_DISPATCH_SUB(L.sub, (k_spmv_p<V, 0, S_><<<L.grid, BT, 0, st>>>(m, p, y, r0, sc, part)));
      if (ev) PF_HIP(hipEventRecord(ev[5 * it + 1], st));
      PF_DISPATCH_SUB(L.sub, (k_spmv_e<V, 0, S_><<<L.grid, BT, 0, st>>>(m, r, y, v, sc, part)));
    }
    if (ev) PF_HIP(hipEventRecord(ev[5 * it + 2], st));
    k_update<V><<<ge, BT, 0, st>>>(m.n, x, r, r0, p, y, v, sc, part, hist);
    if (ev) PF_HIP(hipEventRecord(ev[5 * it + 3], st));
    k_pupdate<V><<<ge, BT, 0, st>>>(m.n, p, r, y, sc);
    if (ev) PF_HIP(hipEventRecord(ev[5 * it + 4], st));
  }
  PF_HIP(hipGetLastError());
  return PF_OK;
}

// Run the iteration loop of one BiCGSTAB (scalars slot `slot`) to its stop.
template <typename V>
int run_loop(pf_ctx *ctx, SolverWork *w, int slot, const Launch &L, const Mat<V> &m, V *x, V *r, const V *r0,
             V *p, V *y, V *v, double *hist, const pf_solve_opts &o) {
  Scal *sc = w->sc.p + slot;
  Scal *hs = w->h_sc + slot;
  const int chunk = o.check_every > 0 ? o.check_every : 32;
  const bool timing = o.kernel_timing != 0;
  bool use_graph = o.use_graph != 0 && !timing;
  if (timing && (int)w->evs.size() < 5 * chunk) {
    while ((int)w->evs.size() < 5 * chunk) {
      hipEvent_t e;
      PF_HIP(hipEventCreate(&e));
      w->evs.push_back(e);
    }
  }
  uint64_t key[16] = {(uint64_t)L.grid, (uint64_t)L.sub, (uint64_t)L.layout, (uint64_t)chunk,
                      (uint64_t)m.n, (uint64_t)(uintptr_t)m.val, (uint64_t)(uintptr_t)m.sval,
                      (uint64_t)(uintptr_t)m.col, (uint64_t)(uintptr_t)x, (uint64_t)(uintptr_t)r,
                      (uint64_t)(uintptr_t)r0, (uint64_t)(uintptr_t)p, (uint64_t)(uintptr_t)y,
                      (uint64_t)(uintptr_t)v, (uint64_t)(uintptr_t)w->part.p, (uint64_t)(uintptr_t)hist};
  if (use_graph && (w->gexec[slot] == nullptr || std::memcmp(w->gkey[slot], key, sizeof(key)) != 0)) {
    if (w->gexec[slot]) {
      PF_HIP(hipGraphExecDestroy(w->gexec[slot]));
      w->gexec[slot] = nullptr;
    }
    hipGraph_t graph;
    PF_HIP(hipStreamBeginCapture(ctx->stream, hipStreamCaptureModeThreadLocal));
    int rc = enqueue_iters<V>(ctx, L, m, x, r, r0, p, y, v, sc, w->part.p, hist, chunk);
    hipError_t ce = hipStreamEndCapture(ctx->stream, &graph);
    if (rc != PF_OK) return rc;
    if (ce != hipSuccess) {
      pf_set_error("hipStreamEndCapture: %s", hipGetErrorString(ce));
      return PF_ERR_HIP;
    }
    hipError_t ie = hipGraphInstantiate(&w->gexec[slot], graph, nullptr, nullptr, 0);
    (void)hipGraphDestroy(graph);
    if (ie != hipSuccess) {
      w->gexec[slot] = nullptr;
      use_graph = false;  // fall back to direct launches of the same kernels
    } else {
      std::memcpy(w->gkey[slot], key, sizeof(key));
    }
  }
  for (;;) {
    const int itr_before = hs->itr;
    if (use_graph) {
      PF_HIP(hipGraphLaunch(w->gexec[slot], ctx->stream));
    } else {
      PF_CHECK(enqueue_iters<V>(ctx, L, m, x, r, r0, p, y, v, sc, w->part.p, hist, chunk,
                                timing ? w->evs.data() : nullptr));
    }
    PF_HIP(hipMemcpyAsync(hs, sc, sizeof(Scal), hipMemcpyDeviceToHost, ctx->stream));
    PF_HIP(hipStreamSynchronize(ctx->stream));
    if (timing) {
      // only iterations that ran (kernels after the stop return at entry)
      const int live = hs->itr - itr_before;
      for (int it = 0; it < live && it < chunk; ++it)
        for (int k = 0; k < 4; ++k) {
          if (k == 3 && hs->done && it == live - 1) continue;  // p update skipped after a stop
          float ms = 0.f;
          PF_HIP(hipEventElapsedTime(&ms, w->evs[5 * it + k], w->evs[5 * it + k + 1]));
          w->t_sum[k] += ms;
          w->t_count[k] += 1;
        }
    }
    if (o.debug) std::printf("BiCGSTAB >> itr %d  sqrt(rr) %.6e\n", hs->itr, std::sqrt(hs->rr));
    if (hs->done) break;
  }
  return PF_OK;
}

int prepare_scalars(pf_ctx *ctx, SolverWork *w, int slot, const pf_solve_opts &o, int itrmax, double rel,
                    int stop_mode, int hist_cap) {
  Scal h;
  std::memset(&h, 0, sizeof(h));
  h.er0 = o.er0;
  h.rel = rel;
  h.itrmax = itrmax;
  h.stop_mode = stop_mode;
  h.hist_cap = hist_cap;
  w->h_sc[slot] = h;
  PF_HIP(hipMemcpyAsync(w->sc.p + slot, w->h_sc + slot, sizeof(Scal), hipMemcpyHostToDevice, ctx->stream));
  return PF_OK;
}

}  // namespace

int pf_bicgstab_device(pf_ctx *ctx, SolverWork *w, int64_t n, const int64_t *row_ptr, const int32_t *col_idx,
                       const double *val, int64_t nnz, const double *b, double *x, const pf_solve_opts &o,
                       bool matrix_changed, SolveResult &res) {
  res = SolveResult();
  for (int k = 0; k < 4; ++k) {
    w->t_sum[k] = 0.0;
    w->t_count[k] = 0;
  }
  if (n <= 0) return PF_OK;
  if (n != w->n) {
    w->n = n;
    matrix_changed = true;
  }
  if (val != w->key_val) matrix_changed = true;
  w->key_val = val;
  PF_CHECK(w->r.ensure(n));
  PF_CHECK(w->r0.ensure(n));
  PF_CHECK(w->p.ensure(n));
  PF_CHECK(w->y.ensure(n));
  PF_CHECK(w->v.ensure(n));
  PF_CHECK(w->sc.ensure(2));
  PF_CHECK(w->flag.ensure(1));
  const int grid_cap = ctx->num_cus * 8;
  PF_CHECK(w->part.ensure((size_t)grid_cap * 2 + 16));

  PF_HIP(hipEventRecord(ctx->ev0, ctx->stream));
  if (matrix_changed) {
    w->have_s64 = w->have_s32 = w->have_v32 = w->have_scale = false;
    w->block_ok = -1;
    std::memset(w->gkey, 0, sizeof(w->gkey));
  }
  const double avg_row = (double)nnz / (double)n;
  int layout = 0;
  if (o.spmv_kind != 1) {
    if (w->block_ok < 0) {
      int ok = 0;
      PF_CHECK(pf_block_structure_check(ctx, n, row_ptr, col_idx, n / 3, w->flag.p, &ok));
      w->block_ok = ok;
    }
    layout = w->block_ok ? 1 : 0;
    if (o.spmv_kind == 2 && !layout) {
      pf_set_error("blocked SpMV requested but the CRS has no DOF=3 node-block structure");
      return PF_INVALID;
    }
  }
  const bool mixed = o.precision == 32;
  const bool scaled = mixed && o.jacobi != 0;
  if (scaled && !w->have_scale) {
    PF_CHECK(w->jscale.ensure(n));
    PF_CHECK(pf_jacobi_scale(ctx, n, row_ptr, col_idx, val, w->jscale.p));
    w->have_scale = true;
  }
  const double *scale = scaled ? w->jscale.p : nullptr;
  if (layout == 1) {
    if (!mixed && !w->have_s64) {
      PF_CHECK(pf_sell3_build<double>(ctx, n, row_ptr, col_idx, val, w->s64));
      w->have_s64 = true;
    }
    if (mixed && (!w->have_s32 || w->s32_scaled != scaled)) {
      PF_CHECK(pf_sell3_build<float>(ctx, n, row_ptr, col_idx, val, w->s32, scale));
      w->have_s32 = true;
      w->s32_scaled = scaled;
    }
  } else if (mixed && (!w->have_v32 || w->v32_scaled != scaled)) {
    PF_CHECK(w->val32.ensure((size_t)nnz + 1));
    PF_CHECK(pf_crs_to_f32(ctx, n, row_ptr, col_idx, val, scale, w->val32.p));
    w->have_v32 = true;
    w->v32_scaled = scaled;
  }
  PF_HIP(hipEventRecord(ctx->ev1, ctx->stream));
  PF_HIP(hipEventSynchronize(ctx->ev1));
  float setup_ms = 0.f;
  PF_HIP(hipEventElapsedTime(&setup_ms, ctx->ev0, ctx->ev1));
  res.setup_ms = setup_ms;

  Mat<double> m64{n, row_ptr, col_idx, val, 0, 0, nullptr, nullptr, nullptr, nullptr};
  Launch L64{0, choose_sub(avg_row), layout};
  if (layout == 1 && w->have_s64) {
    m64.nn = w->s64.nn;
    m64.nslices = w->s64.nslices;
    m64.soff = w->s64.soff.p;
    m64.width = w->s64.width.p;
    m64.col = w->s64.col.p;
    m64.sval = w->s64.val.p;
  }
  // FP64 matrix for the mixed refinement residual: CRS rows (the blocked copy is FP32 only)
  Launch Lres{fused_grid(ctx, n * (int64_t)choose_sub(avg_row)), choose_sub(avg_row), 0};
  L64.grid = fused_grid(ctx, layout == 1 ? ((n / 3 + 63) / 64) * 64 : n * (int64_t)L64.sub);
  res.kind = layout == 1 ? 2 : 1;

  PF_HIP(hipEventRecord(ctx->ev0, ctx->stream));
  if (!mixed) {
    double *hist = nullptr;
    PF_CHECK(prepare_scalars(ctx, w, 0, o, o.itrmax, o.relative_er, o.stop_mode == 1 ? 1 : 0, 0));
    PF_CHECK(enqueue_init<double>(ctx, L64, m64, b, x, w->r.p, w->r0.p, w->p.p, w->sc.p, w->part.p, false));
    PF_CHECK(run_loop<double>(ctx, w, 0, L64, m64, x, w->r.p, w->r0.p, w->p.p, w->y.p, w->v.p, hist, o));
    const Scal &h = w->h_sc[0];
    res.status = h.status;
    // the reference's loop counter: a v.v == 0 stop happens inside iteration itr+1
    res.iterations = h.itr + (h.status == PF_BREAKDOWN_VV ? 1 : 0);
    res.rr = h.rr;
    res.rr1 = h.rr1;
  } else {
    PF_CHECK(w->rref.ensure(n));
    PF_CHECK(w->x32.ensure(n));
    PF_CHECK(w->b32.ensure(n));
    PF_CHECK(w->r32.ensure(n));
    PF_CHECK(w->r032.ensure(n));
    PF_CHECK(w->p32.ensure(n));
    PF_CHECK(w->y32.ensure(n));
    PF_CHECK(w->v32.ensure(n));
    Mat<float> m32{n, row_ptr, col_idx, w->have_v32 ? w->val32.p : nullptr, 0, 0, nullptr, nullptr, nullptr, nullptr};
    Launch L32{0, choose_sub(avg_row), layout};
    if (layout == 1) {
      m32.nn = w->s32.nn;
      m32.nslices = w->s32.nslices;
      m32.soff = w->s32.soff.p;
      m32.width = w->s32.width.p;
      m32.col = w->s32.col.p;
      m32.sval = w->s32.val.p;
    }
    L32.grid = fused_grid(ctx, layout == 1 ? ((n / 3 + 63) / 64) * 64 : n * (int64_t)L32.sub);
    const unsigned ge = fused_grid(ctx, n);
    PF_CHECK(prepare_scalars(ctx, w, 0, o, 0, 0.0, 1, 0));
    int total_inner = 0, outer = 0, inner_status = PF_OK;
    double relres = 0.0;
    const int max_outer = o.max_outer > 0 ? o.max_outer : 50;
    for (;;) {
      PF_DISPATCH_SUB(Lres.sub, (k_residual64<0, S_><<<Lres.grid, BT, 0, ctx->stream>>>(
                                    m64, b, x, w->rref.p, w->sc.p, w->part.p)));
      PF_HIP(hipGetLastError());
      PF_HIP(hipMemcpyAsync(w->h_sc, w->sc.p, sizeof(Scal), hipMemcpyDeviceToHost, ctx->stream));
      PF_HIP(hipStreamSynchronize(ctx->stream));
      const double bn = std::sqrt(w->h_sc[0].bnorm2);
      relres = bn > 0 ? std::sqrt(w->h_sc[0].rnorm2) / bn : std::sqrt(w->h_sc[0].rnorm2);
      if (o.debug) std::printf("refinement %d  relres %.6e  inner %d\n", outer, relres, total_inner);
      if (!(relres >= o.target_relres) || w->h_sc[0].rnorm2 == 0.0) break;  // also stops on NaN
      if (outer >= max_outer || total_inner >= o.itrmax) {
        inner_status = PF_NOT_CONVERGED;
        break;
      }
      k_scale_to_f32<<<ge, BT, 0, ctx->stream>>>(n, w->rref.p, w->sc.p, scale, w->b32.p);
      PF_HIP(hipGetLastError());
      // restart cap: a stagnating FP32 recurrence is better restarted from the FP64 residual
      const int inner_cap = std::min(o.itrmax - total_inner, 20000);
      PF_CHECK(prepare_scalars(ctx, w, 1, o, inner_cap, o.inner_tol, 1, 0));
      PF_CHECK(enqueue_init<float>(ctx, L32, m32, w->b32.p, w->x32.p, w->r32.p, w->r032.p, w->p32.p,
                                   w->sc.p + 1, w->part.p, true));
      PF_CHECK(run_loop<float>(ctx, w, 1, L32, m32, w->x32.p, w->r32.p, w->r032.p, w->p32.p, w->y32.p, w->v32.p,
                               nullptr, o));
      total_inner += w->h_sc[1].itr;
      const int ist = w->h_sc[1].status;
      ++outer;
      if (!std::isfinite(w->h_sc[1].rr)) {
        // the FP32 recurrence broke down numerically: keep x, report the stop
        inner_status = PF_BREAKDOWN_C2C3;
        break;
      }
      k_correct64<<<ge, BT, 0, ctx->stream>>>(n, x, w->x32.p, w->sc.p, scale);
      PF_HIP(hipGetLastError());
      if (ist == PF_BREAKDOWN_VV || ist == PF_BREAKDOWN_C2C3) {
        // the correction made so far is kept; a breakdown right after a restart is final
        if (w->h_sc[1].itr <= 1) {
          inner_status = ist;
          break;
        }
      }
    }
    res.status = inner_status;
    res.iterations = total_inner;
    res.outer = outer;
    res.rr = w->h_sc[0].rnorm2;
    res.rr1 = 0.0;
    res.true_relres = relres;
  }
  PF_HIP(hipEventRecord(ctx->ev1, ctx->stream));
  PF_HIP(hipEventSynchronize(ctx->ev1));
  float ms = 0.f;
  PF_HIP(hipEventElapsedTime(&ms, ctx->ev0, ctx->ev1));
  res.solve_ms = ms;
  for (int k = 0; k < 4; ++k) res.t_ms[k] = w->t_count[k] ? w->t_sum[k] / (double)w->t_count[k] : 0.0;
  res.timed = w->t_count[0];

  // true residual of the solved system in FP64 (CRS rows, independent of the solve layout)
  if (!mixed) {
    PF_CHECK(w->rref.ensure(n));
    PF_DISPATCH_SUB(Lres.sub, (k_residual64<0, S_><<<Lres.grid, BT, 0, ctx->stream>>>(
                                  m64, b, x, w->rref.p, w->sc.p, w->part.p)));
    PF_HIP(hipGetLastError());
    PF_HIP(hipMemcpyAsync(w->h_sc, w->sc.p, sizeof(Scal), hipMemcpyDeviceToHost, ctx->stream));
    PF_HIP(hipStreamSynchronize(ctx->stream));
    const double bn = std::sqrt(w->h_sc[0].bnorm2);
    res.true_relres = bn > 0 ? std::sqrt(w->h_sc[0].rnorm2) / bn : std::sqrt(w->h_sc[0].rnorm2);
  }
  return PF_OK;
}

namespace {
__global__ void k_f32_to_f64(int64_t n, const float *__restrict__ a, double *__restrict__ b) {
  for (int64_t i = blockIdx.x * (int64_t)blockDim.x + threadIdx.x; i < n; i += (int64_t)gridDim.x * blockDim.x)
    b[i] = (double)a[i];
}
}  // namespace

int pf_work_spmv(pf_ctx *ctx, SolverWork *w, int64_t n, const int64_t *row_ptr, const int32_t *col_idx,
                 const double *val, int64_t nnz, bool matrix_changed, int kind, const double *x, double *y,
                 int repeat, double *ms) {
  if (ms) *ms = 0.0;
  if (n <= 0) return PF_OK;
  if (kind < 1 || kind > 3) {
    pf_set_error("spmv kind must be 1, 2 or 3");
    return PF_INVALID;
  }
  if (n != w->n) {
    w->n = n;
    matrix_changed = true;
  }
  if (val != w->key_val) matrix_changed = true;
  w->key_val = val;
  PF_CHECK(w->flag.ensure(1));
  if (matrix_changed) {
    w->have_s64 = w->have_s32 = w->have_v32 = w->have_scale = false;
    w->block_ok = -1;
    std::memset(w->gkey, 0, sizeof(w->gkey));
  }
  if (kind != 1) {
    if (w->block_ok < 0) {
      int ok = 0;
      PF_CHECK(pf_block_structure_check(ctx, n, row_ptr, col_idx, n / 3, w->flag.p, &ok));
      w->block_ok = ok;
    }
    if (!w->block_ok) {
      pf_set_error("blocked SpMV requested but the CRS has no DOF=3 node-block structure");
      return PF_INVALID;
    }
    if (kind == 2 && !w->have_s64) {
      PF_CHECK(pf_sell3_build<double>(ctx, n, row_ptr, col_idx, val, w->s64));
      w->have_s64 = true;
    }
    if (kind == 3 && (!w->have_s32 || w->s32_scaled)) {
      PF_CHECK(pf_sell3_build<float>(ctx, n, row_ptr, col_idx, val, w->s32));
      w->have_s32 = true;
      w->s32_scaled = false;
    }
  }
  if (repeat < 1) repeat = 1;
  const unsigned ge = pf_blocks(n, 256, 65536);
  if (kind == 3) {
    PF_CHECK(w->x32.ensure(n));
    PF_CHECK(w->y32.ensure(n));
    k_f64_to_f32<<<ge, 256, 0, ctx->stream>>>(n, x, w->x32.p);
    PF_HIP(hipGetLastError());
  }
  PF_HIP(hipEventRecord(ctx->ev0, ctx->stream));
  for (int rep = 0; rep < repeat; ++rep) {
    if (kind == 1)
      PF_CHECK(pf_spmv_crs_launch<double>(ctx, n, row_ptr, col_idx, val, x, y, (double)nnz / (double)n));
    else if (kind == 2)
      PF_CHECK(pf_spmv_sell3_launch<double>(ctx, w->s64, x, y));
    else
      PF_CHECK(pf_spmv_sell3_launch<float>(ctx, w->s32, w->x32.p, w->y32.p));
  }
  PF_HIP(hipEventRecord(ctx->ev1, ctx->stream));
  PF_HIP(hipEventSynchronize(ctx->ev1));
  float t = 0.f;
  PF_HIP(hipEventElapsedTime(&t, ctx->ev0, ctx->ev1));
  if (ms) *ms = t / repeat;
  if (kind == 3) {
    k_f32_to_f64<<<ge, 256, 0, ctx->stream>>>(n, w->y32.p, y);
    PF_HIP(hipGetLastError());
    PF_HIP(hipStreamSynchronize(ctx->stream));
  }
  return PF_OK;
}
