// bicgstab.hip -- bicgstab_CRS_2 (FEMSolverClass.f90:2199-2298) on the device.
//
// One iteration of the reference is
//   c1 = r0.r ; y = A p ; c2 = r0.y ; alp = c1/c2 ; e = r - alp y ; v = A e ;
//   c3 = (e.v)/(v.v) ; x += alp p + c3 e ; r = e - c3 v ; rr = r.r ; stop tests ;
//   c1 = r0.r ; bet = c1/(c2 c3) ; p = r + bet (p - c3 y)
// and runs as four kernels:
//   K1  y = A p                                      partials r0.y
//   K2  c2, alpha from K1's partials; v = A e with e = r - alpha y formed while gathering
//                                                    partials e.v, v.v
//   K3  omega (or the v.v == 0 stop) from K2's partials; x += alpha p + omega e;
//       r = e - omega v                              partials r.r, r0.r
//   K4  rr, stop tests, beta from K3's partials; p = r + beta (p - omega y)
// Reductions need no atomics and no fences: every block writes one partial sum per dot
// product, and every block of the NEXT kernel reads all partials and sums them in the same
// fixed order, so each block holds bitwise the same scalar (and results are reproducible
// run to run).  Kernel boundaries make the partials visible across the chip's XCDs.  A
// scalar is written by block 0 of one kernel and read only by later kernels; the iteration
// counter alternates between two slots by iteration parity (a kernel argument) so no block
// reads a value its own launch writes.  After a stop every kernel returns at entry, so the
// host enqueues iterations in chunks (one hipGraph per chunk) and polls a pinned copy of
// the scalars once per chunk.  The redundant loop-top c1 = r0.r of the reference equals
// the c1 of the previous iteration's end and is not recomputed.  e is never stored.
//
// The same kernels, instantiated for float vectors and matrix values (FP32 SpMV with FP32
// accumulation) with double partial sums, are the inner solver of the mixed-precision
// form; the refinement loop keeps the residual r = b - Ax, the solution update and all
// global dot products in FP64.
#include <algorithm>
#include <cmath>
#include <cstdio>
#include <cstring>
#include <type_traits>
#include <vector>

#include "spmv_dev.hpp"

namespace {

constexpr int BT = 256;  // threads per block of every fused kernel (4 wavefronts)
constexpr int KEY_WORDS = 45;      // words of the key of a captured iteration chunk
constexpr int ILU_KEY_WORDS = 32;  // the same for the ILU chunk

struct Scal {
  double c1, c2, alpha, omega, beta;
  double rr, rr1, bnorm2, er0, rel;
  double rr0;          // r.r of the initial residual (stop rule of bicgstab_CRS)
  double rnorm2;       // refinement: |b - A x|^2
  double rr_best[2];   // stagnation watch: smallest r.r so far (slot by parity)
  double maxr[2];      // reliable updates: largest |r| since the last replacement (slot by parity)
  double relres;       // reliable updates: FP64 |b - A x| / |b| at the last replacement
  double target, delta;
  int32_t pending;     // reliable updates: a replacement is due after this iteration
  int32_t nupd;        // replacements done
  int32_t itr_rep;     // iteration of the last replacement
  int32_t itr[2];      // completed iterations, slot by parity (written by K4)
  int32_t itr_best[2]; // iteration of rr_best
  int32_t itr_last;    // host-facing copy of the latest count
  int32_t itrmax, done, status, stop_mode, hist_cap, guard, stall;
  int32_t restart;     // reliable updates: also restart the recurrence (r0 = p = r) at a replacement
};

// V: vector and accumulation type of the products; M: stored matrix values (M = V, or FP32
// values with FP64 vectors for precision 48)
template <typename V, typename M = V>
struct Mat {
  int64_t n;
  const int64_t *row_ptr;
  const int32_t *col_idx;
  const M *val;
  int64_t nn, nslices;
  const int64_t *soff;
  const int32_t *width;
  const int32_t *col;
  const M *sval;
  pfdev::Stage st;  // L = 2: the row blocks' staged columns
};

// Visit every row of y = A*xf with the calling thread's share of the grid; op(row, sum) is
// called by exactly one thread per row.  L = 1: Sell3 (3 rows per lane), L = 0: CRS rows
// reduced by SUB-lane subgroups, L = 2: the same CRS sums with x staged in LDS per row block
// (blocks over row blocks, every thread of a block takes part).
template <typename V, int L, int SUB, typename M, typename XF, typename OP>
__device__ inline void for_rows(const Mat<V, M> &m, XF xf, OP op) {
  if constexpr (L == 2) {
    __shared__ V sx[pfdev::ST_UMAX];
    for (int64_t b = blockIdx.x; b < m.st.nb; b += gridDim.x)
      pfdev::staged_block<M, V, SUB, BT>(m.row_ptr, m.col_idx, m.val, m.st, m.n, b, sx, xf, op);
  } else if constexpr (L == 1) {
    const int64_t wave = (blockIdx.x * (int64_t)BT + threadIdx.x) >> 6;
    const int lane = threadIdx.x & 63;
    const int64_t nwaves = ((int64_t)gridDim.x * BT) >> 6;
    for (int64_t s = wave; s < m.nslices; s += nwaves) {
      V y0, y1, y2;
      pfdev::sell3_node<M, V>(m.soff, m.width, m.col, m.sval, s, lane, xf, y0, y1, y2);
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
      V s = pfdev::crs_row<M, V, SUB>(m.row_ptr, m.col_idx, m.val, row, sub, xf);
      if (sub == 0) op(row, s);
    }
  }
}

// block sum of NV values, result in every thread
template <int NV>
__device__ inline void block_sum(double (&v)[NV]) {
  __shared__ double sh[NV][BT / 64];
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
  for (int k = 0; k < NV; ++k) v[k] = ((sh[k][0] + sh[k][1]) + sh[k][2]) + sh[k][3];
  __syncthreads();
}

// one partial per block: part[k*gridDim.x + blockIdx.x]
template <int NV>
__device__ inline void write_partials(double (&v)[NV], double *part) {
  block_sum<NV>(v);
  if (threadIdx.x == 0)
#pragma unroll
    for (int k = 0; k < NV; ++k) part[(int64_t)k * gridDim.x + blockIdx.x] = v[k];
}

// all G partials of the previous kernel, summed in a fixed order (same result in every block)
template <int NV>
__device__ inline void read_partials(const double *part, int G, double (&out)[NV]) {
#pragma unroll
  for (int k = 0; k < NV; ++k) out[k] = 0.0;
  for (int b = threadIdx.x; b < G; b += BT)
#pragma unroll
    for (int k = 0; k < NV; ++k) out[k] += part[(int64_t)k * G + b];
  block_sum<NV>(out);
}

// ---- K0: r = b - A x (or r = b when x0 = 0); r0 = p = r  -> partials r.r, b.b ----------
template <typename V, int L, int SUB, bool ZEROX, typename M>
__global__ void __launch_bounds__(BT) k_init(Mat<V, M> m, const V *__restrict__ b, V *__restrict__ x,
                                             V *__restrict__ r, V *__restrict__ r0, V *__restrict__ p,
                                             double *__restrict__ part) {
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
  write_partials<2>(acc, part);
}

// one block: c1 = r.r, |b|^2, counters
__global__ void __launch_bounds__(BT) k_init_finish(const double *__restrict__ part, int G, Scal *__restrict__ sc) {
  double t[2];
  read_partials<2>(part, G, t);
  if (threadIdx.x == 0) {
    sc->c1 = t[0];
    sc->rr = t[0];
    sc->rr0 = t[0];
    sc->bnorm2 = t[1];
    sc->itr[0] = sc->itr[1] = 0;
    sc->itr_best[0] = sc->itr_best[1] = 0;
    sc->rr_best[0] = sc->rr_best[1] = t[0];
    sc->itr_last = 0;
    sc->done = sc->itrmax <= 0 ? 1 : 0;
    sc->status = PF_NOT_CONVERGED;
    // bicgstab_CRS returns before its loop when r.r < er (LinearSolverClass.f90:1527)
    if (sc->stop_mode == 3 && t[0] < sc->er0) {
      sc->done = 1;
      sc->status = PF_OK;
    }
  }
}

// ---- K1: y = A p -> partials r0.y ------------------------------------------------------
template <typename V, int L, int SUB, typename M>
__global__ void __launch_bounds__(BT) k_spmv_p(Mat<V, M> m, const V *__restrict__ p, V *__restrict__ y,
                                               const V *__restrict__ r0, const Scal *__restrict__ sc,
                                               double *__restrict__ part) {
  if (sc->done) return;
  double acc[1] = {0.0};
  auto op = [&](int64_t row, V yv) {
    y[row] = yv;
    acc[0] += (double)r0[row] * (double)yv;
  };
  for_rows<V, L, SUB>(m, pfdev::PlainX<V>{p}, op);
  write_partials<1>(acc, part);
}

template <typename V>
struct EFromRY {
  const V *__restrict__ r;
  const V *__restrict__ y;
  V alpha;
  __device__ V operator()(int64_t j) const { return r[j] - alpha * y[j]; }
};

// ---- K2: c2, alpha ; v = A e, e = r - alpha y -> partials e.v, v.v ----------------------
template <typename V, int L, int SUB, typename M>
__global__ void __launch_bounds__(BT) k_spmv_e(Mat<V, M> m, const V *__restrict__ r, const V *__restrict__ y,
                                               V *__restrict__ v, Scal *__restrict__ sc,
                                               const double *__restrict__ pin, int Gin,
                                               double *__restrict__ part) {
  if (sc->done) return;
  double t[1];
  read_partials<1>(pin, Gin, t);
  const double c2 = t[0];
  const double alpha_d = sc->c1 / c2;
  const bool first = threadIdx.x == 0 && blockIdx.x == 0;
  if (first) {
    sc->c2 = c2;
    sc->alpha = alpha_d;
  }
  if (sc->guard && !isfinite(alpha_d)) {  // r0.Ap lost: stop before x is touched
    if (first) {
      sc->status = PF_BREAKDOWN_C2C3;
      sc->done = 1;
    }
    return;
  }
  const EFromRY<V> ef{r, y, (V)alpha_d};
  double acc[2] = {0.0, 0.0};
  auto op = [&](int64_t row, V vv) {
    v[row] = vv;
    const V e = ef(row);
    acc[0] += (double)e * (double)vv;
    acc[1] += (double)vv * (double)vv;
  };
  for_rows<V, L, SUB>(m, ef, op);
  write_partials<2>(acc, part);
}

// K2 in two parts for the CRS layout, where every nonzero gathers its own x entry and the
// fused form would gather r and y for each: c2, alpha and e = r - alpha y as a stream, then
// v = A e with plain gathers.  e and v come out bit-identical to the fused kernel's.
template <typename V>
__global__ void __launch_bounds__(BT) k_evec(int64_t n, const V *__restrict__ r, const V *__restrict__ y,
                                             V *__restrict__ e, Scal *__restrict__ sc,
                                             const double *__restrict__ pin, int Gin) {
  if (sc->done) return;
  double t[1];
  read_partials<1>(pin, Gin, t);
  const double c2 = t[0];
  const double alpha_d = sc->c1 / c2;
  const bool first = threadIdx.x == 0 && blockIdx.x == 0;
  if (first) {
    sc->c2 = c2;
    sc->alpha = alpha_d;
  }
  if (sc->guard && !isfinite(alpha_d)) {  // r0.Ap lost: stop before x is touched
    if (first) {
      sc->status = PF_BREAKDOWN_C2C3;
      sc->done = 1;
    }
    return;
  }
  const EFromRY<V> ef{r, y, (V)alpha_d};
  for (int64_t i = blockIdx.x * (int64_t)BT + threadIdx.x; i < n; i += (int64_t)gridDim.x * BT) e[i] = ef(i);
}

// CRS forms: at most 64 VGPRs (8 waves per SIMD), so the grid of 8 blocks per CU is resident at
// once (the staged form at 32 lanes per row took 66 and ran 0.241 instead of 0.177 ms on a
// 465k-row overset system); the blocked form keeps its registers
template <typename V, int L, int SUB, typename M>
__global__ void __launch_bounds__(BT, L == 1 ? 1 : 8) k_spmv_ev(Mat<V, M> m, const V *__restrict__ e, V *__restrict__ v,
                                                Scal *__restrict__ sc, double *__restrict__ part) {
  if (sc->done) return;
  const pfdev::PlainX<V> ef{e};
  double acc[2] = {0.0, 0.0};
  auto op = [&](int64_t row, V vv) {
    v[row] = vv;
    const V ee = e[row];
    acc[0] += (double)ee * (double)vv;
    acc[1] += (double)vv * (double)vv;
  };
  for_rows<V, L, SUB>(m, ef, op);
  write_partials<2>(acc, part);
}

// ---- K3: omega ; x += alpha p + omega e ; r = e - omega v -> partials r.r, r0.r --------
template <typename V>
__global__ void __launch_bounds__(BT) k_update(int64_t n, V *__restrict__ x, V *__restrict__ r,
                                               const V *__restrict__ r0, const V *__restrict__ p,
                                               const V *__restrict__ y, const V *__restrict__ v,
                                               Scal *__restrict__ sc, const double *__restrict__ pin, int Gin,
                                               double *__restrict__ part) {
  if (sc->done) return;
  double t[2];
  read_partials<2>(pin, Gin, t);
  const bool first = threadIdx.x == 0 && blockIdx.x == 0;
  if (t[1] == 0.0) {  // "Bicgstab devide by zero" (FEMSolverClass.f90:2266)
    if (first) {
      sc->status = PF_BREAKDOWN_VV;
      sc->done = 1;
    }
    return;
  }
  const double omega_d = t[0] / t[1];
  if (sc->guard && !isfinite(omega_d)) {
    if (first) {
      sc->status = PF_BREAKDOWN_VV;
      sc->done = 1;
    }
    return;
  }
  if (first) sc->omega = omega_d;
  const V alpha = (V)sc->alpha, omega = (V)omega_d;
  double acc[2] = {0.0, 0.0};
  for (int64_t i = blockIdx.x * (int64_t)BT + threadIdx.x; i < n; i += (int64_t)gridDim.x * BT) {
    const V e = r[i] - alpha * y[i];
    x[i] = x[i] + alpha * p[i] + omega * e;
    const V rn = e - omega * v[i];
    r[i] = rn;
    acc[0] += (double)rn * (double)rn;
    acc[1] += (double)r0[i] * (double)rn;
  }
  write_partials<2>(acc, part);
}

// ---- K4: rr, stop tests, beta ; p = r + beta (p - omega y) -----------------------------
template <typename V>
__global__ void __launch_bounds__(BT) k_pupdate(int64_t n, V *__restrict__ p, const V *__restrict__ r,
                                                const V *__restrict__ y, Scal *__restrict__ sc,
                                                const double *__restrict__ pin, int Gin, int par,
                                                double *__restrict__ hist, double *__restrict__ xhi,
                                                V *__restrict__ xlo) {
  if (sc->done) return;
  double t[2];
  read_partials<2>(pin, Gin, t);
  const bool first = threadIdx.x == 0 && blockIdx.x == 0;
  const double rr = t[0];
  const int itr = sc->itr[par] + 1;
  const double rr1 = (itr == 1) ? rr : sc->rr1;
  if (first) {
    sc->itr[par ^ 1] = itr;
    sc->itr_last = itr;
    sc->rr = rr;
    if (itr == 1) sc->rr1 = rr;
    if (hist && itr <= sc->hist_cap) hist[itr - 1] = sqrt(rr);
  }
  bool stop;
  bool replace = false;
  if (sc->stop_mode == 0) {  // FEMSolverClass.f90:2275-2289
    stop = (sqrt(rr / rr1) < sc->rel) || (sqrt(rr) < sc->er0);
  } else if (sc->stop_mode == 1) {
    stop = sqrt(rr) < sc->rel * sqrt(sc->bnorm2);
  } else if (sc->stop_mode == 3) {  // bicgstab_CRS: rr / init_rr < er, no square root (:1561)
    stop = rr / sc->rr0 < sc->er0;
  } else {
    // reliable updates: convergence is decided on the FP64 true residual after a
    // replacement; here only decide whether one is due (|r| fell by delta since its max)
    stop = false;
    const double nr = sqrt(rr);
    const double mx = fmax(sc->maxr[par], nr);
    if (first) sc->maxr[par ^ 1] = mx;
    // also replace when the FP32 recurrence has run `stall` iterations since the last one
    replace = nr < sc->delta * mx || (sc->stall > 0 && itr - sc->itr_rep >= sc->stall);
  }
  if (stop) {
    if (first) {
      sc->status = PF_OK;
      sc->done = 1;
    }
    return;
  }
  if (sc->stall > 0 && sc->stop_mode != 2) {
    // inner FP32 solves: end the solve once |r| has not dropped by 10% for `stall`
    // iterations (the recurrence sits at its rounding floor and BiCGSTAB then tends to
    // diverge) or jumps 100x above its best; the FP64 refinement restarts from there
    double best = sc->rr_best[par];
    int ib = sc->itr_best[par];
    if (rr < 0.81 * best) {
      best = rr;
      ib = itr;
    }
    if (first) {
      sc->rr_best[par ^ 1] = best;
      sc->itr_best[par ^ 1] = ib;
    }
    if (itr - ib > sc->stall || rr > 1e4 * best) {
      if (first) {
        sc->status = PF_NOT_CONVERGED;
        sc->done = 1;
      }
      return;
    }
  }
  const double c2w = sc->c2 * sc->omega;
  const double beta_d = t[1] / c2w;
  if (c2w == 0.0 || (sc->guard && !(isfinite(beta_d) && isfinite(rr)))) {  // FEMSolverClass.f90:2295
    if (first) {
      sc->status = PF_BREAKDOWN_C2C3;
      sc->done = 1;
    }
    return;
  }
  if (itr >= sc->itrmax) {
    if (first) {
      sc->status = PF_NOT_CONVERGED;
      sc->done = 1;
    }
    return;
  }
  if (first) {
    sc->c1 = t[1];
    sc->beta = beta_d;
    if (replace) sc->pending = 1;
  }
  const V beta = (V)beta_d, omega = (V)sc->omega;
  if (replace) {
    // fold the FP32 correction into the FP64 solution before the residual is replaced
    for (int64_t i = blockIdx.x * (int64_t)BT + threadIdx.x; i < n; i += (int64_t)gridDim.x * BT) {
      p[i] = r[i] + beta * (p[i] - omega * y[i]);
      xhi[i] += (double)xlo[i];
      xlo[i] = (V)0;
    }
  } else {
    for (int64_t i = blockIdx.x * (int64_t)BT + threadIdx.x; i < n; i += (int64_t)gridDim.x * BT)
      p[i] = r[i] + beta * (p[i] - omega * y[i]);
  }
}

// ---- bicgstab_CRS_ILU (LinearSolverClass.f90:4104-4229) ---------------------------------
// Right-preconditioned: pd = M^-1 p and ps = M^-1 e come from the level-scheduled sweeps
// (the p update and e = r - alpha y are formed inside them), so one iteration is
//   S1 sweep (p, pd) ; K1 y = A pd ; KA alpha ; S2 sweep (e, ps) ; K2 v = A ps ;
//   K3 x += alpha pd + omega ps, r = e - omega v ; K4 stop tests, beta (k_pupdate on no rows)
// First iteration: beta = omega = 0 and y = 0 make the S1 update p = r + 0 (p - 0) = r.
__global__ void __launch_bounds__(BT) k_ilu_start(Scal *__restrict__ sc, double *__restrict__ y, int64_t n) {
  for (int64_t i = blockIdx.x * (int64_t)BT + threadIdx.x; i < n; i += (int64_t)gridDim.x * BT) y[i] = 0.0;
  if (blockIdx.x == 0 && threadIdx.x == 0) {
    sc->beta = 0.0;
    sc->omega = 0.0;
    // "if (c1 < er0) return" on r.r before the loop (LinearSolverClass.f90:4160)
    const bool stop = sc->stop_mode == 1 ? sqrt(sc->c1) < sc->rel * sqrt(sc->bnorm2) : sc->c1 < sc->er0;
    if (stop && sc->itrmax > 0) {
      sc->done = 1;
      sc->status = PF_OK;
    }
  }
}

// c2 = r0.y, alpha = c1 / c2 (one block)
__global__ void __launch_bounds__(BT) k_ilu_alpha(const double *__restrict__ pin, int Gin, Scal *__restrict__ sc) {
  if (sc->done) return;
  double t[1];
  read_partials<1>(pin, Gin, t);
  if (threadIdx.x == 0) {
    const double alpha = sc->c1 / t[0];
    sc->c2 = t[0];
    sc->alpha = alpha;
    if (sc->guard && !isfinite(alpha)) {
      sc->status = PF_BREAKDOWN_C2C3;
      sc->done = 1;
    }
  }
}

// v = A ps -> partials e.v, v.v
template <int L, int SUB>
__global__ void __launch_bounds__(BT) k_spmv_ps(Mat<double> m, const double *__restrict__ ps,
                                                const double *__restrict__ e, double *__restrict__ v,
                                                const Scal *__restrict__ sc, double *__restrict__ part) {
  if (sc->done) return;
  double acc[2] = {0.0, 0.0};
  auto op = [&](int64_t row, double vv) {
    v[row] = vv;
    acc[0] += e[row] * vv;
    acc[1] += vv * vv;
  };
  for_rows<double, L, SUB>(m, pfdev::PlainX<double>{ps}, op);
  write_partials<2>(acc, part);
}

// omega ; x += alpha pd + omega ps ; r = e - omega v -> partials r.r, r0.r
__global__ void __launch_bounds__(BT) k_update_ilu(int64_t n, double *__restrict__ x, double *__restrict__ r,
                                                   const double *__restrict__ r0, const double *__restrict__ pd,
                                                   const double *__restrict__ ps, const double *__restrict__ e,
                                                   const double *__restrict__ v, Scal *__restrict__ sc,
                                                   const double *__restrict__ pin, int Gin, double *__restrict__ part) {
  if (sc->done) return;
  double t[2];
  read_partials<2>(pin, Gin, t);
  const bool first = threadIdx.x == 0 && blockIdx.x == 0;
  if (t[1] == 0.0) {  // "Bicgstab devide by zero" (LinearSolverClass.f90:4196)
    if (first) {
      sc->status = PF_BREAKDOWN_VV;
      sc->done = 1;
    }
    return;
  }
  const double omega = t[0] / t[1];
  if (sc->guard && !isfinite(omega)) {
    if (first) {
      sc->status = PF_BREAKDOWN_VV;
      sc->done = 1;
    }
    return;
  }
  if (first) sc->omega = omega;
  const double alpha = sc->alpha;
  double acc[2] = {0.0, 0.0};
  for (int64_t i = blockIdx.x * (int64_t)BT + threadIdx.x; i < n; i += (int64_t)gridDim.x * BT) {
    x[i] = x[i] + alpha * pd[i] + omega * ps[i];
    const double rn = e[i] - omega * v[i];
    r[i] = rn;
    acc[0] += rn * rn;
    acc[1] += r0[i] * rn;
  }
  write_partials<2>(acc, part);
}

// ---- refinement (mixed precision) -----------------------------------------------------
// r64 = b - A64 x64 -> partials |r|^2, |b|^2
template <int L, int SUB>
__global__ void __launch_bounds__(BT) k_residual64(Mat<double> m, const double *__restrict__ b,
                                                   const double *__restrict__ x, double *__restrict__ r,
                                                   double *__restrict__ part) {
  double acc[2] = {0.0, 0.0};
  auto op = [&](int64_t row, double ax) {
    const double rv = b[row] - ax;
    r[row] = rv;
    acc[0] += rv * rv;
    acc[1] += b[row] * b[row];
  };
  for_rows<double, L, SUB>(m, pfdev::PlainX<double>{x}, op);
  write_partials<2>(acc, part);
}

__global__ void __launch_bounds__(BT) k_residual_finish(const double *__restrict__ part, int G, Scal *__restrict__ sc) {
  double t[2];
  read_partials<2>(part, G, t);
  if (threadIdx.x == 0) {
    sc->rnorm2 = t[0];
    sc->bnorm2 = t[1];
  }
}

// ---- mixed precision: FP32 BiCGSTAB on S A S with FP64 reliable updates -------------
// The FP32 recurrence solves (S A S) z = S b, S = diag(scale).  The solution is kept as
// z_hi (FP64) + z_lo (FP32); whenever the FP32 residual norm has dropped by `delta` since
// its last maximum, z_lo is folded into z_hi and the recursive residual is replaced by the
// FP64 true residual S (b - A S z_hi).  This is iterative refinement without restarting
// the Krylov process; x = S (z_hi + z_lo).
struct ScaledX {
  const double *__restrict__ scale;
  const double *__restrict__ z;
  __device__ double operator()(int64_t j) const { return scale[j] * z[j]; }
};

__global__ void k_rel_setz(int64_t n, const double *__restrict__ x, const double *__restrict__ scale,
                           double *__restrict__ zhi) {
  for (int64_t i = blockIdx.x * (int64_t)BT + threadIdx.x; i < n; i += (int64_t)gridDim.x * BT)
    zhi[i] = x[i] / scale[i];
}

// r' = S (b - A C z_hi) in FP64 (S row scale, C column scale), stored as FP32 r (INIT: also
// r0 = p = r, z_lo = 0; a restarting replacement also sets r0 = p = r)
// partials: |r'|^2, |b - A x|^2, r0.r', |b|^2
template <int L, int SUB, bool INIT, typename VI>
__global__ void __launch_bounds__(BT) k_rel_residual(Mat<double> m, const double *__restrict__ b,
                                                     const double *__restrict__ scale,
                                                     const double *__restrict__ cscale, const double *__restrict__ zhi,
                                                     VI *__restrict__ r, VI *__restrict__ r0,
                                                     VI *__restrict__ p, VI *__restrict__ zlo,
                                                     const Scal *__restrict__ sc, double *__restrict__ part) {
  if (!INIT && !sc->pending) return;
  const bool restart = !INIT && sc->restart;
  double acc[4] = {0.0, 0.0, 0.0, 0.0};
  auto op = [&](int64_t row, double ax) {
    const double bu = b[row] - ax;
    const double rp = scale[row] * bu;
    const VI rf = (VI)rp;
    r[row] = rf;
    if (restart) {
      r0[row] = rf;
      p[row] = rf;
      acc[2] += (double)rf * (double)rf;
    } else if (INIT) {
      r0[row] = rf;
      p[row] = rf;
      zlo[row] = (VI)0;
      acc[2] += (double)rf * (double)rf;
      acc[3] += b[row] * b[row];
    } else {
      acc[2] += (double)r0[row] * (double)rf;
    }
    acc[0] += (double)rf * (double)rf;
    acc[1] += bu * bu;
  };
  for_rows<double, L, SUB>(m, ScaledX{cscale, zhi}, op);
  write_partials<4>(acc, part);
}

// one block: new c1 and |r|, FP64 relative residual, convergence on the true residual
template <bool INIT>
__global__ void __launch_bounds__(BT) k_rel_finish(const double *__restrict__ part, int G, Scal *__restrict__ sc) {
  if (!INIT && !sc->pending) return;
  double t[4];
  read_partials<4>(part, G, t);
  if (threadIdx.x != 0) return;
  sc->itr_rep = INIT ? 0 : sc->itr_last;
  if (INIT) {
    sc->bnorm2 = t[3];
    sc->itr[0] = sc->itr[1] = 0;
    sc->itr_last = 0;
    sc->nupd = 0;
    sc->done = sc->itrmax <= 0 ? 1 : 0;
    sc->status = PF_NOT_CONVERGED;
  } else {
    sc->nupd += 1;
  }
  const double bn = sqrt(sc->bnorm2);
  const double relres = bn > 0.0 ? sqrt(t[1]) / bn : sqrt(t[1]);
  sc->relres = relres;
  sc->c1 = t[2];
  sc->rr = t[0];
  sc->maxr[0] = sc->maxr[1] = sqrt(t[0]);
  sc->pending = 0;
  if (!(relres >= sc->target)) {  // converged, or NaN
    sc->status = isfinite(relres) ? PF_OK : PF_BREAKDOWN_C2C3;
    sc->done = 1;
  }
}

// x = S (z_hi + z_lo)
template <typename VI>
__global__ void k_rel_final(int64_t n, const double *__restrict__ zhi, const VI *__restrict__ zlo,
                            const double *__restrict__ scale, double *__restrict__ x) {
  for (int64_t i = blockIdx.x * (int64_t)BT + threadIdx.x; i < n; i += (int64_t)gridDim.x * BT)
    x[i] = scale[i] * (zhi[i] + (double)zlo[i]);
}

__global__ void k_ones(int64_t n, double *__restrict__ a) {
  for (int64_t i = blockIdx.x * (int64_t)BT + threadIdx.x; i < n; i += (int64_t)gridDim.x * BT) a[i] = 1.0;
}

__global__ void k_f64_to_f32(int64_t n, const double *__restrict__ a, float *__restrict__ b) {
  for (int64_t i = blockIdx.x * (int64_t)blockDim.x + threadIdx.x; i < n; i += (int64_t)gridDim.x * blockDim.x)
    b[i] = (float)a[i];
}

__global__ void k_f32_to_f64(int64_t n, const float *__restrict__ a, double *__restrict__ b) {
  for (int64_t i = blockIdx.x * (int64_t)blockDim.x + threadIdx.x; i < n; i += (int64_t)gridDim.x * blockDim.x)
    b[i] = (double)a[i];
}

}  // namespace

// ========================================================================================
struct SolverWork {
  int64_t n = 0;
  DevBuf<double> r, r0, p, y, v, rref;  // FP64 recurrence vectors (+ FP64 residual)
  DevBuf<double> zhi;                   // mixed: FP64 part of the scaled solution
  DevBuf<double> zlo64;                 // precision 48: the inner (FP64) correction
  DevBuf<float> x32, r32, r032, p32, y32, v32, val32;
  DevBuf<double> part, hist;
  DevBuf<Scal> sc;  // [0] FP64 / outer, [1] FP32 inner
  Scal *h_sc = nullptr;
  Sell3<double> s64;
  Sell3<float> s32;
  CrsStage stage;  // CRS product with x staged in LDS per row block
  DevBuf<int32_t> flag;
  DevBuf<double> jscale;  // FP32 system row scale: 1/sqrt|a_ii| (symmetric) or 1/a_ii (row)
  DevBuf<double> jones;   // column scale of the row-scaled form
  // ILU(0) form: the factor is kept across solves, as FEMSolver_%ILU_MATRIX is
  // (FEMSolverClass.f90:1777-1784), until pf_ilu_reset
  IluFactor ilu;
  DevBuf<double> pd, ps, e;
  hipGraphExec_t gexec_ilu = nullptr;
  uint64_t gkey_ilu[ILU_KEY_WORDS] = {};
  bool have_s64 = false, have_s32 = false, have_v32 = false, have_scale = false;
  int scale_mode = 0;            // what jscale holds: 1 symmetric, 2 row
  int s32_mode = 0, v32_mode = 0;  // scaling of the FP32 copies: 0 none, 1 symmetric, 2 row
  int block_ok = -1;  // -1 unknown
  const void *key_val = nullptr;
  hipGraphExec_t gexec[2] = {nullptr, nullptr};
  uint64_t gkey[2][KEY_WORDS] = {};  // everything a captured chunk bakes in
  std::vector<hipEvent_t> evs;  // kernel_timing: 5 per iteration of a chunk
  double t_sum[4] = {0, 0, 0, 0};
  int64_t t_count[4] = {0, 0, 0, 0};
  ~SolverWork() {
    for (int k = 0; k < 2; ++k)
      if (gexec[k]) (void)hipGraphExecDestroy(gexec[k]);
    if (gexec_ilu) (void)hipGraphExecDestroy(gexec_ilu);
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
  unsigned grid;    // SpMV kernels
  unsigned grid_e;  // element-wise kernels
  int sub;
  int layout;  // 0 CRS, 1 Sell3, 2 CRS with x staged in LDS per row block
};

// partial-sum rows of G doubles each: K1 (1 dot), K2 (2), K3 (2), and one region shared by
// the init, true-residual and replacement kernels (at most 4 dots: k_rel_residual)
constexpr int PART_P1 = 0, PART_P2 = 1, PART_P3 = 3, PART_PI = 5, PART_PI_ROWS = 4;
constexpr int PART_SLOTS = PART_PI + PART_PI_ROWS;
static_assert(PART_P2 >= PART_P1 + 1 && PART_P3 >= PART_P2 + 2 && PART_PI >= PART_P3 + 2,
              "partial rows overlap");

unsigned fused_grid(pf_ctx *ctx, int64_t work_threads) {
  // enough resident waves to cover HBM latency on every CU; the next kernel re-reads one
  // partial per block, so the grid is capped at 8 blocks per CU
  int64_t cap = (int64_t)ctx->num_cus * 8;
  int64_t b = (work_threads + BT - 1) / BT;
  if (b > cap) b = cap;
  if (b < 1) b = 1;
  return (unsigned)b;
}

// element-wise kernels (K3, K4, the e stream, initialisation): every block of the next kernel
// re-sums all partials of the previous one, so these stream with fewer, fuller blocks (measured
// at 100^3: K3 0.042 ms with 8 blocks per CU, 0.035 ms with 4)
constexpr int ELEM_BLOCKS_PER_CU = 4;
unsigned elem_grid(pf_ctx *ctx, int64_t n) {
  int64_t b = (n + BT - 1) / BT;
  const int64_t cap = (int64_t)ctx->num_cus * ELEM_BLOCKS_PER_CU;
  if (b > cap) b = cap;
  if (b < 1) b = 1;
  return (unsigned)b;
}

// every partial row holds one double per block of the writing kernel: grids above G would
// write past w->part
int check_grid(pf_ctx *ctx, const Launch &L) {
  const unsigned G = (unsigned)ctx->num_cus * 8;
  if (L.grid > G || L.grid_e > G || L.grid == 0 || L.grid_e == 0) {
    pf_set_error("internal: launch grid %u/%u outside 1..%u", L.grid, L.grid_e, G);
    return PF_INVALID;
  }
  return PF_OK;
}

int choose_sub(double avg_row) {  // ~five nonzeros per lane, 4..64 lanes per CRS row
  int sub = 4;
  while (sub < 64 && sub * 1.4142 * 5.0 < avg_row) sub *= 2;
  return sub;
}

#define PF_DISPATCH_SUB(SUBV, ...)                          \
  switch (SUBV) {                                           \
    case 4: { constexpr int S_ = 4; __VA_ARGS__; } break;   \
    case 8: { constexpr int S_ = 8; __VA_ARGS__; } break;   \
    case 16: { constexpr int S_ = 16; __VA_ARGS__; } break; \
    case 32: { constexpr int S_ = 32; __VA_ARGS__; } break; \
    default: { constexpr int S_ = 64; __VA_ARGS__; } break; \
  }

// FP64 side of the mixed solve: matrix, launch shape and vectors of the replacements; VI is
// the inner recurrence's vector type (FP32 for precision 32, FP64 for precision 48)
template <typename VI>
struct Rel {
  Mat<double> m;
  Launch L;
  const double *b, *scale, *cscale;  // row and column scale of the inner system
  double *zhi;
  VI *r, *r0, *p, *zlo;
};

template <bool INIT, typename VI>
int enqueue_replacement(pf_ctx *ctx, const Rel<VI> &R, Scal *sc, double *part) {
  static_assert(PART_PI_ROWS >= 4, "k_rel_residual writes 4 partial rows");
  PF_CHECK(check_grid(ctx, R.L));
  hipStream_t st = ctx->stream;
  if (R.L.layout == 1) {
    k_rel_residual<1, 64, INIT><<<R.L.grid, BT, 0, st>>>(R.m, R.b, R.scale, R.cscale, R.zhi, R.r, R.r0, R.p,
                                                         R.zlo, sc, part);
  } else {
    PF_DISPATCH_SUB(R.L.sub, (k_rel_residual<0, S_, INIT><<<R.L.grid, BT, 0, st>>>(
                                 R.m, R.b, R.scale, R.cscale, R.zhi, R.r, R.r0, R.p, R.zlo, sc, part)));
  }
  k_rel_finish<INIT><<<1, BT, 0, st>>>(part, (int)R.L.grid, sc);
  PF_HIP(hipGetLastError());
  return PF_OK;
}

template <typename V, typename M>
int enqueue_init(pf_ctx *ctx, const Launch &L, const Mat<V, M> &m, const V *b, V *x, V *r, V *r0, V *p, Scal *sc,
                 double *part, bool zerox) {
  PF_CHECK(check_grid(ctx, L));
  hipStream_t st = ctx->stream;
  double *pi = part + (size_t)PART_PI * ctx->num_cus * 8;
  const unsigned g = zerox ? L.grid_e : L.grid;
  if (zerox) {
    k_init<V, 0, 4, true><<<g, BT, 0, st>>>(m, b, x, r, r0, p, pi);
  } else if (L.layout == 1) {
    k_init<V, 1, 64, false><<<g, BT, 0, st>>>(m, b, x, r, r0, p, pi);
  } else if (L.layout == 2) {
    PF_DISPATCH_SUB(L.sub, (k_init<V, 2, S_, false><<<g, BT, 0, st>>>(m, b, x, r, r0, p, pi)));
  } else {
    PF_DISPATCH_SUB(L.sub, (k_init<V, 0, S_, false><<<g, BT, 0, st>>>(m, b, x, r, r0, p, pi)));
  }
  k_init_finish<<<1, BT, 0, st>>>(pi, (int)g, sc);
  PF_HIP(hipGetLastError());
  return PF_OK;
}

template <typename V, typename M>
int enqueue_iters(pf_ctx *ctx, const Launch &L, const Mat<V, M> &m, V *x, V *r, const V *r0, V *p, V *y, V *v,
                  Scal *sc, double *part, double *hist, int count, const Rel<V> *rel, V *ebuf,
                  hipEvent_t *ev = nullptr, bool ev_all = true) {
  PF_CHECK(check_grid(ctx, L));
  // ev: events 0 and 1 bracket y = A p; with ev_all, events 2-4 follow the other kernels
  hipEvent_t *eva = ev_all ? ev : nullptr;
  if (L.layout == 2 && !ebuf) {
    pf_set_error("internal: the staged CRS layout needs the e buffer");
    return PF_INVALID;
  }
  hipStream_t st = ctx->stream;
  const size_t G = (size_t)ctx->num_cus * 8;
  double *p1 = part + PART_P1 * G, *p2 = part + PART_P2 * G, *p3 = part + PART_P3 * G;
  const int g1 = (int)L.grid, ge = (int)L.grid_e;
  for (int it = 0; it < count; ++it) {
    const int par = it & 1;
    if (ev) PF_HIP(hipEventRecord(ev[5 * it + 0], st));
    if (L.layout == 1) {
      k_spmv_p<V, 1, 64><<<L.grid, BT, 0, st>>>(m, p, y, r0, sc, p1);
      if (ev) PF_HIP(hipEventRecord(ev[5 * it + 1], st));
      if (ebuf) {  // e streamed, then v = A e gathering e alone
        k_evec<V><<<L.grid_e, BT, 0, st>>>(m.n, r, y, ebuf, sc, p1, g1);
        k_spmv_ev<V, 1, 64><<<L.grid, BT, 0, st>>>(m, ebuf, v, sc, p2);
      } else {
        k_spmv_e<V, 1, 64><<<L.grid, BT, 0, st>>>(m, r, y, v, sc, p1, g1, p2);
      }
    } else if (L.layout == 2) {  // e streamed, then staged per row block like p
      PF_DISPATCH_SUB(L.sub, (k_spmv_p<V, 2, S_><<<L.grid, BT, 0, st>>>(m, p, y, r0, sc, p1)));
      if (ev) PF_HIP(hipEventRecord(ev[5 * it + 1], st));
      k_evec<V><<<L.grid_e, BT, 0, st>>>(m.n, r, y, ebuf, sc, p1, g1);
      PF_DISPATCH_SUB(L.sub, (k_spmv_ev<V, 2, S_><<<L.grid, BT, 0, st>>>(m, ebuf, v, sc, p2)));
    } else {
      PF_DISPATCH_SUB(L.sub, (k_spmv_p<V, 0, S_><<<L.grid, BT, 0, st>>>(m, p, y, r0, sc, p1)));
      if (ev) PF_HIP(hipEventRecord(ev[5 * it + 1], st));
      if (ebuf) {
        k_evec<V><<<L.grid_e, BT, 0, st>>>(m.n, r, y, ebuf, sc, p1, g1);
        PF_DISPATCH_SUB(L.sub, (k_spmv_ev<V, 0, S_><<<L.grid, BT, 0, st>>>(m, ebuf, v, sc, p2)));
      } else {
        PF_DISPATCH_SUB(L.sub, (k_spmv_e<V, 0, S_><<<L.grid, BT, 0, st>>>(m, r, y, v, sc, p1, g1, p2)));
      }
    }
    if (eva) PF_HIP(hipEventRecord(eva[5 * it + 2], st));
    k_update<V><<<L.grid_e, BT, 0, st>>>(m.n, x, r, r0, p, y, v, sc, p2, g1, p3);
    if (eva) PF_HIP(hipEventRecord(eva[5 * it + 3], st));
    k_pupdate<V><<<L.grid_e, BT, 0, st>>>(m.n, p, r, y, sc, p3, ge, par, hist, rel ? rel->zhi : nullptr,
                                          rel ? x : nullptr);
    if (eva) PF_HIP(hipEventRecord(eva[5 * it + 4], st));
    if (rel) PF_CHECK(enqueue_replacement<false>(ctx, *rel, sc, part + PART_PI * G));
  }
  PF_HIP(hipGetLastError());
  return PF_OK;
}

// Run the iteration loop of one BiCGSTAB (scalars slot `slot`) to its stop.
template <typename V, typename M>
int run_loop(pf_ctx *ctx, SolverWork *w, int slot, const Launch &L, const Mat<V, M> &m, V *x, V *r, const V *r0,
             V *p, V *y, V *v, double *hist, const pf_solve_opts &o, const Rel<V> *rel = nullptr, V *ebuf = nullptr) {
  Scal *sc = w->sc.p + slot;
  Scal *hs = w->h_sc + slot;
  int chunk = o.check_every > 0 ? o.check_every : 32;
  chunk += chunk & 1;  // even: the parity of the iteration counter slots carries across chunks
  const bool timing = o.kernel_timing != 0;
  bool use_graph = o.use_graph != 0 && !timing && pf_graphs_allowed();
  if (timing) {
    while ((int)w->evs.size() < 5 * chunk) {
      hipEvent_t e;
      PF_HIP(hipEventCreate(&e));
      w->evs.push_back(e);
    }
  }
  // everything the captured chunk bakes in: a replay is valid only while all of it is
  const uint64_t key[KEY_WORDS] = {
      (uint64_t)(uintptr_t)ebuf, (uint64_t)L.grid, (uint64_t)L.grid_e * 131 + (uint64_t)L.sub, (uint64_t)L.layout,
      (uint64_t)chunk, (uint64_t)m.n, (uint64_t)(uintptr_t)m.row_ptr, (uint64_t)(uintptr_t)m.col_idx,
      (uint64_t)(uintptr_t)m.val, (uint64_t)m.nn, (uint64_t)m.nslices, (uint64_t)(uintptr_t)m.soff,
      (uint64_t)(uintptr_t)m.width, (uint64_t)(uintptr_t)m.col, (uint64_t)(uintptr_t)m.sval,
      (uint64_t)(uintptr_t)x, (uint64_t)(uintptr_t)r, (uint64_t)(uintptr_t)r0, (uint64_t)(uintptr_t)p,
      (uint64_t)(uintptr_t)y, (uint64_t)(uintptr_t)v, (uint64_t)(uintptr_t)sc, (uint64_t)(uintptr_t)w->part.p,
      (uint64_t)(uintptr_t)hist,
      rel ? (uint64_t)(uintptr_t)rel->m.row_ptr : 0, rel ? (uint64_t)(uintptr_t)rel->m.col_idx : 0,
      rel ? (uint64_t)(uintptr_t)rel->m.val : 0, rel ? (uint64_t)(uintptr_t)rel->m.soff : 0,
      rel ? (uint64_t)(uintptr_t)rel->m.width : 0, rel ? (uint64_t)(uintptr_t)rel->m.col : 0,
      rel ? (uint64_t)(uintptr_t)rel->m.sval : 0, rel ? (uint64_t)rel->L.grid * 131 + rel->L.sub : 0,
      rel ? (uint64_t)(uintptr_t)rel->b : 0, rel ? (uint64_t)(uintptr_t)rel->scale : 0,
      rel ? (uint64_t)(uintptr_t)rel->cscale : 0, rel ? (uint64_t)(uintptr_t)rel->zhi : 0,
      rel ? (uint64_t)(uintptr_t)rel->r : 0, rel ? (uint64_t)(uintptr_t)rel->r0 : 0,
      rel ? (uint64_t)(uintptr_t)rel->p : 0, rel ? (uint64_t)(uintptr_t)rel->zlo : 0,
      (uint64_t)(uintptr_t)m.st.lidx, (uint64_t)(uintptr_t)m.st.ucnt, (uint64_t)(uintptr_t)m.st.ubase,
      (uint64_t)(uintptr_t)m.st.ucol, (uint64_t)m.st.nb * 4099 + (uint64_t)m.st.rpb};
  if (use_graph && (w->gexec[slot] == nullptr || std::memcmp(w->gkey[slot], key, sizeof(key)) != 0)) {
    if (w->gexec[slot]) {
      PF_HIP(hipGraphExecDestroy(w->gexec[slot]));
      w->gexec[slot] = nullptr;
    }
    PF_CHECK(pf_capture_graph(
        ctx, [&] { return enqueue_iters<V>(ctx, L, m, x, r, r0, p, y, v, sc, w->part.p, hist, chunk, rel, ebuf); },
        &w->gexec[slot]));
    if (w->gexec[slot])
      std::memcpy(w->gkey[slot], key, sizeof(key));
    else
      use_graph = false;  // direct launches of the same kernels
  }
  for (;;) {
    const int itr_before = hs->itr_last;
    if (use_graph) {
      PF_HIP(hipGraphLaunch(w->gexec[slot], ctx->stream));
    } else {
      PF_CHECK(enqueue_iters<V>(ctx, L, m, x, r, r0, p, y, v, sc, w->part.p, hist, chunk, rel, ebuf,
                                timing ? w->evs.data() : nullptr, o.kernel_timing != 2));
    }
    PF_HIP(hipMemcpyAsync(hs, sc, sizeof(Scal), hipMemcpyDeviceToHost, ctx->stream));
    PF_HIP(hipStreamSynchronize(ctx->stream));
    if (timing) {
      // iterations whose four kernels ran (kernels after a stop return at entry)
      const int live = hs->itr_last - itr_before;
      const int nk = o.kernel_timing == 2 ? 1 : 4;  // 2: y = A p only
      for (int it = 0; it < live && it < chunk; ++it)
        for (int k = 0; k < nk; ++k) {
          float ms = 0.f;
          PF_HIP(hipEventElapsedTime(&ms, w->evs[5 * it + k], w->evs[5 * it + k + 1]));
          w->t_sum[k] += ms;
          w->t_count[k] += 1;
        }
    }
    if (o.debug)
      std::printf("BiCGSTAB >> itr %d  sqrt(rr) %.6e%s\n", hs->itr_last, std::sqrt(hs->rr), rel ? "" : "");
    if (o.debug && rel) std::printf("            replacements %d  FP64 relres %.6e\n", hs->nupd, hs->relres);
    if (hs->done) break;
  }
  return PF_OK;
}

int prepare_scalars(pf_ctx *ctx, SolverWork *w, int slot, const pf_solve_opts &o, int itrmax, double rel,
                    int stop_mode, int hist_cap, int guard, int stall = 0) {
  Scal h;
  std::memset(&h, 0, sizeof(h));
  h.er0 = o.er0;
  h.rel = rel;
  h.itrmax = itrmax;
  h.stop_mode = stop_mode;
  h.hist_cap = hist_cap;
  h.guard = guard;
  h.stall = stall;
  w->h_sc[slot] = h;
  PF_HIP(hipMemcpyAsync(w->sc.p + slot, w->h_sc + slot, sizeof(Scal), hipMemcpyHostToDevice, ctx->stream));
  return PF_OK;
}

// reset cached layouts when the matrix (or its size) changed
void note_matrix(SolverWork *w, int64_t n, const double *val, bool &matrix_changed) {
  if (n != w->n) {
    w->n = n;
    matrix_changed = true;
  }
  if (val != w->key_val) matrix_changed = true;
  w->key_val = val;
  if (matrix_changed) {
    w->have_s64 = w->have_s32 = w->have_v32 = w->have_scale = false;
    w->stage.ready = false;
    w->block_ok = -1;
    std::memset(w->gkey, 0, sizeof(w->gkey));
  }
}

int check_blocks(pf_ctx *ctx, SolverWork *w, int64_t n, const int64_t *row_ptr, const int32_t *col_idx) {
  if (w->block_ok < 0) {
    int ok = 0;
    PF_CHECK(pf_block_structure_check(ctx, n, row_ptr, col_idx, n / 3, w->flag.p, &ok));
    w->block_ok = ok;
  }
  return PF_OK;
}

template <typename T>
struct NonDeduced {
  using type = T;
};

template <typename V, typename M = V>
Mat<V, M> make_mat(int64_t n, const int64_t *row_ptr, const int32_t *col_idx, const M *val,
                   const typename NonDeduced<Sell3<M>>::type *s) {
  Mat<V, M> m{n, row_ptr, col_idx, val, 0, 0, nullptr, nullptr, nullptr, nullptr};
  if (s) {
    m.nn = s->nn;
    m.nslices = s->nslices;
    m.soff = s->soff.p;
    m.width = s->width.p;
    m.col = s->col.p;
    m.sval = s->val.p;
  }
  return m;
}

// one chunk of bicgstab_CRS_ILU iterations (the factor in w->ilu)
int enqueue_iters_ilu(pf_ctx *ctx, SolverWork *w, const Launch &L, const Mat<double> &m, double *x, double *hist,
                      int count) {
  PF_CHECK(check_grid(ctx, L));
  hipStream_t st = ctx->stream;
  Scal *sc = w->sc.p;
  const size_t G = (size_t)ctx->num_cus * 8;
  double *p1 = w->part.p + PART_P1 * G, *p2 = w->part.p + PART_P2 * G, *p3 = w->part.p + PART_P3 * G;
  const int g1 = (int)L.grid, ge = (int)L.grid_e;
  IluSweepArgs s1;
  s1.r = w->r.p;
  s1.y = w->y.p;
  s1.pe = w->p.p;
  s1.beta = &sc->beta;
  s1.omega = &sc->omega;
  s1.done = &sc->done;
  s1.out = w->pd.p;
  IluSweepArgs s2;
  s2.r = w->r.p;
  s2.y = w->y.p;
  s2.pe = w->e.p;
  s2.alpha = &sc->alpha;
  s2.done = &sc->done;
  s2.out = w->ps.p;
  for (int it = 0; it < count; ++it) {
    const int par = it & 1;
    PF_CHECK(pf_ilu_sweep_enqueue(ctx, w->ilu, 1, s1));
    if (L.layout == 1)
      k_spmv_p<double, 1, 64><<<L.grid, BT, 0, st>>>(m, w->pd.p, w->y.p, w->r0.p, sc, p1);
    else
      PF_DISPATCH_SUB(L.sub, (k_spmv_p<double, 0, S_><<<L.grid, BT, 0, st>>>(m, w->pd.p, w->y.p, w->r0.p, sc, p1)));
    k_ilu_alpha<<<1, BT, 0, st>>>(p1, g1, sc);
    PF_CHECK(pf_ilu_sweep_enqueue(ctx, w->ilu, 2, s2));
    if (L.layout == 1)
      k_spmv_ps<1, 64><<<L.grid, BT, 0, st>>>(m, w->ps.p, w->e.p, w->v.p, sc, p2);
    else
      PF_DISPATCH_SUB(L.sub, (k_spmv_ps<0, S_><<<L.grid, BT, 0, st>>>(m, w->ps.p, w->e.p, w->v.p, sc, p2)));
    k_update_ilu<<<L.grid_e, BT, 0, st>>>(m.n, x, w->r.p, w->r0.p, w->pd.p, w->ps.p, w->e.p, w->v.p, sc, p2, g1, p3);
    k_pupdate<double><<<1, BT, 0, st>>>(0, w->p.p, w->r.p, w->y.p, sc, p3, ge, par, hist, nullptr, nullptr);
  }
  PF_HIP(hipGetLastError());
  return PF_OK;
}

// iterations of bicgstab_CRS_ILU to the stop, in hipGraph chunks of 2 (an iteration is
// 2 * levels + 5 launches, so polls are cheap next to it)
int run_loop_ilu(pf_ctx *ctx, SolverWork *w, const Launch &L, const Mat<double> &m, double *x,
                 const pf_solve_opts &o) {
  Scal *sc = w->sc.p;
  Scal *hs = w->h_sc;
  const int chunk = 2;
  bool use_graph = o.use_graph != 0 && pf_graphs_allowed();
  const IluFactor &F = w->ilu;
  uint64_t lev_hash = 1469598103934665603ull ^ (uint64_t)F.G;  // level bounds are launch arguments
  for (int64_t v : F.lev) lev_hash = (lev_hash ^ (uint64_t)v) * 1099511628211ull;
  const uint64_t key[ILU_KEY_WORDS] = {
      (uint64_t)L.grid, (uint64_t)L.grid_e * 131 + (uint64_t)L.sub, (uint64_t)L.layout, (uint64_t)m.n,
      (uint64_t)(uintptr_t)m.row_ptr, (uint64_t)(uintptr_t)m.col_idx, (uint64_t)(uintptr_t)m.val,
      (uint64_t)(uintptr_t)m.soff, (uint64_t)(uintptr_t)m.width, (uint64_t)(uintptr_t)m.col,
      (uint64_t)(uintptr_t)m.sval, (uint64_t)(uintptr_t)x, (uint64_t)(uintptr_t)w->sc.p,
      (uint64_t)(uintptr_t)w->part.p, (uint64_t)(uintptr_t)w->r.p, (uint64_t)(uintptr_t)w->r0.p,
      (uint64_t)(uintptr_t)w->p.p, (uint64_t)(uintptr_t)w->y.p, (uint64_t)(uintptr_t)w->v.p,
      (uint64_t)(uintptr_t)w->e.p, (uint64_t)(uintptr_t)w->pd.p, (uint64_t)(uintptr_t)w->ps.p,
      (uint64_t)(uintptr_t)F.grp.p, (uint64_t)(uintptr_t)F.lptr.p, (uint64_t)(uintptr_t)F.lcol.p,
      (uint64_t)(uintptr_t)F.lval.p, (uint64_t)(uintptr_t)F.wv.p, (uint64_t)(uintptr_t)F.dd.p,
      (uint64_t)(uintptr_t)F.t.p, (uint64_t)F.lev.size(), (uint64_t)F.lnnz, lev_hash};
  if (use_graph && (w->gexec_ilu == nullptr || std::memcmp(w->gkey_ilu, key, sizeof(key)) != 0)) {
    if (w->gexec_ilu) {
      PF_HIP(hipGraphExecDestroy(w->gexec_ilu));
      w->gexec_ilu = nullptr;
    }
    PF_CHECK(pf_capture_graph(ctx, [&] { return enqueue_iters_ilu(ctx, w, L, m, x, nullptr, chunk); },
                              &w->gexec_ilu));
    if (w->gexec_ilu)
      std::memcpy(w->gkey_ilu, key, sizeof(key));
    else
      use_graph = false;
  }
  for (;;) {
    if (use_graph) {
      PF_HIP(hipGraphLaunch(w->gexec_ilu, ctx->stream));
    } else {
      PF_CHECK(enqueue_iters_ilu(ctx, w, L, m, x, nullptr, chunk));
    }
    PF_HIP(hipMemcpyAsync(hs, sc, sizeof(Scal), hipMemcpyDeviceToHost, ctx->stream));
    PF_HIP(hipStreamSynchronize(ctx->stream));
    if (o.debug) std::printf("BiCGSTAB ILU >> itr %d  sqrt(rr) %.6e\n", hs->itr_last, std::sqrt(hs->rr));
    if (hs->done) break;
  }
  return PF_OK;
}

// Iterative refinement with an inner BiCGSTAB on the FP32 copy of the (scaled) matrix: VI =
// float keeps the inner vectors and sums in FP32 (precision 32), VI = double keeps them in
// FP64 and only the matrix values in FP32 (precision 48).  The FP64 side holds the solution
// z_hi, the true residual and the stop test on ||b - A x|| / ||b||.
template <typename VI>
int solve_refined(pf_ctx *ctx, SolverWork *w, int64_t n, const int64_t *row_ptr, const int32_t *col_idx,
                  const double *val, const double *b, double *x, const pf_solve_opts &o, int layout, int sub,
                  int smode, const double *scale, const double *cscale, SolveResult &res) {
  const size_t G = (size_t)ctx->num_cus * 8;
  const unsigned ge = elem_grid(ctx, n);
  const int64_t sell_threads = ((n / 3 + 63) / 64) * 64;
  PF_CHECK(w->zhi.ensure(n));
  VI *zlo, *r, *r0, *p, *y, *v;
  if constexpr (std::is_same<VI, float>::value) {
    PF_CHECK(w->x32.ensure(n));
    PF_CHECK(w->r32.ensure(n));
    PF_CHECK(w->r032.ensure(n));
    PF_CHECK(w->p32.ensure(n));
    PF_CHECK(w->y32.ensure(n));
    PF_CHECK(w->v32.ensure(n));
    zlo = w->x32.p;
    r = w->r32.p;
    r0 = w->r032.p;
    p = w->p32.p;
    y = w->y32.p;
    v = w->v32.p;
  } else {
    PF_CHECK(w->zlo64.ensure(n));
    zlo = w->zlo64.p;
    r = w->r.p;
    r0 = w->r0.p;
    p = w->p.p;
    y = w->y.p;
    v = w->v.p;
  }
  const Mat<VI, float> mi = make_mat<VI, float>(n, row_ptr, col_idx, w->have_v32 ? w->val32.p : nullptr,
                                                layout == 1 ? &w->s32 : nullptr);
  const Launch Li{fused_grid(ctx, layout == 1 ? sell_threads : n * (int64_t)sub), ge, sub, layout};
  Rel<VI> R;
  R.m = make_mat<double>(n, row_ptr, col_idx, val, layout == 1 ? &w->s64 : nullptr);
  R.L = Launch{fused_grid(ctx, layout == 1 ? sell_threads : n * (int64_t)sub), ge, sub, layout};
  R.b = b;
  R.scale = scale;
  R.cscale = cscale;
  R.zhi = w->zhi.p;
  R.r = r;
  R.r0 = r0;
  R.p = p;
  R.zlo = zlo;
  PF_CHECK(prepare_scalars(ctx, w, 1, o, o.itrmax, 0.0, 2, 0, 1, o.inner_max > 0 ? o.inner_max : 0));
  w->h_sc[1].target = o.target_relres;
  w->h_sc[1].delta = o.inner_tol > 0.0 && o.inner_tol < 1.0 ? o.inner_tol : 0.1;
  w->h_sc[1].restart = smode == 2 ? 1 : 0;
  PF_HIP(hipMemcpyAsync(w->sc.p + 1, w->h_sc + 1, sizeof(Scal), hipMemcpyHostToDevice, ctx->stream));
  k_rel_setz<<<ge, BT, 0, ctx->stream>>>(n, x, cscale, w->zhi.p);
  PF_HIP(hipGetLastError());
  PF_CHECK(enqueue_replacement<true>(ctx, R, w->sc.p + 1, w->part.p + PART_PI * G));
  // precision 48 on the blocked layout: e = r - alpha y streamed once, then v = A e gathers e
  // alone (its FP32 values make the two FP64 gathers of the fused form a larger share)
  VI *ebuf = nullptr;
  if constexpr (std::is_same<VI, double>::value) {
    if (layout == 1) {
      PF_CHECK(w->e.ensure(n));
      ebuf = w->e.p;
    }
  }
  PF_CHECK((run_loop<VI, float>(ctx, w, 1, Li, mi, zlo, r, r0, p, y, v, nullptr, o, &R, ebuf)));
  k_rel_final<<<ge, BT, 0, ctx->stream>>>(n, w->zhi.p, zlo, cscale, x);
  PF_HIP(hipGetLastError());
  const Scal &hi = w->h_sc[1];
  res.status = hi.status;
  res.iterations = hi.itr_last;
  res.outer = hi.nupd;
  res.rr = hi.rr;
  res.rr1 = 0.0;
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
  if (o.stop_mode < 0 || o.stop_mode > 2) {
    pf_set_error("stop_mode must be 0 (bicgstab_CRS_2), 1 (|r| < relative_er |b|) or 2 (bicgstab_CRS)");
    return PF_INVALID;
  }
  if (o.stop_mode == 2 && (o.precision != 64 || o.preconditioning == 2)) {
    pf_set_error("stop_mode 2 (bicgstab_CRS) applies to the unpreconditioned FP64 solver");
    return PF_INVALID;
  }
  if (n <= 0) return PF_OK;
  note_matrix(w, n, val, matrix_changed);
  PF_CHECK(w->r.ensure(n));
  PF_CHECK(w->r0.ensure(n));
  PF_CHECK(w->p.ensure(n));
  PF_CHECK(w->y.ensure(n));
  PF_CHECK(w->v.ensure(n));
  PF_CHECK(w->sc.ensure(2));
  PF_CHECK(w->flag.ensure(1));
  PF_CHECK(w->part.ensure((size_t)ctx->num_cus * 8 * PART_SLOTS + 16));

  PF_HIP(hipEventRecord(ctx->ev0, ctx->stream));
  const double avg_row = (double)nnz / (double)n;
  int layout = 0;
  if (o.spmv_kind < 0 || o.spmv_kind > 4 || o.spmv_kind == 3) {
    pf_set_error("spmv_kind must be 0 (auto), 1 (CRS rows), 2 (blocked) or 4 (CRS rows, x staged in LDS)");
    return PF_INVALID;
  }
  if (o.spmv_kind != 1 && o.spmv_kind != 4) {
    PF_CHECK(check_blocks(ctx, w, n, row_ptr, col_idx));
    layout = w->block_ok ? 1 : 0;
    if (o.spmv_kind == 2 && !layout) {
      pf_set_error("blocked SpMV requested but the CRS has no DOF=3 node-block structure");
      return PF_INVALID;
    }
    if (layout == 1 && o.spmv_kind == 0) {
      if (!w->have_s64) {
        PF_CHECK(pf_sell3_build<double>(ctx, n, row_ptr, col_idx, val, w->s64));
        w->have_s64 = true;
      }
      // the sliced copy pads every 64-node slice to its widest row: when row lengths vary a
      // lot (the rows coupled by overset penalties) it outgrows the CRS, which then streams
      // fewer bytes (matrices below a million entries stay on the sliced copy: launch
      // latency, not bandwidth, sets their time).  Measured on overset plants: values padded
      // 1.48x, y = A p + v = A e 0.495 ms blocked against 0.384 ms on CRS rows.
      if (nnz >= ((int64_t)1 << 20) && (double)w->s64.nslots * 64.0 * 9.0 > 1.2 * (double)nnz) layout = 0;
    }
  }
  if (o.precision != 64 && o.precision != 32 && o.precision != 48) {
    pf_set_error("precision must be 64 (FP64), 32 (FP32 values and vectors) or 48 (FP32 values, FP64 vectors)");
    return PF_INVALID;
  }
  const bool mixed = o.precision == 32;
  const bool refined = o.precision != 64;  // precision 32 or 48: refinement over an FP32 matrix copy
  // FP32 system: 0 A, 1 S A S (S = diag^-1/2, unknown scaled), 2 D^-1 A (rows only, restarted)
  const int smode = refined ? (o.jacobi == 2 ? 2 : (o.jacobi != 0 ? 1 : 0)) : 0;
  if (smode != 0 && (!w->have_scale || w->scale_mode != smode)) {
    PF_CHECK(w->jscale.ensure(n));
    if (smode == 1) {
      PF_CHECK(pf_jacobi_scale(ctx, n, row_ptr, col_idx, val, w->jscale.p));
    } else {
      PF_CHECK(pf_jacobi_row_scale(ctx, n, row_ptr, col_idx, val, w->jscale.p));
      PF_CHECK(w->jones.ensure(n));
      k_ones<<<fused_grid(ctx, n), BT, 0, ctx->stream>>>(n, w->jones.p);
      PF_HIP(hipGetLastError());
    }
    w->have_scale = true;
    w->scale_mode = smode;
  }
  const double *scale = smode ? w->jscale.p : nullptr;
  const double *cscale = smode == 2 ? w->jones.p : scale;
  if (layout == 1) {
    if (!w->have_s64) {  // FP64 products (mixed: the residual replacements)
      PF_CHECK(pf_sell3_build<double>(ctx, n, row_ptr, col_idx, val, w->s64));
      w->have_s64 = true;
    }
    if (refined && (!w->have_s32 || w->s32_mode != smode)) {
      PF_CHECK(pf_sell3_build<float>(ctx, n, row_ptr, col_idx, val, w->s32, scale, nullptr, cscale));
      w->have_s32 = true;
      w->s32_mode = smode;
    }
  } else if (refined && (!w->have_v32 || w->v32_mode != smode)) {
    PF_CHECK(w->val32.ensure((size_t)nnz + 1));
    PF_CHECK(pf_crs_to_f32(ctx, n, row_ptr, col_idx, val, scale, w->val32.p, cscale));
    w->have_v32 = true;
    w->v32_mode = smode;
  }
  const bool ilu = o.preconditioning == 2;
  if (ilu) {
    if (refined) {
      pf_set_error("ILU(0) preconditioning runs in FP64 only (precision 64)");
      return PF_INVALID;
    }
    if (w->block_ok < 0) PF_CHECK(check_blocks(ctx, w, n, row_ptr, col_idx));
    if (!w->ilu.ready || w->ilu.n != n || w->ilu.nnz != nnz)
      PF_CHECK(pf_ilu_factor(ctx, n, row_ptr, col_idx, val, nnz, w->block_ok == 1 ? 3 : 1, w->ilu));
    PF_CHECK(w->pd.ensure(n));
    PF_CHECK(w->ps.ensure(n));
    PF_CHECK(w->e.ensure(n));
  }
  // the FP64 solve on CRS rows stages x per row block (spmv_kind 4; auto from 2^18 rows: y = A p
  // plus v = A e 1.458 -> 1.188 ms on the 3.09M-row hex box and 0.383 -> 0.346 ms on a 465k-row
  // overset system, but 0.095 -> 0.100 ms on a 148k-row one, whose few row blocks per
  // workgroup leave the staging gathers exposed)
  if (o.spmv_kind == 4 && (refined || ilu)) {
    pf_set_error("spmv_kind 4 (x staged in LDS) runs the FP64 solve without ILU only");
    return PF_INVALID;
  }
  if (layout == 0 && !refined && !ilu && (o.spmv_kind == 4 || (o.spmv_kind == 0 && n >= ((int64_t)1 << 18)))) {
    if (!w->stage.ready) PF_CHECK(pf_crs_stage_build(ctx, n, row_ptr, col_idx, nnz, w->stage));
    layout = 2;
  }
  PF_HIP(hipEventRecord(ctx->ev1, ctx->stream));
  PF_HIP(hipEventSynchronize(ctx->ev1));
  float setup_ms = 0.f;
  PF_HIP(hipEventElapsedTime(&setup_ms, ctx->ev0, ctx->ev1));
  res.setup_ms = setup_ms;

  const int sub = choose_sub(avg_row);
  const int64_t sell_threads = ((n / 3 + 63) / 64) * 64;
  // FP64 CRS rows: refinement residual and the final true residual
  Mat<double> mcrs = make_mat<double>(n, row_ptr, col_idx, val, nullptr);
  Launch Lres{fused_grid(ctx, n * (int64_t)sub), elem_grid(ctx, n), sub, 0};
  res.kind = layout == 1 ? 2 : (layout == 2 ? 4 : 1);
  const size_t G = (size_t)ctx->num_cus * 8;
  double *pres = w->part.p + PART_PI * G;

  PF_HIP(hipEventRecord(ctx->ev0, ctx->stream));
  if (ilu) {
    Mat<double> m64 = make_mat<double>(n, row_ptr, col_idx, val, layout == 1 ? &w->s64 : nullptr);
    Launch L64{fused_grid(ctx, layout == 1 ? sell_threads : n * (int64_t)sub), elem_grid(ctx, n), sub, layout};
    const int guard = o.stop_mode == 1 ? 1 : 0;
    PF_CHECK(prepare_scalars(ctx, w, 0, o, o.itrmax, o.relative_er, o.stop_mode == 1 ? 1 : 0, 0, guard));
    PF_CHECK(enqueue_init<double>(ctx, L64, m64, b, x, w->r.p, w->r0.p, w->p.p, w->sc.p, w->part.p, false));
    k_ilu_start<<<L64.grid_e, BT, 0, ctx->stream>>>(w->sc.p, w->y.p, n);
    PF_HIP(hipGetLastError());
    PF_CHECK(run_loop_ilu(ctx, w, L64, m64, x, o));
    const Scal &h = w->h_sc[0];
    res.status = h.status;
    res.iterations = h.itr_last + (h.status == PF_BREAKDOWN_VV ? 1 : 0);
    res.rr = h.rr;
    res.rr1 = h.rr1;
  } else if (!refined) {
    Mat<double> m64 = make_mat<double>(n, row_ptr, col_idx, val, layout == 1 ? &w->s64 : nullptr);
    Launch L64{fused_grid(ctx, layout == 1 ? sell_threads : n * (int64_t)sub), elem_grid(ctx, n), sub, layout};
    if (layout == 2) {
      m64.st = pfdev::Stage{w->stage.lidx.p, w->stage.ucnt.p, w->stage.ubase.p, w->stage.ucol.p, w->stage.nb,
                            w->stage.rpb};
      L64.grid = fused_grid(ctx, w->stage.nb * (int64_t)BT);
      L64.sub = w->stage.sub;
    }
    const int guard = o.stop_mode == 1 ? 1 : 0;  // modes 0 and 2 keep the reference's unguarded arithmetic
    // kernel stop rules: 0 bicgstab_CRS_2, 1 |r| < relative_er |b|, 3 bicgstab_CRS (er = er0)
    const int kmode = o.stop_mode == 2 ? 3 : o.stop_mode;
    PF_CHECK(prepare_scalars(ctx, w, 0, o, o.itrmax, o.relative_er, kmode, 0, guard));
    PF_CHECK(enqueue_init<double>(ctx, L64, m64, b, x, w->r.p, w->r0.p, w->p.p, w->sc.p, w->part.p, false));
    double *ebuf = nullptr;
    if (layout != 1) {  // CRS rows: e as a stream, then the product gathers (or stages) e alone
      PF_CHECK(w->e.ensure(n));
      ebuf = w->e.p;
    }
    PF_CHECK(run_loop<double>(ctx, w, 0, L64, m64, x, w->r.p, w->r0.p, w->p.p, w->y.p, w->v.p, nullptr, o, nullptr,
                              ebuf));
    const Scal &h = w->h_sc[0];
    res.status = h.status;
    // the reference's loop counter: a v.v == 0 stop happens inside iteration itr+1
    res.iterations = h.itr_last + (h.status == PF_BREAKDOWN_VV ? 1 : 0);
    res.rr = h.rr;
    res.rr1 = h.rr1;
  } else {
    // BiCGSTAB on S A C with the FP32 copy of the matrix (S = C = Jacobi scaling or identity;
    // or S = D^-1, C = I with a restart at every replacement) and FP64 reliable updates:
    // FP32 vectors for precision 32, FP64 vectors for precision 48
    if (!scale) {  // unscaled: S = C = I
      PF_CHECK(w->jscale.ensure(n));
      k_ones<<<fused_grid(ctx, n), BT, 0, ctx->stream>>>(n, w->jscale.p);
      PF_HIP(hipGetLastError());
      scale = cscale = w->jscale.p;
      w->have_scale = false;
    }
    if (mixed)
      PF_CHECK(solve_refined<float>(ctx, w, n, row_ptr, col_idx, val, b, x, o, layout, sub, smode, scale, cscale,
                                    res));
    else
      PF_CHECK(solve_refined<double>(ctx, w, n, row_ptr, col_idx, val, b, x, o, layout, sub, smode, scale, cscale,
                                     res));
  }
  PF_HIP(hipEventRecord(ctx->ev1, ctx->stream));
  PF_HIP(hipEventSynchronize(ctx->ev1));
  float ms = 0.f;
  PF_HIP(hipEventElapsedTime(&ms, ctx->ev0, ctx->ev1));
  res.solve_ms = ms;
  if (res.status == PF_BREAKDOWN_VV)
    pf_set_error("BiCGSTAB: v.v == 0 in iteration %d", res.iterations);
  else if (res.status == PF_BREAKDOWN_C2C3)
    pf_set_error("BiCGSTAB: c2*omega == 0 or a non-finite scalar after %d iterations", res.iterations);
  else if (res.status == PF_NOT_CONVERGED)
    pf_set_error("BiCGSTAB: not converged after %d iterations", res.iterations);
  for (int k = 0; k < 4; ++k) res.t_ms[k] = w->t_count[k] ? w->t_sum[k] / (double)w->t_count[k] : 0.0;
  res.timed = w->t_count[0];

  // true residual of the solved system in FP64 (CRS rows, independent of the solve layout)
  {
    PF_CHECK(w->rref.ensure(n));
    PF_CHECK(check_grid(ctx, Lres));
    PF_DISPATCH_SUB(Lres.sub, (k_residual64<0, S_><<<Lres.grid, BT, 0, ctx->stream>>>(mcrs, b, x, w->rref.p, pres)));
    k_residual_finish<<<1, BT, 0, ctx->stream>>>(pres, (int)Lres.grid, w->sc.p);
    PF_HIP(hipGetLastError());
    PF_HIP(hipMemcpyAsync(w->h_sc, w->sc.p, sizeof(Scal), hipMemcpyDeviceToHost, ctx->stream));
    PF_HIP(hipStreamSynchronize(ctx->stream));
    const double bn = std::sqrt(w->h_sc[0].bnorm2);
    res.true_relres = bn > 0 ? std::sqrt(w->h_sc[0].rnorm2) / bn : std::sqrt(w->h_sc[0].rnorm2);
  }
  return PF_OK;
}

size_t pf_bicgstab_workspace_doubles(pf_ctx *ctx, int64_t n) {
  const size_t G = (size_t)ctx->num_cus * 8;
  const size_t scal = (sizeof(Scal) + sizeof(double) - 1) / sizeof(double);
  return 7 * (size_t)n + PART_SLOTS * G + 16 + 8 + scal;
}

int pf_bicgstab_crs_workspace(pf_ctx *ctx, int64_t n, const int64_t *row_ptr, const int32_t *col_idx,
                              const double *val, int64_t nnz, const double *b, double *x, const pf_solve_opts &o,
                              double *ws, size_t ws_doubles, SolveResult &res) {
  res = SolveResult();
  if (o.precision != 64 || o.preconditioning != 0 || o.kernel_timing != 0) {
    pf_set_error("workspace solver: precision 64, no preconditioning and no kernel timing only");
    return PF_INVALID;
  }
  if (o.stop_mode < 0 || o.stop_mode > 2) {
    pf_set_error("stop_mode must be 0 (bicgstab_CRS_2), 1 (|r| < relative_er |b|) or 2 (bicgstab_CRS)");
    return PF_INVALID;
  }
  if (n <= 0) return PF_OK;
  const size_t G = (size_t)ctx->num_cus * 8;
  if (pf_bicgstab_workspace_doubles(ctx, n) > ws_doubles) {
    pf_set_error("workspace solver: %zu doubles of workspace needed, %zu available",
                 pf_bicgstab_workspace_doubles(ctx, n), ws_doubles);
    return PF_INVALID;
  }
  // workspace: r, r0, p, y, v, e, rref (n each), the partial rows, the scalars (64-byte aligned)
  double *r = ws, *r0 = r + n, *p = r0 + n, *y = p + n, *v = y + n, *e = v + n, *rref = e + n;
  double *part = rref + n;
  double *sc_d = part + PART_SLOTS * G + 16;
  sc_d += (8 - ((uintptr_t)sc_d / sizeof(double)) % 8) % 8;
  Scal *sc = reinterpret_cast<Scal *>(sc_d);

  const int sub = choose_sub((double)nnz / (double)n);
  const Mat<double> m = make_mat<double>(n, row_ptr, col_idx, val, nullptr);
  const Launch L{fused_grid(ctx, n * (int64_t)sub), elem_grid(ctx, n), sub, 0};
  Scal h;
  std::memset(&h, 0, sizeof(h));
  h.er0 = o.er0;
  h.rel = o.relative_er;
  h.itrmax = o.itrmax;
  h.stop_mode = o.stop_mode == 2 ? 3 : o.stop_mode;
  h.guard = o.stop_mode == 1 ? 1 : 0;
  PF_HIP(hipEventRecord(ctx->ev0, ctx->stream));
  PF_HIP(hipMemcpyAsync(sc, &h, sizeof(Scal), hipMemcpyHostToDevice, ctx->stream));
  PF_CHECK(enqueue_init<double>(ctx, L, m, b, x, r, r0, p, sc, part, false));
  int chunk = o.check_every > 0 ? o.check_every : 32;
  chunk += chunk & 1;  // even: the iteration counter slots alternate by parity
  for (;;) {
    PF_CHECK(enqueue_iters<double>(ctx, L, m, x, r, r0, p, y, v, sc, part, nullptr, chunk, nullptr, e));
    PF_HIP(hipMemcpyAsync(&h, sc, sizeof(Scal), hipMemcpyDeviceToHost, ctx->stream));
    PF_HIP(hipStreamSynchronize(ctx->stream));
    if (o.debug) std::printf("BiCGSTAB >> itr %d  sqrt(rr) %.6e\n", h.itr_last, std::sqrt(h.rr));
    if (h.done) break;
  }
  PF_HIP(hipEventRecord(ctx->ev1, ctx->stream));
  PF_HIP(hipEventSynchronize(ctx->ev1));
  float ms = 0.f;
  PF_HIP(hipEventElapsedTime(&ms, ctx->ev0, ctx->ev1));
  res.solve_ms = ms;
  res.kind = 1;
  res.status = h.status;
  res.iterations = h.itr_last + (h.status == PF_BREAKDOWN_VV ? 1 : 0);
  res.rr = h.rr;
  res.rr1 = h.rr1;
  if (res.status == PF_BREAKDOWN_VV)
    pf_set_error("BiCGSTAB: v.v == 0 in iteration %d", res.iterations);
  else if (res.status == PF_BREAKDOWN_C2C3)
    pf_set_error("BiCGSTAB: c2*omega == 0 or a non-finite scalar after %d iterations", res.iterations);
  else if (res.status == PF_NOT_CONVERGED)
    pf_set_error("BiCGSTAB: not converged after %d iterations", res.iterations);
  // true residual in FP64
  double *pres = part + (size_t)PART_PI * G;
  PF_CHECK(check_grid(ctx, L));
  PF_DISPATCH_SUB(L.sub, (k_residual64<0, S_><<<L.grid, BT, 0, ctx->stream>>>(m, b, x, rref, pres)));
  k_residual_finish<<<1, BT, 0, ctx->stream>>>(pres, (int)L.grid, sc);
  PF_HIP(hipGetLastError());
  PF_HIP(hipMemcpyAsync(&h, sc, sizeof(Scal), hipMemcpyDeviceToHost, ctx->stream));
  PF_HIP(hipStreamSynchronize(ctx->stream));
  const double bn = std::sqrt(h.bnorm2);
  res.true_relres = bn > 0 ? std::sqrt(h.rnorm2) / bn : std::sqrt(h.rnorm2);
  return PF_OK;
}

int pf_work_ilu(pf_ctx *ctx, SolverWork *w, int64_t n, const int64_t *row_ptr, const int32_t *col_idx,
                const double *val, int64_t nnz, bool matrix_changed, IluFactor **out) {
  note_matrix(w, n, val, matrix_changed);
  PF_CHECK(w->flag.ensure(1));
  if (w->block_ok < 0) PF_CHECK(check_blocks(ctx, w, n, row_ptr, col_idx));
  if (!w->ilu.ready || w->ilu.n != n || w->ilu.nnz != nnz)
    PF_CHECK(pf_ilu_factor(ctx, n, row_ptr, col_idx, val, nnz, w->block_ok == 1 ? 3 : 1, w->ilu));
  *out = &w->ilu;
  return PF_OK;
}

void pf_work_ilu_reset(SolverWork *w) {
  if (w) w->ilu.ready = false;
}

int pf_work_spmv(pf_ctx *ctx, SolverWork *w, int64_t n, const int64_t *row_ptr, const int32_t *col_idx,
                 const double *val, int64_t nnz, bool matrix_changed, int kind, const double *x, double *y,
                 int repeat, double *ms) {
  if (ms) *ms = 0.0;
  if (n <= 0) return PF_OK;
  if (kind < 1 || kind > 4) {
    pf_set_error("spmv kind must be 1, 2, 3 or 4");
    return PF_INVALID;
  }
  note_matrix(w, n, val, matrix_changed);
  PF_CHECK(w->flag.ensure(1));
  if (kind == 4 && !w->stage.ready) PF_CHECK(pf_crs_stage_build(ctx, n, row_ptr, col_idx, nnz, w->stage));
  if (kind == 2 || kind == 3) {
    PF_CHECK(check_blocks(ctx, w, n, row_ptr, col_idx));
    if (!w->block_ok) {
      pf_set_error("blocked SpMV requested but the CRS has no DOF=3 node-block structure");
      return PF_INVALID;
    }
    if (kind == 2 && !w->have_s64) {
      PF_CHECK(pf_sell3_build<double>(ctx, n, row_ptr, col_idx, val, w->s64));
      w->have_s64 = true;
    }
    if (kind == 3 && (!w->have_s32 || w->s32_mode != 0)) {
      PF_CHECK(pf_sell3_build<float>(ctx, n, row_ptr, col_idx, val, w->s32));
      w->have_s32 = true;
      w->s32_mode = 0;
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
    else if (kind == 4)
      PF_CHECK(pf_spmv_crs_stage_launch<double>(ctx, w->stage, row_ptr, col_idx, val, x, y));
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
