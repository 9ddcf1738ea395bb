// multi.hip -- bicgstab_CRS_2 (FEMSolverClass.f90:2199-2298) for many right-hand sides of one
// matrix at once: the load cases of a forward solve and the adjoint solves of the
// differentiable form.
//
// Vectors are interleaved, entry (dof i, case l) at i*R + l, so one wavefront holds the R
// cases of a node in consecutive lanes: the 3x3 blocks of the matrix row of node i are read
// once per wavefront (same address in every lane) and applied to R cases, and every gather
// of x is a contiguous R-wide load.  Each case keeps its own scalars (alpha, omega, beta,
// r.r, stop state); a case that has stopped leaves its x, r, p untouched while the others
// continue.  Per iteration: five streaming kernels (the two products with their fused dot
// products, e = r - alpha y, the x/r update, the p update) and three one-block kernels that
// reduce the per-block partial sums of every case in a fixed order and advance its scalars.
//
// The matrix is read in a node-block CSR copy of the system's CRS (9 values of a 3x3
// block contiguous, one int32 node column per block), built once per matrix.
#include <cmath>
#include <cstring>
#include <vector>

#include "pf_internal.hpp"

namespace {

constexpr int MBT = 256;  // threads per block of the streaming kernels
constexpr int MKEY_WORDS = 20;  // words of the key of a captured iteration chunk

struct MScal {
  double c1, c2, alpha, omega, beta, rr, rr1, bnorm2;
  int32_t itr, done, status, pad;
};

struct MCtl {
  int32_t all_done, itrmax, stop_mode, nrhs;
  double er0, rel;
};

struct Blk {
  int64_t nn;
  const int64_t *noff;   // nn+1 block offsets
  const int32_t *bcol;   // node column per block
  const double *bval;    // 9 values per block, row-major
};

// per-block partial sums of case l for dot k: part[(k * R + l) * G + block]
template <int RP, int NV>
__device__ inline void case_partials(double (&v)[NV], int R, double *part) {
  __shared__ double sh[NV][MBT / 64][RP];
  const int lane = threadIdx.x & 63, w = threadIdx.x >> 6;
#pragma unroll
  for (int k = 0; k < NV; ++k)
#pragma unroll
    for (int off = RP; off < 64; off <<= 1) v[k] += __shfl_xor(v[k], off, 64);
  if (lane < RP)
#pragma unroll
    for (int k = 0; k < NV; ++k) sh[k][w][lane] = v[k];
  __syncthreads();
  if (threadIdx.x < RP && (int)threadIdx.x < R) {
#pragma unroll
    for (int k = 0; k < NV; ++k) {
      double t = 0.0;
#pragma unroll
      for (int j = 0; j < MBT / 64; ++j) t += sh[k][j][threadIdx.x];
      part[((int64_t)k * R + threadIdx.x) * gridDim.x + blockIdx.x] = t;
    }
  }
}

// blocks per step of the row loop: their node columns, x gathers and values are all in
// flight before the first product of the step (a hex8 node row holds at most 27 blocks).
// The node-block arrays end with MCHUNK_MAX zero blocks, so a step may read past its row.
constexpr int MCHUNK_MAX = 9;
template <int RP>
constexpr int mchunk() {
  return RP == 64 ? 9 : 3;  // block values in scalar registers with one node per wavefront
}

// every node row of y = A x for the calling thread's case lane; op(dof, case, sum).  The
// blocks are summed in stored order.  Lanes past the last case load case 0 (in bounds) and
// drop the result, so no load waits behind a lane condition.
template <int RP, typename XF, typename OP>
__device__ inline void node_rows(const Blk &B, int R, XF xf, OP op) {
  constexpr int MC = mchunk<RP>();
  static_assert(MC <= MCHUNK_MAX, "row step longer than the zero blocks after the last row");
  const int l = threadIdx.x % RP;
  const bool live = l < R;
  const int lc = live ? l : 0;
  // Workgroups are dispatched to the 8 XCDs in turn (block b on XCD b % 8).  With a grid of
  // a multiple of 8 blocks, XCD k walks its own eighth of the nodes in order, so the nodes in
  // flight on one XCD are consecutive and the R-wide x rows they gather (up to 27 per node,
  // 1.5 KB each with 64 cases) are shared in that XCD's L2 instead of fetched by several.
  int64_t i_lo = 0, i_hi = B.nn, g0, ng;
  if (gridDim.x % 8 == 0) {
    const int64_t k = blockIdx.x % 8;
    i_lo = k * B.nn / 8;
    i_hi = (k + 1) * B.nn / 8;
    g0 = i_lo + (blockIdx.x / 8) * (MBT / RP) + threadIdx.x / RP;
    ng = (int64_t)(gridDim.x / 8) * (MBT / RP);
  } else {
    g0 = (blockIdx.x * (int64_t)MBT + threadIdx.x) / RP;
    ng = (int64_t)gridDim.x * (MBT / RP);
  }
  for (int64_t i0 = g0; i0 < i_hi; i0 += ng) {
    // with 64 cases a wavefront owns one node: make the node (and so every matrix address
    // below) wave-uniform so the block values come through scalar loads
    const int64_t i = RP == 64 ? (int64_t)__builtin_amdgcn_readfirstlane((int)i0) : i0;
    const int64_t b0 = B.noff[i], b1 = B.noff[i + 1];
    double a0 = 0.0, a1 = 0.0, a2 = 0.0;
    for (int64_t blk = b0; blk < b1; blk += MC) {
      int64_t j[MC];
#pragma unroll
      for (int u = 0; u < MC; ++u) j[u] = 3 * (int64_t)B.bcol[blk + u];
      double x[MC][3];
#pragma unroll
      for (int u = 0; u < MC; ++u) {
        x[u][0] = xf(j[u], lc);
        x[u][1] = xf(j[u] + 1, lc);
        x[u][2] = xf(j[u] + 2, lc);
      }
#pragma unroll
      for (int u = 0; u < MC; ++u) {
        const double *q = B.bval + (blk + u) * 9;
        const double t0 = q[0] * x[u][0] + q[1] * x[u][1] + q[2] * x[u][2];
        const double t1 = q[3] * x[u][0] + q[4] * x[u][1] + q[5] * x[u][2];
        const double t2 = q[6] * x[u][0] + q[7] * x[u][1] + q[8] * x[u][2];
        const bool own = blk + u < b1;  // a select, not a branch: no load waits behind it
        a0 = own ? a0 + t0 : a0;
        a1 = own ? a1 + t1 : a1;
        a2 = own ? a2 + t2 : a2;
      }
    }
    if (live) {
      op(3 * i, l, a0);
      op(3 * i + 1, l, a1);
      op(3 * i + 2, l, a2);
    }
  }
}

// every (dof, case) of the calling thread, in the same thread -> case mapping as node_rows
template <int RP, typename OP>
__device__ inline void node_entries(int64_t nn, int R, OP op) {
  const int l = threadIdx.x % RP;
  if (l >= R) return;
  const int64_t g0 = (blockIdx.x * (int64_t)MBT + threadIdx.x) / RP;
  const int64_t ng = (int64_t)gridDim.x * (MBT / RP);
  for (int64_t i = g0; i < nn; i += ng)
#pragma unroll
    for (int d = 0; d < 3; ++d) op(3 * i + d, l);
}

// one-block reductions: 16 lanes per case (cases 0..63 in a 1024-thread block) sum the G
// partials of their case in a fixed order and leave the total in the case's lane 0
constexpr int RBT = 1024, RLANES = 16;

__device__ inline double case_total(const double *__restrict__ part, int64_t row, int G) {
  const int j = threadIdx.x % RLANES;
  double t = 0.0;
  for (int b = j; b < G; b += RLANES) t += part[row * G + b];
#pragma unroll
  for (int off = RLANES / 2; off > 0; off >>= 1) t += __shfl_xor(t, off, RLANES);
  return t;
}

struct Vec {
  const double *__restrict__ v;
  int R;
  __device__ double operator()(int64_t j, int l) const { return v[j * R + l]; }
};


// ---- setup: b~ = b with the fixed rows set to 0 (ubar = 0); x = 0; r = r0 = p = b~ ------
template <int RP>
__global__ void __launch_bounds__(MBT) k_m_init(int64_t nn, int R, const double *__restrict__ b,
                                                const uint8_t *__restrict__ fixed, double *__restrict__ bt,
                                                double *__restrict__ x, double *__restrict__ r,
                                                double *__restrict__ r0, double *__restrict__ p,
                                                double *__restrict__ part) {
  double acc[2] = {0.0, 0.0};
  node_entries<RP>(nn, R, [&](int64_t i, int l) {
    const int64_t g = i * R + l;
    const double v = (fixed && fixed[i]) ? 0.0 : b[g];
    bt[g] = v;
    x[g] = 0.0;
    r[g] = v;
    r0[g] = v;
    p[g] = v;
    acc[0] += v * v;
  });
  acc[1] = acc[0];
  case_partials<RP, 2>(acc, R, part);
}

// with an initial guess: x_fixed = 0, r = b~ - A x, r0 = p = r (b~ already written by k_m_init)
template <int RP>
__global__ void __launch_bounds__(MBT) k_m_init_x0(Blk B, int R, const double *__restrict__ bt,
                                                   const double *__restrict__ x, double *__restrict__ r,
                                                   double *__restrict__ r0, double *__restrict__ p,
                                                   double *__restrict__ part) {
  double acc[2] = {0.0, 0.0};
  node_rows<RP>(B, R, Vec{x, R}, [&](int64_t i, int l, double ax) {
    const int64_t g = i * R + l;
    const double v = bt[g] - ax;
    r[g] = v;
    r0[g] = v;
    p[g] = v;
    acc[0] += v * v;
    acc[1] += bt[g] * bt[g];
  });
  case_partials<RP, 2>(acc, R, part);
}

template <int RP>
__global__ void __launch_bounds__(MBT) k_m_fix_x0(int64_t nn, int R, const double *__restrict__ b,
                                                  const uint8_t *__restrict__ fixed, double *__restrict__ bt,
                                                  double *__restrict__ x) {
  node_entries<RP>(nn, R, [&](int64_t i, int l) {
    const int64_t g = i * R + l;
    const bool f = fixed && fixed[i];
    bt[g] = f ? 0.0 : b[g];
    if (f) x[g] = 0.0;
  });
}

__global__ void __launch_bounds__(RBT) k_m_init_finish(const double *__restrict__ part, int G, MScal *__restrict__ sc,
                                                       MCtl *__restrict__ ctl) {
  const int l = threadIdx.x / RLANES;
  const int R = ctl->nrhs;
  if (l < R) {
    const double t = case_total(part, l, G);      // r.r
    const double bb = case_total(part, R + l, G);  // b.b
    if (threadIdx.x % RLANES == 0) {
      MScal s{};
      s.c1 = t;
      s.rr = t;
      s.bnorm2 = bb;
      s.status = PF_NOT_CONVERGED;
      s.done = (ctl->itrmax <= 0 || !(t > 0.0)) ? 1 : 0;  // r = 0: x (0, or the guess) solves it
      if (!(t > 0.0)) s.status = PF_OK;
      sc[l] = s;
    }
  }
  __syncthreads();
  if (threadIdx.x == 0) {
    int all = 1;
    for (int k = 0; k < R; ++k) all &= sc[k].done;
    ctl->all_done = all;
  }
}

// ---- K1: y = A p ; partials r0.y -----------------------------------------------------
template <int RP>
__global__ void __launch_bounds__(MBT) k_m_spmv_p(Blk B, int R, const double *__restrict__ p, double *__restrict__ y,
                                                  const double *__restrict__ r0, const MCtl *__restrict__ ctl,
                                                  double *__restrict__ part) {
  if (ctl->all_done) return;
  double acc[1] = {0.0};
  node_rows<RP>(B, R, Vec{p, R}, [&](int64_t i, int l, double yv) {
    y[i * R + l] = yv;
    acc[0] += r0[i * R + l] * yv;
  });
  case_partials<RP, 1>(acc, R, part);
}

__global__ void __launch_bounds__(RBT) k_m_alpha(const double *__restrict__ part, int G, MScal *__restrict__ sc,
                                                 const MCtl *__restrict__ ctl) {
  const int l = threadIdx.x / RLANES;
  if (ctl->all_done || l >= ctl->nrhs || sc[l].done) return;
  const double t = case_total(part, l, G);
  if (threadIdx.x % RLANES == 0) {
    sc[l].c2 = t;
    sc[l].alpha = sc[l].c1 / t;
  }
}

// ---- K2: v = A e, e = r - alpha y ; partials e.v, v.v ---------------------------------
// e = r - alpha y, stored once so the product below gathers one vector instead of two (with
// 64 cases the gathers, 27 x 1.5 KB per node, dominate the product's traffic)
template <int RP>
__global__ void __launch_bounds__(MBT) k_m_evec(int64_t nn, int R, const double *__restrict__ r,
                                                const double *__restrict__ y, double *__restrict__ e,
                                                const MScal *__restrict__ sc, const MCtl *__restrict__ ctl) {
  if (ctl->all_done) return;
  const int lc = threadIdx.x % RP;
  const double al = lc < R ? sc[lc].alpha : 0.0;
  node_entries<RP>(nn, R, [&](int64_t i, int l) {
    const int64_t g = i * R + l;
    e[g] = r[g] - al * y[g];
  });
}

template <int RP>
__global__ void __launch_bounds__(MBT) k_m_spmv_e(Blk B, int R, const double *__restrict__ e,
                                                  double *__restrict__ v, const MCtl *__restrict__ ctl,
                                                  double *__restrict__ part) {
  if (ctl->all_done) return;
  double acc[2] = {0.0, 0.0};
  node_rows<RP>(B, R, Vec{e, R}, [&](int64_t i, int l, double vv) {
    v[i * R + l] = vv;
    acc[0] += e[i * R + l] * vv;
    acc[1] += vv * vv;
  });
  case_partials<RP, 2>(acc, R, part);
}

__global__ void __launch_bounds__(RBT) k_m_omega(const double *__restrict__ part, int G, MScal *__restrict__ sc,
                                                 const MCtl *__restrict__ ctl) {
  const int l = threadIdx.x / RLANES;
  if (ctl->all_done || l >= ctl->nrhs || sc[l].done) return;
  const int R = ctl->nrhs;
  const double ev = case_total(part, l, G);
  const double vv = case_total(part, R + l, G);
  if (threadIdx.x % RLANES != 0) return;
  if (vv == 0.0 || !isfinite(sc[l].alpha)) {  // FEMSolverClass.f90:2266
    sc[l].status = vv == 0.0 ? PF_BREAKDOWN_VV : PF_BREAKDOWN_C2C3;
    sc[l].itr += vv == 0.0 ? 1 : 0;
    sc[l].done = 1;
    return;
  }
  sc[l].omega = ev / vv;
}

// ---- K3: x += alpha p + omega e ; r = e - omega v ; partials r.r, r0.r ------------------
template <int RP>
__global__ void __launch_bounds__(MBT) k_m_update(int64_t nn, int R, double *__restrict__ x, double *__restrict__ r,
                                                  const double *__restrict__ r0, const double *__restrict__ p,
                                                  const double *__restrict__ e, const double *__restrict__ v,
                                                  const MScal *__restrict__ sc, const MCtl *__restrict__ ctl,
                                                  double *__restrict__ part) {
  if (ctl->all_done) return;
  const int lc = threadIdx.x % RP;
  const bool skip = lc >= R || sc[lc].done;
  const double al = skip ? 0.0 : sc[lc].alpha, om = skip ? 0.0 : sc[lc].omega;
  double acc[2] = {0.0, 0.0};
  node_entries<RP>(nn, R, [&](int64_t i, int l) {
    if (skip) return;
    const int64_t g = i * R + l;
    const double eg = e[g];
    x[g] = x[g] + al * p[g] + om * eg;
    const double rn = eg - om * v[g];
    r[g] = rn;
    acc[0] += rn * rn;
    acc[1] += r0[g] * rn;
  });
  case_partials<RP, 2>(acc, R, part);
}

__global__ void __launch_bounds__(RBT) k_m_beta(const double *__restrict__ part, int G, MScal *__restrict__ sc,
                                                MCtl *__restrict__ ctl) {
  const int l = threadIdx.x / RLANES;
  const int R = ctl->nrhs;
  const bool live = !ctl->all_done && l < R && !sc[l].done;
  if (live) {
    const double rr = case_total(part, l, G);
    const double r0r = case_total(part, R + l, G);
    if (threadIdx.x % RLANES == 0) {
      MScal s = sc[l];
      s.itr += 1;
      s.rr = rr;
      if (s.itr == 1) s.rr1 = rr;
      const bool stop = ctl->stop_mode == 0 ? (sqrt(rr / s.rr1) < ctl->rel || sqrt(rr) < ctl->er0)
                                            : sqrt(rr) < ctl->rel * sqrt(s.bnorm2);
      const double c2w = s.c2 * s.omega;
      const double beta = r0r / c2w;
      if (stop) {
        s.status = PF_OK;
        s.done = 1;
      } else if (c2w == 0.0 || !isfinite(beta) || !isfinite(rr)) {  // FEMSolverClass.f90:2295
        s.status = PF_BREAKDOWN_C2C3;
        s.done = 1;
      } else if (s.itr >= ctl->itrmax) {
        s.status = PF_NOT_CONVERGED;
        s.done = 1;
      } else {
        s.c1 = r0r;
        s.beta = beta;
      }
      sc[l] = s;
    }
  }
  __syncthreads();
  if (threadIdx.x == 0 && !ctl->all_done) {
    int all = 1;
    for (int k = 0; k < R; ++k) all &= sc[k].done;
    ctl->all_done = all;
  }
}

// ---- K4: p = r + beta (p - omega y) -----------------------------------------------------
template <int RP>
__global__ void __launch_bounds__(MBT) k_m_pupdate(int64_t nn, int R, double *__restrict__ p,
                                                   const double *__restrict__ r, const double *__restrict__ y,
                                                   const MScal *__restrict__ sc, const MCtl *__restrict__ ctl) {
  if (ctl->all_done) return;
  const int lc = threadIdx.x % RP;
  if (lc >= R || sc[lc].done) return;
  const double be = sc[lc].beta, om = sc[lc].omega;
  node_entries<RP>(nn, R, [&](int64_t i, int l) {
    const int64_t g = i * R + l;
    p[g] = r[g] + be * (p[g] - om * y[g]);
  });
}

// ---- after the loop: x_fixed = 0 and the FP64 true residual of every case --------------
template <int RP>
__global__ void __launch_bounds__(MBT) k_m_final(Blk B, int R, const uint8_t *__restrict__ fixed,
                                                 double *__restrict__ x) {
  node_entries<RP>(B.nn, R, [&](int64_t i, int l) {
    if (fixed && fixed[i]) x[i * R + l] = 0.0;
  });
}

template <int RP>
__global__ void __launch_bounds__(MBT) k_m_residual(Blk B, int R, const double *__restrict__ bt,
                                                    const double *__restrict__ x, double *__restrict__ part) {
  double acc[2] = {0.0, 0.0};
  node_rows<RP>(B, R, Vec{x, R}, [&](int64_t i, int l, double ax) {
    const double bv = bt[i * R + l];
    const double rv = bv - ax;
    acc[0] += rv * rv;
    acc[1] += bv * bv;
  });
  case_partials<RP, 2>(acc, R, part);
}

// node-block CSR from a DOF=3 CRS with node-block structure (one thread per node)
__global__ void k_bcsr_build(int64_t nn, const int64_t *__restrict__ row_ptr, const int32_t *__restrict__ col_idx,
                             const double *__restrict__ val, int64_t *__restrict__ noff, int32_t *__restrict__ bcol,
                             double *__restrict__ bval) {
  const int64_t i = blockIdx.x * (int64_t)blockDim.x + threadIdx.x;
  if (i > nn) return;
  const int64_t o = (row_ptr[3 * i] - 1) / 9;
  noff[i] = o;
  if (i == nn) {  // the zero blocks after the last row
    for (int t = 0; t < MCHUNK_MAX; ++t) {
      bcol[o + t] = 0;
      for (int k = 0; k < 9; ++k) bval[(o + t) * 9 + k] = 0.0;
    }
    return;
  }
  const int64_t w = (row_ptr[3 * i + 1] - row_ptr[3 * i]) / 3;
  for (int64_t t = 0; t < w; ++t) {
    bcol[o + t] = (col_idx[row_ptr[3 * i] - 1 + 3 * t] - 1) / 3;
    for (int d = 0; d < 3; ++d)
      for (int k = 0; k < 3; ++k) bval[(o + t) * 9 + 3 * d + k] = val[row_ptr[3 * i + d] - 1 + 3 * t + k];
  }
}

}  // namespace

// ========================================================================================
struct MultiWork {
  int64_t n = 0;
  int R = 0;
  DevBuf<int64_t> noff;
  DevBuf<int32_t> bcol;
  DevBuf<double> bval;
  bool have_bcsr = false;
  uint64_t mat_version = ~0ull;
  DevBuf<double> bt, r, r0, p, y, v, e, part;
  DevBuf<MScal> sc;
  DevBuf<MCtl> ctl;
  MCtl *h_ctl = nullptr;
  hipGraphExec_t gexec = nullptr;
  uint64_t gkey[MKEY_WORDS] = {};
  ~MultiWork() {
    if (gexec) (void)hipGraphExecDestroy(gexec);
    if (h_ctl) (void)hipHostFree(h_ctl);
  }
};

void pf_multi_work_destroy(MultiWork *m) { delete m; }

namespace {

int rp_of(int R) {
  int rp = 1;
  while (rp < R) rp <<= 1;
  return rp;
}

#define PF_DISPATCH_RP(RPV, ...)                              \
  switch (RPV) {                                              \
    case 1: { constexpr int RP_ = 1; __VA_ARGS__; } break;    \
    case 2: { constexpr int RP_ = 2; __VA_ARGS__; } break;    \
    case 4: { constexpr int RP_ = 4; __VA_ARGS__; } break;    \
    case 8: { constexpr int RP_ = 8; __VA_ARGS__; } break;    \
    case 16: { constexpr int RP_ = 16; __VA_ARGS__; } break;  \
    case 32: { constexpr int RP_ = 32; __VA_ARGS__; } break;  \
    default: { constexpr int RP_ = 64; __VA_ARGS__; } break;  \
  }

struct MLaunch {
  unsigned grid;
  int rp;
};

int enqueue_chunk(pf_ctx *ctx, MultiWork *w, const MLaunch &L, const Blk &B, double *x, int count) {
  hipStream_t st = ctx->stream;
  const int R = w->R;
  const int G = (int)L.grid;
  double *p1 = w->part.p, *p2 = w->part.p + (size_t)R * G, *p3 = w->part.p + (size_t)3 * R * G;
  for (int it = 0; it < count; ++it) {
    PF_DISPATCH_RP(L.rp, (k_m_spmv_p<RP_><<<L.grid, MBT, 0, st>>>(B, R, w->p.p, w->y.p, w->r0.p, w->ctl.p, p1)));
    k_m_alpha<<<1, RBT, 0, st>>>(p1, G, w->sc.p, w->ctl.p);
    PF_DISPATCH_RP(L.rp, (k_m_evec<RP_><<<L.grid, MBT, 0, st>>>(B.nn, R, w->r.p, w->y.p, w->e.p, w->sc.p, w->ctl.p)));
    PF_DISPATCH_RP(L.rp, (k_m_spmv_e<RP_><<<L.grid, MBT, 0, st>>>(B, R, w->e.p, w->v.p, w->ctl.p, p2)));
    k_m_omega<<<1, RBT, 0, st>>>(p2, G, w->sc.p, w->ctl.p);
    PF_DISPATCH_RP(L.rp, (k_m_update<RP_><<<L.grid, MBT, 0, st>>>(B.nn, R, x, w->r.p, w->r0.p, w->p.p, w->e.p, w->v.p,
                                                                  w->sc.p, w->ctl.p, p3)));
    k_m_beta<<<1, RBT, 0, st>>>(p3, G, w->sc.p, w->ctl.p);
    PF_DISPATCH_RP(L.rp, (k_m_pupdate<RP_><<<L.grid, MBT, 0, st>>>(B.nn, R, w->p.p, w->r.p, w->y.p, w->sc.p, w->ctl.p)));
  }
  PF_HIP(hipGetLastError());
  return PF_OK;
}

}  // namespace

int pf_multi_solve_device(pf_ctx *ctx, MultiWork **pw, int64_t n, const int64_t *row_ptr, const int32_t *col_idx,
                          const double *val, uint64_t mat_version, const uint8_t *fixed, int32_t nrhs,
                          const double *B_dev, double *X_dev, const pf_solve_opts &o, pf_stats *per_rhs,
                          pf_stats *total) {
  if (nrhs < 1 || nrhs > 64) {
    pf_set_error("multi-RHS solve: 1 <= nrhs <= 64");
    return PF_INVALID;
  }
  if (n % 3 != 0) {
    pf_set_error("multi-RHS solve needs a DOF=3 system");
    return PF_INVALID;
  }
  if (o.stop_mode != 0 && o.stop_mode != 1) {
    pf_set_error("multi-RHS solve: stop_mode must be 0 (bicgstab_CRS_2) or 1 (|r| < relative_er |b|)");
    return PF_INVALID;
  }
  if (!*pw) {
    *pw = new MultiWork();
    hipError_t e = hipHostMalloc((void **)&(*pw)->h_ctl, sizeof(MCtl), hipHostMallocDefault);
    if (e != hipSuccess) {
      pf_set_error("hipHostMalloc failed: %s", hipGetErrorString(e));
      return PF_ERR_NOMEM;
    }
  }
  MultiWork *w = *pw;
  const int64_t nn = n / 3;
  if (w->n != n || w->mat_version != mat_version || !w->have_bcsr) {
    int ok = 0;
    PF_CHECK(w->bcol.ensure(1));
    DevBuf<int32_t> flag;
    PF_CHECK(flag.ensure(1));
    PF_CHECK(pf_block_structure_check(ctx, n, row_ptr, col_idx, nn, flag.p, &ok));
    if (!ok) {
      pf_set_error("multi-RHS solve needs the DOF=3 node-block CRS that ZeroMatrix(3) builds");
      return PF_INVALID;
    }
    int64_t nnz1 = 0;
    PF_CHECK(pf_read_i64(ctx, row_ptr + n, &nnz1));
    const int64_t nblk = (nnz1 - 1) / 9;
    PF_CHECK(w->noff.ensure((size_t)nn + 1));
    PF_CHECK(w->bcol.ensure((size_t)nblk + MCHUNK_MAX));
    PF_CHECK(w->bval.ensure(((size_t)nblk + MCHUNK_MAX) * 9));
    k_bcsr_build<<<pf_blocks(nn + 1, 256, INT32_MAX), 256, 0, ctx->stream>>>(nn, row_ptr, col_idx, val, w->noff.p,
                                                                           w->bcol.p, w->bval.p);
    PF_HIP(hipGetLastError());
    w->have_bcsr = true;
    w->mat_version = mat_version;
    w->n = n;
    std::memset(w->gkey, 0, sizeof(w->gkey));
  }
  const int R = nrhs;
  w->R = R;
  const size_t nv = (size_t)n * R;
  PF_CHECK(w->bt.ensure(nv));
  PF_CHECK(w->r.ensure(nv));
  PF_CHECK(w->r0.ensure(nv));
  PF_CHECK(w->p.ensure(nv));
  PF_CHECK(w->y.ensure(nv));
  PF_CHECK(w->v.ensure(nv));
  PF_CHECK(w->e.ensure(nv));
  PF_CHECK(w->sc.ensure(64));
  PF_CHECK(w->ctl.ensure(1));
  MLaunch L;
  L.rp = rp_of(R);
  {
    // at most 4 blocks per CU (16 waves): every block leaves one partial per case that the
    // one-block reductions sum (measured at 40^3, 64 cases: 8 blocks/CU -> 106 us of
    // reductions per iteration)
    const int64_t nodes_per_block = MBT / L.rp;
    int64_t g = (nn + nodes_per_block - 1) / nodes_per_block;
    const int64_t cap = (int64_t)ctx->num_cus * 4;
    L.grid = (unsigned)std::max<int64_t>(1, std::min(g, cap));
  }
  PF_CHECK(w->part.ensure((size_t)5 * R * L.grid + 16));
  const Blk B{nn, w->noff.p, w->bcol.p, w->bval.p};
  MCtl hc;
  std::memset(&hc, 0, sizeof(hc));
  hc.itrmax = o.itrmax;
  hc.stop_mode = o.stop_mode == 1 ? 1 : 0;  // 0 or 1 (checked at entry)
  hc.nrhs = R;
  hc.er0 = o.er0;
  hc.rel = o.relative_er;
  *w->h_ctl = hc;
  PF_HIP(hipMemcpyAsync(w->ctl.p, w->h_ctl, sizeof(MCtl), hipMemcpyHostToDevice, ctx->stream));
  PF_HIP(hipEventRecord(ctx->ev0, ctx->stream));
  if (o.use_x0) {
    PF_DISPATCH_RP(L.rp, (k_m_fix_x0<RP_><<<L.grid, MBT, 0, ctx->stream>>>(nn, R, B_dev, fixed, w->bt.p, X_dev)));
    PF_DISPATCH_RP(L.rp, (k_m_init_x0<RP_><<<L.grid, MBT, 0, ctx->stream>>>(B, R, w->bt.p, X_dev, w->r.p, w->r0.p,
                                                                           w->p.p, w->part.p)));
  } else {
    PF_DISPATCH_RP(L.rp, (k_m_init<RP_><<<L.grid, MBT, 0, ctx->stream>>>(nn, R, B_dev, fixed, w->bt.p, X_dev,
                                                                        w->r.p, w->r0.p, w->p.p, w->part.p)));
  }
  k_m_init_finish<<<1, RBT, 0, ctx->stream>>>(w->part.p, (int)L.grid, w->sc.p, w->ctl.p);
  PF_HIP(hipGetLastError());
  int chunk = o.check_every > 0 ? o.check_every : 32;
  const bool use_graph = o.use_graph != 0 && pf_graphs_allowed();
  // everything the captured chunk bakes in
  const uint64_t key[MKEY_WORDS] = {
      (uint64_t)L.grid, (uint64_t)L.rp, (uint64_t)R, (uint64_t)chunk, (uint64_t)n, (uint64_t)(uintptr_t)X_dev,
      (uint64_t)B.nn, (uint64_t)(uintptr_t)B.noff, (uint64_t)(uintptr_t)B.bcol, (uint64_t)(uintptr_t)B.bval,
      (uint64_t)(uintptr_t)w->r.p, (uint64_t)(uintptr_t)w->r0.p, (uint64_t)(uintptr_t)w->p.p,
      (uint64_t)(uintptr_t)w->y.p, (uint64_t)(uintptr_t)w->v.p, (uint64_t)(uintptr_t)w->e.p,
      (uint64_t)(uintptr_t)w->part.p, (uint64_t)(uintptr_t)w->sc.p, (uint64_t)(uintptr_t)w->ctl.p,
      (uint64_t)(uintptr_t)w->bt.p};
  bool graph_ok = false;
  if (use_graph) {
    if (!w->gexec || std::memcmp(w->gkey, key, sizeof(key)) != 0) {
      if (w->gexec) {
        PF_HIP(hipGraphExecDestroy(w->gexec));
        w->gexec = nullptr;
      }
      PF_CHECK(pf_capture_graph(ctx, [&] { return enqueue_chunk(ctx, w, L, B, X_dev, chunk); }, &w->gexec));
      if (w->gexec) std::memcpy(w->gkey, key, sizeof(key));
    }
    graph_ok = w->gexec != nullptr;
  }
  for (;;) {
    if (graph_ok) {
      PF_HIP(hipGraphLaunch(w->gexec, ctx->stream));
    } else {
      PF_CHECK(enqueue_chunk(ctx, w, L, B, X_dev, chunk));
    }
    PF_HIP(hipMemcpyAsync(w->h_ctl, w->ctl.p, sizeof(MCtl), hipMemcpyDeviceToHost, ctx->stream));
    PF_HIP(hipStreamSynchronize(ctx->stream));
    if (w->h_ctl->all_done) break;
  }
  PF_HIP(hipEventRecord(ctx->ev1, ctx->stream));
  PF_HIP(hipEventSynchronize(ctx->ev1));
  float ms = 0.f;
  PF_HIP(hipEventElapsedTime(&ms, ctx->ev0, ctx->ev1));
  // x_fixed = 0, FP64 true residual of every case
  double *pres = w->part.p + (size_t)3 * R * L.grid;
  PF_DISPATCH_RP(L.rp, (k_m_final<RP_><<<L.grid, MBT, 0, ctx->stream>>>(B, R, fixed, X_dev)));
  PF_DISPATCH_RP(L.rp, (k_m_residual<RP_><<<L.grid, MBT, 0, ctx->stream>>>(B, R, w->bt.p, X_dev, pres)));
  PF_HIP(hipGetLastError());
  std::vector<MScal> hs(R);
  std::vector<double> hp((size_t)2 * R * L.grid);
  PF_HIP(hipMemcpyAsync(hs.data(), w->sc.p, sizeof(MScal) * R, hipMemcpyDeviceToHost, ctx->stream));
  PF_HIP(hipMemcpyAsync(hp.data(), pres, sizeof(double) * hp.size(), hipMemcpyDeviceToHost, ctx->stream));
  PF_HIP(hipStreamSynchronize(ctx->stream));
  int status = PF_OK;
  int64_t iters = 0;
  int maxit = 0;
  double worst = 0.0;
  for (int l = 0; l < R; ++l) {
    double r2 = 0.0, b2 = 0.0;
    for (unsigned b = 0; b < L.grid; ++b) {
      r2 += hp[(size_t)l * L.grid + b];
      b2 += hp[(size_t)(R + l) * L.grid + b];
    }
    const double rel = b2 > 0.0 ? std::sqrt(r2 / b2) : std::sqrt(r2);
    worst = std::max(worst, rel);
    iters += hs[l].itr;
    maxit = std::max(maxit, hs[l].itr);
    if (status == PF_OK && hs[l].status != PF_OK) status = hs[l].status;
    if (per_rhs) {
      pf_stats &s = per_rhs[l];
      std::memset(&s, 0, sizeof(s));
      s.status = hs[l].status;
      s.iterations = hs[l].itr;
      s.spmv_kind_used = 2;
      s.rr = hs[l].rr;
      s.rr1 = hs[l].rr1;
      s.true_relres = rel;
      s.solve_ms = ms;
    }
  }
  if (total) {
    std::memset(total, 0, sizeof(*total));
    total->status = status;
    total->iterations = maxit;
    total->spmv_kind_used = 2;
    total->true_relres = worst;
    total->solve_ms = ms;
    total->timed_iterations = iters;
  }
  if (status != PF_OK) pf_set_error("multi-RHS BiCGSTAB: a case ended with status %d", status);
  return status;
}
