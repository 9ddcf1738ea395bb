// spmv.hip -- y = A x kernels and the blocked (Sell3) layout built from a DOF=3 CRS.
#include <vector>

#include "spmv_dev.hpp"

namespace {

// y = A x, or y = y + A x (ACC: the reference's C kernels crs_spmv_real64/real32,
// SparseClass-c.c:38-66, add into new_vector)
template <typename V, int SUB, bool ACC>
__global__ void __launch_bounds__(256) k_spmv_crs(int64_t n, const int64_t *__restrict__ row_ptr,
                                                  const int32_t *__restrict__ col_idx,
                                                  const V *__restrict__ val, const V *__restrict__ x,
                                                  V *__restrict__ y) {
  const int64_t g = blockIdx.x * (int64_t)blockDim.x + threadIdx.x;
  const int64_t row = g / SUB;
  const int sub = (int)(g % SUB);
  if (row >= n) return;  // whole subgroups leave together (SUB divides 64)
  V s = pfdev::crs_row<V, V, SUB>(row_ptr, col_idx, val, row, sub, pfdev::PlainX<V>{x});
  if (sub == 0) y[row] = ACC ? y[row] + s : s;
}

// the same row sums for several vectors (X, Y column-major, leading dimension n): a subgroup
// takes one row and every column in turn, so the row's values and column indices come from
// HBM once and from cache for the other columns
template <typename V, int SUB>
__global__ void __launch_bounds__(256) k_spmm_crs(int64_t n, int64_t cols, const int64_t *__restrict__ row_ptr,
                                                  const int32_t *__restrict__ col_idx,
                                                  const V *__restrict__ val, const V *__restrict__ x,
                                                  V *__restrict__ y) {
  const int64_t g = blockIdx.x * (int64_t)blockDim.x + threadIdx.x;
  const int64_t row = g / SUB;
  const int sub = (int)(g % SUB);
  if (row >= n) return;  // whole subgroups leave together (SUB divides 64)
  for (int64_t c = 0; c < cols; ++c) {
    const V s = pfdev::crs_row<V, V, SUB>(row_ptr, col_idx, val, row, sub, pfdev::PlainX<V>{x + c * n});
    if (sub == 0) y[row + c * n] = s;
  }
}

template <typename V>
__global__ void __launch_bounds__(256) k_spmv_sell3(int64_t nn, int64_t nslices,
                                                    const int64_t *__restrict__ soff,
                                                    const int32_t *__restrict__ width,
                                                    const int32_t *__restrict__ col,
                                                    const V *__restrict__ val, const V *__restrict__ x,
                                                    V *__restrict__ y) {
  const int64_t i = blockIdx.x * (int64_t)blockDim.x + threadIdx.x;
  const int64_t s = i >> 6;
  if (s >= nslices) return;
  V y0, y1, y2;
  pfdev::sell3_node<V, V>(soff, width, col, val, s, (int)(i & 63), pfdev::PlainX<V>{x}, y0, y1, y2);
  if (i < nn) {
    y[3 * i] = y0;
    y[3 * i + 1] = y1;
    y[3 * i + 2] = y2;
  }
}

// rows 3i..3i+2 share one node-column list and each node column c appears as the aligned
// triple 3c+1, 3c+2, 3c+3 (1-based) -- the layout ZeroMatrix(DOF=3) produces.  One SUB-lane
// subgroup per node, lanes over the node's column triples.
template <int SUB>
__global__ void __launch_bounds__(256) k_block_check(int64_t nn, const int64_t *__restrict__ row_ptr,
                                                     const int32_t *__restrict__ col_idx, int32_t *__restrict__ bad) {
  const int64_t g = blockIdx.x * (int64_t)blockDim.x + threadIdx.x;
  const int64_t i = g / SUB;
  const int lane = (int)(g % SUB);
  if (i >= nn) return;
  const int64_t p0 = row_ptr[3 * i], p1 = row_ptr[3 * i + 1], p2 = row_ptr[3 * i + 2], p3 = row_ptr[3 * i + 3];
  const int64_t L = p1 - p0;
  if (L % 3 != 0 || p2 - p1 != L || p3 - p2 != L) {
    if (lane == 0) atomicOr(bad, 1);
    return;
  }
  // the three rows have L entries each: all nine loads of a triple are in bounds and in flight
  const int32_t *a = col_idx + (p0 - 1), *b = col_idx + (p1 - 1), *e = col_idx + (p2 - 1);
  bool ok = true;
  for (int64_t t = 3 * (int64_t)lane; t < L; t += 3 * SUB) {
    const int32_t a0 = a[t], a1 = a[t + 1], a2 = a[t + 2];
    const int32_t b0 = b[t], b1 = b[t + 1], b2 = b[t + 2];
    const int32_t e0 = e[t], e1 = e[t + 1], e2 = e[t + 2];
    ok = ok && (a0 - 1) % 3 == 0 && a1 == a0 + 1 && a2 == a0 + 2 && b0 == a0 && b1 == a0 + 1 && b2 == a0 + 2 &&
         e0 == a0 && e1 == a0 + 1 && e2 == a0 + 2;
  }
  if (!ok) atomicOr(bad, 1);
}

// first node and end node of slice s (plain: 64 consecutive nodes; mapped: a batch slice)
struct SliceNodes {
  int64_t nn;
  const int64_t *snode, *send;
  __device__ int64_t first(int64_t s) const { return snode ? snode[s] : s * 64; }
  __device__ int64_t end(int64_t s) const {
    if (snode) return send[s];
    return s * 64 + 64 < nn ? s * 64 + 64 : nn;
  }
};

__global__ void k_sell3_width(SliceNodes sn, int64_t nslices, const int64_t *__restrict__ row_ptr,
                              int32_t *__restrict__ width) {
  const int64_t s = blockIdx.x * (int64_t)blockDim.x + threadIdx.x;
  if (s >= nslices) return;
  int32_t w = 0;
  const int64_t e = sn.end(s);
  for (int64_t i = sn.first(s); i < e; ++i) {
    const int32_t dg = (int32_t)((row_ptr[3 * i + 1] - row_ptr[3 * i]) / 3);
    w = dg > w ? dg : w;
  }
  width[s] = w;
}

// one thread per cell (slot, lane): node column index, or -1 for padding
__global__ void k_sell3_cols(SliceNodes sn, int64_t nslices, int64_t ncells, const int64_t *__restrict__ soff,
                             const int64_t *__restrict__ row_ptr, const int32_t *__restrict__ col_idx,
                             int32_t *__restrict__ col) {
  for (int64_t g = blockIdx.x * (int64_t)blockDim.x + threadIdx.x; g < ncells;
       g += (int64_t)gridDim.x * blockDim.x) {
    const int64_t slot = g >> 6;
    const int lane = (int)(g & 63);
    int64_t lo = 0, hi = nslices;  // largest s with soff[s] <= slot
    while (hi - lo > 1) {
      const int64_t mid = (lo + hi) >> 1;
      if (soff[mid] <= slot) lo = mid; else hi = mid;
    }
    const int64_t i = sn.first(lo) + lane;
    const int64_t t = slot - soff[lo];
    int32_t c = -1;
    if (i < sn.end(lo)) {
      const int64_t rb = row_ptr[3 * i] - 1;
      if (3 * t < row_ptr[3 * i + 1] - 1 - rb) c = (col_idx[rb + 3 * t] - 1) / 3;
    }
    col[g] = c;
  }
}

template <typename V>
__global__ void k_sell3_vals(SliceNodes sn, int64_t nslices, int64_t ncells, const int64_t *__restrict__ soff,
                             const int64_t *__restrict__ row_ptr, const int32_t *__restrict__ col,
                             const double *__restrict__ val, const double *__restrict__ scale,
                             const double *__restrict__ cscale, V *__restrict__ out) {
  for (int64_t g = blockIdx.x * (int64_t)blockDim.x + threadIdx.x; g < ncells;
       g += (int64_t)gridDim.x * blockDim.x) {
    const int64_t slot = g >> 6;
    const int lane = (int)(g & 63);
    int64_t lo = 0, hi = nslices;
    while (hi - lo > 1) {
      const int64_t mid = (lo + hi) >> 1;
      if (soff[mid] <= slot) lo = mid; else hi = mid;
    }
    const int64_t i = sn.first(lo) + lane;
    const int64_t t = slot - soff[lo];
    V *o = out + slot * 64 * 9;
    bool live = false;
    if (i < sn.end(lo)) {
      const int64_t len = row_ptr[3 * i + 1] - row_ptr[3 * i];
      live = 3 * t < len;
    }
    for (int d = 0; d < 3; ++d)
      for (int k = 0; k < 3; ++k) {
        double a = live ? val[row_ptr[3 * i + d] - 1 + 3 * t + k] : 0.0;
        if (live && scale) a = (a * scale[3 * i + d]) * cscale[3 * (int64_t)col[g] + k];
        o[pfdev::sell3_value_index<V>(3 * d + k, lane)] = (V)a;
      }
  }
}

// Position of the diagonal entry of row i (the first one in storage order), or -1: a SUB-lane
// subgroup per row, each lane its strided share of the row, then the smallest position found.
template <int SUB>
__device__ inline int64_t diag_pos(int64_t i, bool live, const int64_t *__restrict__ row_ptr,
                                   const int32_t *__restrict__ col_idx, int lane) {
  int64_t kmin = INT64_MAX;
  if (live) {
    const int64_t e = row_ptr[i + 1] - 1;
    for (int64_t k = row_ptr[i] - 1 + lane; k < e; k += SUB)
      if (col_idx[k] == i + 1) {
        kmin = k;
        break;
      }
  }
#pragma unroll
  for (int o = SUB / 2; o > 0; o >>= 1) {
    const int64_t other = __shfl_xor(kmin, o, SUB);
    kmin = other < kmin ? other : kmin;
  }
  return kmin == INT64_MAX ? -1 : kmin;
}

// s_i = 1/sqrt|a_ii| (1 where the diagonal is missing or zero)
template <int SUB>
__global__ void __launch_bounds__(256) k_jacobi_scale(int64_t n, const int64_t *__restrict__ row_ptr,
                                                      const int32_t *__restrict__ col_idx,
                                                      const double *__restrict__ val, double *__restrict__ scale) {
  const int64_t g = blockIdx.x * (int64_t)blockDim.x + threadIdx.x;
  const int64_t i = g / SUB;
  const int lane = (int)(g % SUB);
  const int64_t k = diag_pos<SUB>(i, i < n, row_ptr, col_idx, lane);
  if (i >= n || lane != 0) return;
  const double d = k >= 0 ? val[k] : 0.0;
  scale[i] = d != 0.0 ? 1.0 / sqrt(fabs(d)) : 1.0;
}

// d_i = 1/a_ii (1 where the diagonal is missing or zero)
template <int SUB>
__global__ void __launch_bounds__(256) k_jacobi_row_scale(int64_t n, const int64_t *__restrict__ row_ptr,
                                                          const int32_t *__restrict__ col_idx,
                                                          const double *__restrict__ val, double *__restrict__ scale) {
  const int64_t g = blockIdx.x * (int64_t)blockDim.x + threadIdx.x;
  const int64_t i = g / SUB;
  const int lane = (int)(g % SUB);
  const int64_t k = diag_pos<SUB>(i, i < n, row_ptr, col_idx, lane);
  if (i >= n || lane != 0) return;
  const double d = k >= 0 ? val[k] : 0.0;
  scale[i] = d != 0.0 ? 1.0 / d : 1.0;
}

// FP32 copy of CRS values, optionally scaled s_i a_ij c_j (one thread per row)
__global__ void k_crs_f32(int64_t n, const int64_t *__restrict__ row_ptr, const int32_t *__restrict__ col_idx,
                          const double *__restrict__ val, const double *__restrict__ scale,
                          const double *__restrict__ cscale, float *__restrict__ out) {
  const int64_t i = blockIdx.x * (int64_t)blockDim.x + threadIdx.x;
  if (i >= n) return;
  for (int64_t k = row_ptr[i] - 1; k < row_ptr[i + 1] - 1; ++k) {
    double a = val[k];
    if (scale) a = (a * scale[i]) * cscale[col_idx[k] - 1];
    out[k] = (float)a;
  }
}

// ---- CRS product with x staged in LDS per row block (pfdev::staged_block) -----------
constexpr int ST_T = 256;      // threads of the staging and product kernels
constexpr int ST_HASH = 4096;  // LDS hash slots of the staging (twice the distinct-column limit)
constexpr int ST_PROBE = 64;   // probes after which a block is left unstaged

__device__ inline int st_hash(int32_t c) { return (int)(((uint32_t)c * 2654435761u) >> 20); }  // 12 bits

__device__ inline int block_excl_scan(int v, int *sh, int &total) {
  const int t = threadIdx.x;
  sh[t] = v;
  __syncthreads();
  for (int off = 1; off < ST_T; off <<= 1) {
    const int w = t >= off ? sh[t - off] : 0;
    __syncthreads();
    sh[t] += w;
    __syncthreads();
  }
  total = sh[ST_T - 1];
  const int incl = sh[t];
  __syncthreads();
  return incl - v;
}

// One workgroup per row block.  The block's 0-based columns go into an LDS hash set (linear
// probing); COUNT writes the number of distinct ones (ucnt[b]; -1 above ST_UMAX or after a
// probe run of ST_PROBE).  FILL sorts them (bitonic, <= ST_UMAX), writes them to
// ucol[ubase[b] ..] and every entry's rank among them to lidx; a block whose insertion now
// needs too long a probe run is marked unstaged instead (slot placement depends on the order
// of the atomics; the set of columns and hence the sorted list do not).
template <bool FILL>
__global__ void __launch_bounds__(ST_T) k_stage_build(int64_t n, int rpb, const int64_t *__restrict__ row_ptr,
                                                      const int32_t *__restrict__ col_idx, int32_t *__restrict__ ucnt,
                                                      const int64_t *__restrict__ ubase, int32_t *__restrict__ ucol,
                                                      uint16_t *__restrict__ lidx) {
  __shared__ int32_t key[ST_HASH];
  __shared__ int32_t rank[FILL ? ST_HASH : 1];
  __shared__ int32_t srt[FILL ? pfdev::ST_UMAX : 1];
  __shared__ int sh[ST_T];
  __shared__ int ovf;
  const int tid = threadIdx.x;
  const int64_t b = blockIdx.x;
  if (FILL && ucnt[b] < 0) return;  // whole block
  const int64_t r0 = b * rpb;
  const int64_t r1 = r0 + rpb < n ? r0 + rpb : n;
  const int64_t k0 = row_ptr[r0] - 1, k1 = row_ptr[r1] - 1;
  for (int i = tid; i < ST_HASH; i += ST_T) key[i] = -1;
  if (tid == 0) ovf = 0;
  __syncthreads();
  for (int64_t q = k0 + tid; q < k1; q += ST_T) {
    const int32_t c = col_idx[q] - 1;
    int h = st_hash(c), t = 0;
    for (; t < ST_PROBE; ++t) {
      const int32_t prev = atomicCAS(&key[h], -1, c);
      if (prev == -1 || prev == c) break;
      h = (h + 1) & (ST_HASH - 1);
    }
    if (t == ST_PROBE) ovf = 1;
  }
  __syncthreads();
  constexpr int CH = ST_HASH / ST_T;
  int cnt = 0;
  for (int i = 0; i < CH; ++i) cnt += key[tid * CH + i] >= 0;
  int total = 0;
  const int ex = block_excl_scan(cnt, sh, total);
  const bool staged = !ovf && total <= pfdev::ST_UMAX;
  if (!FILL || !staged) {  // COUNT, or FILL of a block that no longer fits
    if (tid == 0) ucnt[b] = staged ? total : -1;
    return;
  }
  int pos = ex;
  for (int i = 0; i < CH; ++i) {
    const int32_t c = key[tid * CH + i];
    if (c >= 0) srt[pos++] = c;
  }
  int P = 1;
  while (P < total) P <<= 1;
  for (int i = total + tid; i < P; i += ST_T) srt[i] = INT32_MAX;
  __syncthreads();
  for (int kk = 2; kk <= P; kk <<= 1)
    for (int j = kk >> 1; j > 0; j >>= 1) {
      for (int i = tid; i < P; i += ST_T) {
        const int l = i ^ j;
        if (l > i) {
          const int32_t x = srt[i], y = srt[l];
          if ((i & kk) == 0 ? x > y : x < y) {
            srt[i] = y;
            srt[l] = x;
          }
        }
      }
      __syncthreads();
    }
  int32_t *uc = ucol + ubase[b];
  for (int i = tid; i < total; i += ST_T) {
    const int32_t c = srt[i];
    uc[i] = c;
    int h = st_hash(c);
    while (key[h] != c) h = (h + 1) & (ST_HASH - 1);  // c is in the set
    rank[h] = i;
  }
  __syncthreads();
  for (int64_t q = k0 + tid; q < k1; q += ST_T) {
    const int32_t c = col_idx[q] - 1;
    int h = st_hash(c);
    while (key[h] != c) h = (h + 1) & (ST_HASH - 1);
    lidx[q] = (uint16_t)rank[h];
  }
}

__global__ void k_stage_sizes(int64_t nb, const int32_t *__restrict__ ucnt, int32_t *__restrict__ usz) {
  const int64_t b = blockIdx.x * (int64_t)blockDim.x + threadIdx.x;
  if (b < nb) usz[b] = ucnt[b] > 0 ? ucnt[b] : 0;
}

template <typename V, int SUB, bool ACC>
__global__ void __launch_bounds__(ST_T) k_spmv_crs_lds(int64_t n, pfdev::Stage st, const int64_t *__restrict__ row_ptr,
                                                       const int32_t *__restrict__ col_idx, const V *__restrict__ val,
                                                       const V *__restrict__ x, V *__restrict__ y) {
  __shared__ V sx[pfdev::ST_UMAX];
  auto op = [&](int64_t row, V s) { y[row] = ACC ? y[row] + s : s; };
  pfdev::staged_block<V, V, SUB, ST_T>(row_ptr, col_idx, val, st, n, (int64_t)blockIdx.x, sx, pfdev::PlainX<V>{x},
                                       op);
}

}  // namespace

template <typename V>
int pf_spmv_crs_launch(pf_ctx *ctx, int64_t n, const int64_t *row_ptr, const int32_t *col_idx,
                       const V *val, const V *x, V *y, double avg_row, bool accumulate) {
  if (n <= 0) return PF_OK;
  const int sub = pf_crs_sub(avg_row);
  const int64_t threads = n * sub;
  const unsigned blocks = pf_blocks(threads, 256, INT32_MAX);
#define PF_SPMV_CRS_CASE(S)                                                                  \
  case S:                                                                                   \
    if (accumulate)                                                                         \
      k_spmv_crs<V, S, true><<<blocks, 256, 0, ctx->stream>>>(n, row_ptr, col_idx, val, x, y); \
    else                                                                                    \
      k_spmv_crs<V, S, false><<<blocks, 256, 0, ctx->stream>>>(n, row_ptr, col_idx, val, x, y); \
    break;
  switch (sub) {
    PF_SPMV_CRS_CASE(4)
    PF_SPMV_CRS_CASE(8)
    PF_SPMV_CRS_CASE(16)
    PF_SPMV_CRS_CASE(32)
    default:
    PF_SPMV_CRS_CASE(64)
  }
#undef PF_SPMV_CRS_CASE
  PF_HIP(hipGetLastError());
  return PF_OK;
}
template int pf_spmv_crs_launch<double>(pf_ctx *, int64_t, const int64_t *, const int32_t *, const double *,
                                        const double *, double *, double, bool);
template int pf_spmv_crs_launch<float>(pf_ctx *, int64_t, const int64_t *, const int32_t *, const float *,
                                       const float *, float *, double, bool);

int pf_spmm_crs_launch(pf_ctx *ctx, int64_t n, int64_t cols, const int64_t *row_ptr, const int32_t *col_idx,
                       const double *val, const double *x, double *y, double avg_row) {
  if (n <= 0 || cols <= 0) return PF_OK;
  const int sub = pf_crs_sub(avg_row);
  const unsigned blocks = pf_blocks(n * sub, 256, INT32_MAX);
  switch (sub) {
    case 4: k_spmm_crs<double, 4><<<blocks, 256, 0, ctx->stream>>>(n, cols, row_ptr, col_idx, val, x, y); break;
    case 8: k_spmm_crs<double, 8><<<blocks, 256, 0, ctx->stream>>>(n, cols, row_ptr, col_idx, val, x, y); break;
    case 16: k_spmm_crs<double, 16><<<blocks, 256, 0, ctx->stream>>>(n, cols, row_ptr, col_idx, val, x, y); break;
    case 32: k_spmm_crs<double, 32><<<blocks, 256, 0, ctx->stream>>>(n, cols, row_ptr, col_idx, val, x, y); break;
    default: k_spmm_crs<double, 64><<<blocks, 256, 0, ctx->stream>>>(n, cols, row_ptr, col_idx, val, x, y); break;
  }
  PF_HIP(hipGetLastError());
  return PF_OK;
}

int pf_crs_sub(double avg_row) {
  // subgroup of ~avg_row/5 lanes (4..64): about five nonzeros per lane (measured at 100^3,
  // 79 nonzeros per row: 16 lanes 0.82 ms, 64 lanes 1.15 ms)
  int sub = 4;
  while (sub < 64 && sub * 1.4142 * 5.0 < avg_row) sub *= 2;
  return sub;
}

int pf_crs_stage_build(pf_ctx *ctx, int64_t n, const int64_t *row_ptr, const int32_t *col_idx, int64_t nnz,
                       CrsStage &st) {
  st.ready = false;
  st.n = n;
  st.nnz = nnz;
  st.sub = pf_crs_sub(n > 0 ? (double)nnz / (double)n : 1.0);
  // ST_RPS rows per subgroup, halved while the matrix has fewer than 8 blocks per CU: every
  // block starts with a dependent gather of its x entries, which other blocks must overlap
  int rps = pfdev::ST_RPS;
  st.rpb = rps * (ST_T / st.sub);
  while (rps > 1 && n > 0 && (n + st.rpb - 1) / st.rpb < 8 * (int64_t)ctx->num_cus) {
    rps /= 2;
    st.rpb = rps * (ST_T / st.sub);
  }
  st.nb = n > 0 ? (n + st.rpb - 1) / st.rpb : 0;
  st.staged = st.unique = 0;
  if (st.nb == 0) {
    st.ready = true;
    return PF_OK;
  }
  if (st.nb > INT32_MAX) {
    pf_set_error("staged CRS product: too many row blocks");
    return PF_INVALID;
  }
  PF_CHECK(st.ucnt.ensure((size_t)st.nb));
  PF_CHECK(st.usz.ensure((size_t)st.nb));
  PF_CHECK(st.ubase.ensure((size_t)st.nb + 1));
  PF_CHECK(st.lidx.ensure((size_t)nnz + 1));
  k_stage_build<false><<<(unsigned)st.nb, ST_T, 0, ctx->stream>>>(n, st.rpb, row_ptr, col_idx, st.ucnt.p, nullptr,
                                                                 nullptr, nullptr);
  k_stage_sizes<<<pf_blocks(st.nb, 256, INT32_MAX), 256, 0, ctx->stream>>>(st.nb, st.ucnt.p, st.usz.p);
  PF_HIP(hipGetLastError());
  PF_CHECK(pf_scan_i32_to_i64(ctx, st.usz.p, st.nb, st.ubase.p, st.tmp));
  int64_t total = 0;
  PF_CHECK(pf_read_i64(ctx, st.ubase.p + st.nb, &total));
  PF_CHECK(st.ucol.ensure((size_t)total + 1));
  k_stage_build<true><<<(unsigned)st.nb, ST_T, 0, ctx->stream>>>(n, st.rpb, row_ptr, col_idx, st.ucnt.p,
                                                                st.ubase.p, st.ucol.p, st.lidx.p);
  PF_HIP(hipGetLastError());
  std::vector<int32_t> h((size_t)st.nb);
  PF_HIP(hipMemcpyAsync(h.data(), st.ucnt.p, sizeof(int32_t) * st.nb, hipMemcpyDeviceToHost, ctx->stream));
  PF_HIP(hipStreamSynchronize(ctx->stream));
  for (int32_t c : h) st.staged += c >= 0;
  st.unique = total;
  st.ready = true;
  return PF_OK;
}

template <typename V>
int pf_spmv_crs_stage_launch(pf_ctx *ctx, const CrsStage &st, const int64_t *row_ptr, const int32_t *col_idx,
                             const V *val, const V *x, V *y, bool accumulate) {
  if (!st.ready) {
    pf_set_error("staged CRS product: staging not built");
    return PF_INVALID;
  }
  if (st.nb == 0) return PF_OK;
  const unsigned blocks = (unsigned)st.nb;
  const pfdev::Stage sv{st.lidx.p, st.ucnt.p, st.ubase.p, st.ucol.p, st.nb, st.rpb};
#define PF_SPMV_LDS_CASE(S)                                                                                   \
  case S:                                                                                                    \
    if (accumulate)                                                                                          \
      k_spmv_crs_lds<V, S, true><<<blocks, ST_T, 0, ctx->stream>>>(st.n, sv, row_ptr, col_idx, val, x, y);  \
    else                                                                                                     \
      k_spmv_crs_lds<V, S, false><<<blocks, ST_T, 0, ctx->stream>>>(st.n, sv, row_ptr, col_idx, val, x, y); \
    break;
  switch (st.sub) {
    PF_SPMV_LDS_CASE(4)
    PF_SPMV_LDS_CASE(8)
    PF_SPMV_LDS_CASE(16)
    PF_SPMV_LDS_CASE(32)
    default:
    PF_SPMV_LDS_CASE(64)
  }
#undef PF_SPMV_LDS_CASE
  PF_HIP(hipGetLastError());
  return PF_OK;
}
template int pf_spmv_crs_stage_launch<double>(pf_ctx *, const CrsStage &, const int64_t *, const int32_t *,
                                              const double *, const double *, double *, bool);
template int pf_spmv_crs_stage_launch<float>(pf_ctx *, const CrsStage &, const int64_t *, const int32_t *,
                                             const float *, const float *, float *, bool);

int pf_block_structure_check(pf_ctx *ctx, int64_t n, const int64_t *row_ptr, const int32_t *col_idx,
                             int64_t nn, int32_t *ok_dev, int *ok_host) {
  *ok_host = 0;
  if (n % 3 != 0 || nn * 3 != n) return PF_OK;
  if (nn == 0) {
    *ok_host = 1;
    return PF_OK;
  }
  PF_HIP(hipMemsetAsync(ok_dev, 0, sizeof(int32_t), ctx->stream));
  k_block_check<16><<<pf_blocks(nn * 16, 256, INT32_MAX), 256, 0, ctx->stream>>>(nn, row_ptr, col_idx, ok_dev);
  int32_t bad = 1;
  PF_HIP(hipMemcpyAsync(&bad, ok_dev, sizeof(int32_t), hipMemcpyDeviceToHost, ctx->stream));
  PF_HIP(hipStreamSynchronize(ctx->stream));
  *ok_host = bad ? 0 : 1;
  return PF_OK;
}

int pf_jacobi_scale(pf_ctx *ctx, int64_t n, const int64_t *row_ptr, const int32_t *col_idx, const double *val,
                    double *scale) {
  if (n <= 0) return PF_OK;
  k_jacobi_scale<16><<<pf_blocks(n * 16, 256, INT32_MAX), 256, 0, ctx->stream>>>(n, row_ptr, col_idx, val, scale);
  PF_HIP(hipGetLastError());
  return PF_OK;
}

int pf_jacobi_row_scale(pf_ctx *ctx, int64_t n, const int64_t *row_ptr, const int32_t *col_idx, const double *val,
                        double *scale) {
  if (n <= 0) return PF_OK;
  k_jacobi_row_scale<16><<<pf_blocks(n * 16, 256, INT32_MAX), 256, 0, ctx->stream>>>(n, row_ptr, col_idx, val,
                                                                                     scale);
  PF_HIP(hipGetLastError());
  return PF_OK;
}

int pf_crs_to_f32(pf_ctx *ctx, int64_t n, const int64_t *row_ptr, const int32_t *col_idx, const double *val,
                  const double *scale, float *out, const double *cscale) {
  if (n <= 0) return PF_OK;
  k_crs_f32<<<pf_blocks(n, 256, INT32_MAX), 256, 0, ctx->stream>>>(n, row_ptr, col_idx, val, scale,
                                                                   cscale ? cscale : scale, out);
  PF_HIP(hipGetLastError());
  return PF_OK;
}

template <typename V>
int pf_sell3_build(pf_ctx *ctx, int64_t n, const int64_t *row_ptr, const int32_t *col_idx,
                   const double *val, Sell3<V> &s, const double *scale, const SliceMap *map,
                   const double *cscale) {
  const int64_t nn = n / 3;
  s.nn = nn;
  s.nslices = map ? map->nslices : (nn + 63) / 64;
  s.snode = map ? map->snode : nullptr;
  s.send = map ? map->send : nullptr;
  const SliceNodes sn{nn, s.snode, s.send};
  DevBuf<int64_t> tmp;
  PF_CHECK(s.width.ensure((size_t)s.nslices + 1));
  PF_CHECK(s.soff.ensure((size_t)s.nslices + 1));
  if (s.nslices > 0)
    k_sell3_width<<<pf_blocks(s.nslices, 256), 256, 0, ctx->stream>>>(sn, s.nslices, row_ptr, s.width.p);
  PF_CHECK(pf_scan_i32_to_i64(ctx, s.width.p, s.nslices, s.soff.p, tmp));
  PF_CHECK(pf_read_i64(ctx, s.soff.p + s.nslices, &s.nslots));
  const int64_t ncells = s.nslots * 64;
  PF_CHECK(s.col.ensure((size_t)ncells + 1));
  PF_CHECK(s.val.ensure((size_t)ncells * 9 + 1));
  if (ncells > 0)
    k_sell3_cols<<<pf_blocks(ncells, 256, 65536), 256, 0, ctx->stream>>>(sn, s.nslices, ncells, s.soff.p,
                                                                       row_ptr, col_idx, s.col.p);
  PF_HIP(hipGetLastError());
  return pf_sell3_values<V>(ctx, n, row_ptr, val, s, scale, cscale);
}

template <typename V>
int pf_sell3_values(pf_ctx *ctx, int64_t n, const int64_t *row_ptr, const double *val, Sell3<V> &s,
                    const double *scale, const double *cscale) {
  (void)n;
  const int64_t ncells = s.nslots * 64;
  const SliceNodes sn{s.nn, s.snode, s.send};
  if (ncells > 0)
    k_sell3_vals<V><<<pf_blocks(ncells, 256, 65536), 256, 0, ctx->stream>>>(sn, s.nslices, ncells, s.soff.p,
                                                                          row_ptr, s.col.p, val, scale,
                                                                          cscale ? cscale : scale, s.val.p);
  PF_HIP(hipGetLastError());
  return PF_OK;
}

template <typename V>
int pf_spmv_sell3_launch(pf_ctx *ctx, const Sell3<V> &s, const V *x, V *y) {
  if (s.nslices == 0) return PF_OK;
  k_spmv_sell3<V><<<pf_blocks(s.nslices * 64, 256, INT32_MAX), 256, 0, ctx->stream>>>(
      s.nn, s.nslices, s.soff.p, s.width.p, s.col.p, s.val.p, x, y);
  PF_HIP(hipGetLastError());
  return PF_OK;
}

template int pf_sell3_build<double>(pf_ctx *, int64_t, const int64_t *, const int32_t *, const double *,
                                    Sell3<double> &, const double *, const SliceMap *, const double *);
template int pf_sell3_build<float>(pf_ctx *, int64_t, const int64_t *, const int32_t *, const double *,
                                   Sell3<float> &, const double *, const SliceMap *, const double *);
template int pf_sell3_values<double>(pf_ctx *, int64_t, const int64_t *, const double *, Sell3<double> &,
                                     const double *, const double *);
template int pf_sell3_values<float>(pf_ctx *, int64_t, const int64_t *, const double *, Sell3<float> &,
                                    const double *, const double *);
template int pf_spmv_sell3_launch<double>(pf_ctx *, const Sell3<double> &, const double *, double *);
template int pf_spmv_sell3_launch<float>(pf_ctx *, const Sell3<float> &, const float *, float *);
