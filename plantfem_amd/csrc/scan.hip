// scan.hip -- exclusive prefix sums used by the graph build (counts -> int64 offsets).
// Three passes: per-tile sums, one-block scan of the tile sums, per-tile apply.
#include "pf_internal.hpp"

namespace {
constexpr int SCAN_T = 256;
constexpr int SCAN_ITEMS = 8;
constexpr int SCAN_TILE = SCAN_T * SCAN_ITEMS;

__device__ inline int64_t block_inclusive_scan(int64_t v, int64_t *sh) {
  const int t = threadIdx.x;
  sh[t] = v;
  __syncthreads();
  for (int off = 1; off < SCAN_T; off <<= 1) {
    int64_t add = (t >= off) ? sh[t - off] : 0;
    __syncthreads();
    sh[t] += add;
    __syncthreads();
  }
  return sh[t];
}

template <typename Tin>
__global__ void __launch_bounds__(SCAN_T) k_tile_sums(const Tin *__restrict__ in, int64_t n,
                                                      int64_t *__restrict__ bsum) {
  __shared__ int64_t sh[SCAN_T];
  const int64_t base = (int64_t)blockIdx.x * SCAN_TILE;
  int64_t s = 0;
  for (int k = 0; k < SCAN_ITEMS; ++k) {
    int64_t i = base + (int64_t)k * SCAN_T + threadIdx.x;
    if (i < n) s += (int64_t)in[i];
  }
  sh[threadIdx.x] = s;
  __syncthreads();
  for (int off = SCAN_T / 2; off > 0; off >>= 1) {
    if ((int)threadIdx.x < off) sh[threadIdx.x] += sh[threadIdx.x + off];
    __syncthreads();
  }
  if (threadIdx.x == 0) bsum[blockIdx.x] = sh[0];
}

// one block of SCAN_T threads: exclusive scan of bsum[0..nb), total into bsum[nb]
__global__ void __launch_bounds__(SCAN_T) k_scan_sums(int64_t *__restrict__ bsum, int64_t nb) {
  __shared__ int64_t sh[SCAN_T];
  int64_t carry = 0;
  for (int64_t base = 0; base < nb; base += SCAN_T) {
    int64_t i = base + threadIdx.x;
    int64_t v = (i < nb) ? bsum[i] : 0;
    int64_t incl = block_inclusive_scan(v, sh);
    if (i < nb) bsum[i] = carry + incl - v;
    carry += sh[SCAN_T - 1];
    __syncthreads();
  }
  if (threadIdx.x == 0) bsum[nb] = carry;
}

template <typename Tin>
__global__ void __launch_bounds__(SCAN_T) k_scan_apply(const Tin *__restrict__ in, int64_t n,
                                                       const int64_t *__restrict__ bsum,
                                                       int64_t nb, int64_t *__restrict__ out) {
  __shared__ int64_t sh[SCAN_T];
  int64_t carry = bsum[blockIdx.x];
  const int64_t base = (int64_t)blockIdx.x * SCAN_TILE;
  for (int k = 0; k < SCAN_ITEMS; ++k) {
    int64_t i = base + (int64_t)k * SCAN_T + threadIdx.x;
    int64_t v = (i < n) ? (int64_t)in[i] : 0;
    int64_t incl = block_inclusive_scan(v, sh);
    if (i < n) out[i] = carry + incl - v;
    carry += sh[SCAN_T - 1];
    __syncthreads();
  }
  if (blockIdx.x == 0 && threadIdx.x == 0) out[n] = bsum[nb];
}

template <typename Tin>
int scan_impl(pf_ctx *ctx, const Tin *in, int64_t n, int64_t *out, DevBuf<int64_t> &tmp) {
  const int64_t nb = (n + SCAN_TILE - 1) / SCAN_TILE;
  PF_CHECK(tmp.ensure((size_t)nb + 1));
  if (nb == 0) {
    PF_HIP(hipMemsetAsync(out, 0, sizeof(int64_t), ctx->stream));
    return PF_OK;
  }
  k_tile_sums<Tin><<<(unsigned)nb, SCAN_T, 0, ctx->stream>>>(in, n, tmp.p);
  k_scan_sums<<<1, SCAN_T, 0, ctx->stream>>>(tmp.p, nb);
  k_scan_apply<Tin><<<(unsigned)nb, SCAN_T, 0, ctx->stream>>>(in, n, tmp.p, nb, out);
  PF_HIP(hipGetLastError());
  return PF_OK;
}
}  // namespace

int pf_scan_i32_to_i64(pf_ctx *ctx, const int32_t *in, int64_t n, int64_t *out, DevBuf<int64_t> &tmp) {
  return scan_impl<int32_t>(ctx, in, n, out, tmp);
}

int pf_scan_i64(pf_ctx *ctx, const int64_t *in, int64_t n, int64_t *out, DevBuf<int64_t> &tmp) {
  return scan_impl<int64_t>(ctx, in, n, out, tmp);
}

int pf_read_i64(pf_ctx *ctx, const int64_t *dev, int64_t *host) {
  PF_HIP(hipMemcpyAsync(host, dev, sizeof(int64_t), hipMemcpyDeviceToHost, ctx->stream));
  PF_HIP(hipStreamSynchronize(ctx->stream));
  return PF_OK;
}
