// noalloc.hip -- BiCGSTAB on device arrays without any allocation, for callers that cannot
// allocate (code under stream capture, tight serving loops, harnesses that forbid hipMalloc).
//
// The solver's work vectors, partial sums and scalars live in `scratch`, one file-scope
// device array per GPU that stays resident until the process ends.  Its device address is
// looked up once per device with hipGetSymbolAddress when a context is created on that device
// (pf_noalloc_setup, called from pf_ctx_create with the device current); host code uses only
// that address, never the array's name.  Calls on one GPU share its array: calls that overlap
// on one GPU (several streams or threads at once) race.  Every value the kernels read from the
// array is written earlier in the same call, so nothing carries over between calls.
#include <cstdio>
#include <cstring>
#include <mutex>

#include "pf_internal.hpp"

namespace {

constexpr int64_t MAX_ROWS = PF_NOALLOC_MAX_ROWS;
// 7 vectors of MAX_ROWS doubles, the partial rows of up to 320 CUs and the scalar block
constexpr size_t SCRATCH_DOUBLES = 7 * (size_t)MAX_ROWS + 9 * 320 * 8 + 16 + 8 + 64;

__device__ __align__(256) double scratch[SCRATCH_DOUBLES];

// row_ptr nondecreasing from 1 to nnz+1 and every column in 1..n; *bad = 1 otherwise
__global__ void k_check_crs(int64_t n, const int64_t *__restrict__ row_ptr, const int32_t *__restrict__ col_idx,
                            int32_t *__restrict__ bad) {
  const int64_t i = blockIdx.x * (int64_t)blockDim.x + threadIdx.x;
  if (i >= n) return;
  const int64_t a = row_ptr[i], b = row_ptr[i + 1];
  bool ok = a >= 1 && b >= a;
  for (int64_t k = a - 1; ok && k < b - 1; ++k) ok = col_idx[k] >= 1 && (int64_t)col_idx[k] <= n;
  if (!ok) *bad = 1;
}

__global__ void k_clear_flag(int32_t *__restrict__ flag) { *flag = 0; }

constexpr int MAX_DEVICES = 64;
double *g_scratch[MAX_DEVICES] = {};
std::mutex g_mutex;

}  // namespace

int pf_noalloc_setup(int device) {
  if (device < 0 || device >= MAX_DEVICES) {
    pf_set_error("device %d outside 0..%d", device, MAX_DEVICES - 1);
    return PF_INVALID;
  }
  std::lock_guard<std::mutex> lock(g_mutex);
  if (g_scratch[device]) return PF_OK;
  void *ptr = nullptr;
  const hipError_t e = hipGetSymbolAddress(&ptr, HIP_SYMBOL(scratch));
  if (e != hipSuccess || !ptr) {
    pf_set_error("hipGetSymbolAddress(scratch) on device %d: %s", device, hipGetErrorString(e));
    return PF_ERR_HIP;
  }
  g_scratch[device] = static_cast<double *>(ptr);
  return PF_OK;
}

extern "C" int pf_dev_bicgstab_crs_noalloc(pf_ctx *ctx, int64_t n, const int64_t *row_ptr, const int32_t *col_idx,
                                           const double *val, const double *rhs, double *x,
                                           const pf_solve_opts *opts, pf_stats *stats) {
  if (stats) std::memset(stats, 0, sizeof(*stats));
  if (!ctx) {
    pf_set_error("null pf_ctx");
    return PF_INVALID;
  }
  if (n < 0 || n > MAX_ROWS) {
    pf_set_error("pf_dev_bicgstab_crs_noalloc: n = %lld outside 0..%lld (the size of its scratch array)",
                 (long long)n, (long long)MAX_ROWS);
    return PF_INVALID;
  }
  if (n > 0 && (!row_ptr || !col_idx || !val || !rhs || !x)) {
    pf_set_error("pf_dev_bicgstab_crs_noalloc: null array");
    return PF_INVALID;
  }
  pf_solve_opts o;
  if (opts) o = *opts; else pf_solve_opts_default(&o);
  if (o.precision != 64 || o.preconditioning != 0 || (o.spmv_kind != 0 && o.spmv_kind != 1)) {
    pf_set_error("pf_dev_bicgstab_crs_noalloc: FP64, no preconditioning, CRS layout only");
    return PF_INVALID;
  }
  o.kernel_timing = 0;
  if (n == 0) return PF_OK;
  int prev = -1;
  if (hipGetDevice(&prev) != hipSuccess) prev = -1;
  if (prev != ctx->device && hipSetDevice(ctx->device) != hipSuccess) {
    pf_set_error("hipSetDevice(%d) failed", ctx->device);
    return PF_ERR_HIP;
  }
  int rc = PF_OK;
  double *ws = (ctx->device >= 0 && ctx->device < MAX_DEVICES) ? g_scratch[ctx->device] : nullptr;
  SolveResult res;
  int64_t ends[2] = {0, 0};
  if (!ws) {
    pf_set_error("pf_dev_bicgstab_crs_noalloc: scratch of device %d not set up", ctx->device);
    rc = PF_ERR_HIP;
  } else if (pf_bicgstab_workspace_doubles(ctx, n) > SCRATCH_DOUBLES - 8) {
    pf_set_error("pf_dev_bicgstab_crs_noalloc: %d CUs need more partial sums than the scratch array holds",
                 ctx->num_cus);
    rc = PF_INVALID;
  } else if (hipMemcpyAsync(ends, row_ptr, sizeof(int64_t), hipMemcpyDeviceToHost, ctx->stream) != hipSuccess ||
             hipMemcpyAsync(ends + 1, row_ptr + n, sizeof(int64_t), hipMemcpyDeviceToHost, ctx->stream) !=
                 hipSuccess ||
             hipStreamSynchronize(ctx->stream) != hipSuccess) {
    pf_set_error("pf_dev_bicgstab_crs_noalloc: reading row_ptr failed");
    rc = PF_ERR_HIP;
  } else if (ends[0] != 1 || ends[1] < 1) {
    pf_set_error("pf_dev_bicgstab_crs_noalloc: row_ptr must start at 1 and end at nnz+1");
    rc = PF_INVALID;
  } else {
    // validate the pattern on the device before any kernel gathers through it (flag in the
    // scratch array's last slot, beyond the solver's workspace)
    int32_t *flag = reinterpret_cast<int32_t *>(ws + SCRATCH_DOUBLES - 8);
    int32_t bad = 0;
    k_clear_flag<<<1, 1, 0, ctx->stream>>>(flag);
    k_check_crs<<<pf_blocks(n, 256, INT32_MAX), 256, 0, ctx->stream>>>(n, row_ptr, col_idx, flag);
    if (hipGetLastError() != hipSuccess ||
        hipMemcpyAsync(&bad, flag, sizeof(int32_t), hipMemcpyDeviceToHost, ctx->stream) != hipSuccess ||
        hipStreamSynchronize(ctx->stream) != hipSuccess) {
      pf_set_error("pf_dev_bicgstab_crs_noalloc: CRS check failed to run");
      rc = PF_ERR_HIP;
    } else if (bad) {
      pf_set_error("pf_dev_bicgstab_crs_noalloc: row_ptr decreases or a column lies outside 1..n");
      rc = PF_INVALID;
    }
  }
  if (rc == PF_OK && ws) {
    rc = pf_bicgstab_crs_workspace(ctx, n, row_ptr, col_idx, val, ends[1] - 1, rhs, x, o, ws,
                                   SCRATCH_DOUBLES - 8, res);
  }
  if (prev >= 0 && prev != ctx->device) (void)hipSetDevice(prev);
  if (rc != PF_OK) return rc;
  if (stats) {
    stats->status = res.status;
    stats->iterations = res.iterations;
    stats->spmv_kind_used = res.kind;
    stats->rr = res.rr;
    stats->rr1 = res.rr1;
    stats->true_relres = res.true_relres;
    stats->solve_ms = res.solve_ms;
  }
  return res.status;
}
