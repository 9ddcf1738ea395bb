// bc.hip -- Dirichlet conditions of FEMSolver_ on the device CRS.
//   fix()   : flags and prescribed values (FEMSolverClass.f90:1073-1136); with repeated ids
//             the last one wins, as in the reference's sequential loop.
//   solve() : steps 1-4 of the elimination (FEMSolverClass.f90:1707-1756) and the final
//             x_fixed = ubar overwrite (:1814-1823).
#include "pf_internal.hpp"

#pragma clang fp contract(off)

namespace {

__global__ void k_fix_last(int64_t nids, const int32_t *__restrict__ ids, int64_t n,
                           int32_t *__restrict__ last) {
  const int64_t t = blockIdx.x * (int64_t)blockDim.x + threadIdx.x;
  if (t >= nids) return;
  const int32_t id = ids[t];
  if (id < 1 || (int64_t)id > n) return;
  atomicMax(&last[id - 1], (int32_t)t);
}

__global__ void k_fix_write(int64_t nids, const int32_t *__restrict__ ids, int64_t n,
                            const int32_t *__restrict__ last, const double *__restrict__ values,
                            double fixvalue, uint8_t *__restrict__ fixed, double *__restrict__ ubar) {
  const int64_t t = blockIdx.x * (int64_t)blockDim.x + threadIdx.x;
  if (t >= nids) return;
  const int32_t id = ids[t];
  if (id < 1 || (int64_t)id > n) return;
  if (last[id - 1] != (int32_t)t) return;
  fixed[id - 1] = 1;
  ubar[id - 1] = values ? values[t] : fixvalue;
}

// one SUB-lane subgroup per row, coalesced over the row's columns:
// (1) b_i -= A_ij*ubar_j over the fixed columns in CRS order, (2) b_i = ubar_i on fixed rows,
// (3) zero fixed rows and columns, (4) diagonal 1 on fixed rows.
// Step (1) is a sequential sum in the reference's order, so it runs on the subgroup's first
// lane, and only for the rows that hold a fixed column (a layer of rows next to each fixed
// set); every other row keeps b.  A fixed row's b is replaced by ubar whatever (1) gave.
template <int SUB>
__global__ void __launch_bounds__(256) k_bc_rows(int64_t n, const int64_t *__restrict__ row_ptr,
                                                 const int32_t *__restrict__ col_idx, double *__restrict__ val,
                                                 double *__restrict__ rhs, const uint8_t *__restrict__ fixed,
                                                 const double *__restrict__ ubar) {
  const int64_t g = blockIdx.x * (int64_t)blockDim.x + threadIdx.x;
  const int64_t i = g / SUB;
  const int lane = (int)(g % SUB);
  if (i >= n) return;  // whole subgroups leave together (SUB divides 64)
  const int64_t k0 = row_ptr[i] - 1, k1 = row_ptr[i + 1] - 1;
  const bool fr = fixed[i] != 0;
  // four column loads, then four flag gathers, in flight per lane
  int anyf = 0;
  int64_t k = k0 + lane;
  for (; k + 3 * SUB < k1; k += 4 * SUB) {
    const int32_t c0 = col_idx[k], c1 = col_idx[k + SUB], c2 = col_idx[k + 2 * SUB], c3 = col_idx[k + 3 * SUB];
    anyf |= (fixed[c0 - 1] | fixed[c1 - 1]) | (fixed[c2 - 1] | fixed[c3 - 1]);
  }
  for (; k < k1; k += SUB) anyf |= fixed[col_idx[k] - 1];
#pragma unroll
  for (int off = SUB / 2; off > 0; off >>= 1) anyf |= __shfl_xor(anyf, off, SUB);
  if (fr) {
    if (lane == 0) rhs[i] = ubar[i];
  } else if (anyf) {
    // b = b - a_ik ubar_k over the fixed columns in CRS order: the lanes load SUB terms at a
    // time, then every lane subtracts them in entry order from shuffles (the same roundings as
    // one sequential loop)
    double b = rhs[i];
    for (int64_t base = k0; base < k1; base += SUB) {
      const int64_t k = base + lane;
      double term = 0.0;
      int has = 0;
      if (k < k1) {
        const int64_t c = col_idx[k] - 1;
        if (fixed[c]) {
          term = val[k] * ubar[c];
          has = 1;
        }
      }
      const int cnt = k1 - base < SUB ? (int)(k1 - base) : SUB;
      for (int t = 0; t < cnt; ++t) {
        const double tt = __shfl(term, t, SUB);
        if (__shfl(has, t, SUB)) b = b - tt;
      }
    }
    if (lane == 0) rhs[i] = b;
  }
  // the reads of val above have been consumed (shuffled) before these stores
  if (fr || anyf)
    for (int64_t k = k0 + lane; k < k1; k += SUB) {
      const int64_t c = col_idx[k] - 1;
      if (fr || fixed[c]) val[k] = (fr && c == i) ? 1.0 : 0.0;
    }
}

__global__ void k_set_fixed_x(int64_t n, const uint8_t *__restrict__ fixed,
                              const double *__restrict__ ubar, double *__restrict__ x) {
  const int64_t i = blockIdx.x * (int64_t)blockDim.x + threadIdx.x;
  if (i < n && fixed[i]) x[i] = ubar[i];
}

}  // namespace

int pf_bc_fix(pf_ctx *ctx, int64_t n, int64_t nids, const int32_t *ids, const double *values,
              double fixvalue, uint8_t *fixed, double *ubar) {
  if (nids <= 0) return PF_OK;
  if (nids >= (int64_t)INT32_MAX) {
    pf_set_error("too many fixed ids");
    return PF_INVALID;
  }
  DevBuf<int32_t> last;
  PF_CHECK(last.ensure((size_t)n + 1));
  PF_HIP(hipMemsetAsync(last.p, 0xff, sizeof(int32_t) * (n + 1), ctx->stream));  // -1
  k_fix_last<<<pf_blocks(nids, 256), 256, 0, ctx->stream>>>(nids, ids, n, last.p);
  k_fix_write<<<pf_blocks(nids, 256), 256, 0, ctx->stream>>>(nids, ids, n, last.p, values, fixvalue,
                                                             fixed, ubar);
  PF_HIP(hipGetLastError());
  PF_HIP(hipStreamSynchronize(ctx->stream));
  return PF_OK;
}

int pf_bc_apply(pf_ctx *ctx, int64_t n, const int64_t *row_ptr, const int32_t *col_idx, double *val,
                double *rhs, const uint8_t *fixed, const double *ubar, double avg_row) {
  if (n <= 0) return PF_OK;
  // about five entries per lane, as the CRS row product
  int sub = 4;
  while (sub < 64 && sub * 1.4142 * 5.0 < avg_row) sub *= 2;
  const unsigned blocks = pf_blocks(n * sub, 256, INT32_MAX);
  switch (sub) {
    case 4: k_bc_rows<4><<<blocks, 256, 0, ctx->stream>>>(n, row_ptr, col_idx, val, rhs, fixed, ubar); break;
    case 8: k_bc_rows<8><<<blocks, 256, 0, ctx->stream>>>(n, row_ptr, col_idx, val, rhs, fixed, ubar); break;
    case 16: k_bc_rows<16><<<blocks, 256, 0, ctx->stream>>>(n, row_ptr, col_idx, val, rhs, fixed, ubar); break;
    case 32: k_bc_rows<32><<<blocks, 256, 0, ctx->stream>>>(n, row_ptr, col_idx, val, rhs, fixed, ubar); break;
    default: k_bc_rows<64><<<blocks, 256, 0, ctx->stream>>>(n, row_ptr, col_idx, val, rhs, fixed, ubar); break;
  }
  PF_HIP(hipGetLastError());
  return PF_OK;
}

int pf_bc_set_fixed_x(pf_ctx *ctx, int64_t n, const uint8_t *fixed, const double *ubar, double *x) {
  if (n <= 0) return PF_OK;
  k_set_fixed_x<<<pf_blocks(n, 256), 256, 0, ctx->stream>>>(n, fixed, ubar, x);
  PF_HIP(hipGetLastError());
  return PF_OK;
}
