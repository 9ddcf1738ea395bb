// spmv_dev.hpp -- device-side building blocks of the sparse matrix-vector products.
//
// Two layouts of the same matrix are used:
//  * CRS as the reference stores it (1-based int64 row_ptr, int32 col_idx, values):
//    a subgroup of SUB lanes reduces one row (sub_crs_matvec_generic,
//    LinearSolverClass.f90:3728-3776, without its per-nonzero atomics).
//  * Sell3: the DOF=3 node blocks that ZeroMatrix produces (FEMDomainClass.f90:10780-10836
//    gives the 3 rows of a node one shared node-column list) stored sliced-ELL: one lane of
//    a wavefront per node, 64 nodes per slice, one int32 node index per 3x3 block and the
//    nine values of a block component-major across the 64 lanes, so every load of the
//    inner loop is a contiguous 64-lane access and no cross-lane reduction is needed.
// Blocked matrix values and indices are read once per product and never reused, so they
// are loaded non-temporally; this keeps the gathered vector entries resident in L2/MALL
// (measured on MI355X: 5.6 -> 6.1 TB/s for the FP64 blocked product at 100^3).
// Padding slots hold zero values and index -1; they are read like live slots (no branch
// before the value loads) and multiply x[0].
#pragma once
#include "pf_internal.hpp"

namespace pfdev {

template <typename T>
__device__ inline T ldnt(const T *p) {
  return __builtin_nontemporal_load(p);
}

template <typename V, int SUB>
__device__ inline V subgroup_sum(V v) {
#pragma unroll
  for (int off = SUB / 2; off > 0; off >>= 1) v += __shfl_xor(v, off, SUB);
  return v;
}

// sum_k val[k] * xf(col[k]-1) over row `row`, lanes of a SUB-wide subgroup split the row
template <typename V, typename A, int SUB, typename XF>
__device__ inline A crs_row(const int64_t *__restrict__ row_ptr, const int32_t *__restrict__ col_idx,
                            const V *__restrict__ val, int64_t row, int sub, XF xf) {
  const int64_t k0 = row_ptr[row] - 1, k1 = row_ptr[row + 1] - 1;
  A acc = 0;
  // plain loads: neighbouring subgroups share cache lines of val/col_idx (rows are not
  // line-aligned), which non-temporal loads would fetch twice (measured 0.82 vs 1.01 ms at 100^3)
  // four loads in flight per lane (four accumulators): more gathers of x outstanding
  A acc1 = 0, acc2 = 0, acc3 = 0;
  int64_t k = k0 + sub;
  for (; k + 3 * SUB < k1; k += 4 * SUB) {
    acc += (A)val[k] * (A)xf((int64_t)col_idx[k] - 1);
    acc1 += (A)val[k + SUB] * (A)xf((int64_t)col_idx[k + SUB] - 1);
    acc2 += (A)val[k + 2 * SUB] * (A)xf((int64_t)col_idx[k + 2 * SUB] - 1);
    acc3 += (A)val[k + 3 * SUB] * (A)xf((int64_t)col_idx[k + 3 * SUB] - 1);
  }
  for (; k < k1; k += SUB) acc += (A)val[k] * (A)xf((int64_t)col_idx[k] - 1);
  return subgroup_sum<A, SUB>((acc + acc1) + (acc2 + acc3));
}

// the three row sums of node (slice s, lane l); xf(j) returns x_j
template <typename V, typename A, typename XF>
__device__ inline void sell3_node(const int64_t *__restrict__ soff, const int32_t *__restrict__ width,
                                  const int32_t *__restrict__ col, const V *__restrict__ val, int64_t s,
                                  int lane, XF xf, A &y0, A &y1, A &y2) {
  const int64_t so = soff[s];
  const int w = width[s];
  A a0 = 0, a1 = 0, a2 = 0;
  const int32_t *cp = col + so * 64 + lane;
  const V *vp = val + so * 64 * 9 + lane;
#pragma unroll 2
  for (int t = 0; t < w; ++t) {
    const int32_t c = ldnt(cp + (int64_t)t * 64);
    const V *v = vp + (int64_t)t * 64 * 9;
    V q[9];
#pragma unroll
    for (int k = 0; k < 9; ++k) q[k] = ldnt(v + 64 * k);
    const int64_t j = 3 * (int64_t)(c < 0 ? 0 : c);
    const A x0 = (A)xf(j), x1 = (A)xf(j + 1), x2 = (A)xf(j + 2);
    a0 += (A)q[0] * x0 + (A)q[1] * x1 + (A)q[2] * x2;
    a1 += (A)q[3] * x0 + (A)q[4] * x1 + (A)q[5] * x2;
    a2 += (A)q[6] * x0 + (A)q[7] * x1 + (A)q[8] * x2;
  }
  y0 = a0;
  y1 = a1;
  y2 = a2;
}

template <typename V>
struct PlainX {
  const V *__restrict__ x;
  __device__ V operator()(int64_t j) const { return x[j]; }
};

}  // namespace pfdev
