// spmv_dev.hpp -- device-side building blocks of the sparse matrix-vector products.
//
// Two layouts of the same matrix are used:
//  * CRS as the reference stores it (1-based int64 row_ptr, int32 col_idx, values):
//    a subgroup of SUB lanes reduces one row (sub_crs_matvec_generic,
//    LinearSolverClass.f90:3728-3776, without its per-nonzero atomics).
//  * Sell3: the DOF=3 node blocks that ZeroMatrix produces (FEMDomainClass.f90:10780-10836
//    gives the 3 rows of a node one shared node-column list) stored sliced-ELL: one lane of
//    a wavefront per node, 64 nodes per slice, one int32 node index per 3x3 block and the
//    nine values of a block in 16-byte groups per lane (sell3_value_index), so every load of
//    the inner loop is a contiguous 64-lane access of 16 bytes per lane (FP64: four pairs and
//    one single, FP32: two quads and one single) and no cross-lane reduction is needed.
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

// sum_k val[k] * xf(ix(k)) over row `row`, lanes of a SUB-wide subgroup split the row; ix(k)
// is the index of entry k's column in whatever holds x (the 0-based column for x itself, a
// position in LDS for a staged copy), so every form sums the same products in the same order
template <typename V, typename A, int SUB, typename IX, typename XF>
__device__ inline A crs_row_ix(const int64_t *__restrict__ row_ptr, IX ix, const V *__restrict__ val,
                               int64_t row, int sub, XF xf) {
  const int64_t k0 = row_ptr[row] - 1, k1 = row_ptr[row + 1] - 1;
  A acc = 0;
  // plain loads: neighbouring subgroups share cache lines of val/col_idx (rows are not
  // line-aligned), which non-temporal loads would fetch twice (measured 0.82 vs 1.01 ms at 100^3)
  // four loads in flight per lane (four accumulators): more gathers of x outstanding
  A acc1 = 0, acc2 = 0, acc3 = 0;
  int64_t k = k0 + sub;
  for (; k + 3 * SUB < k1; k += 4 * SUB) {
    acc += (A)val[k] * (A)xf(ix(k));
    acc1 += (A)val[k + SUB] * (A)xf(ix(k + SUB));
    acc2 += (A)val[k + 2 * SUB] * (A)xf(ix(k + 2 * SUB));
    acc3 += (A)val[k + 3 * SUB] * (A)xf(ix(k + 3 * SUB));
  }
  for (; k < k1; k += SUB) acc += (A)val[k] * (A)xf(ix(k));
  return subgroup_sum<A, SUB>((acc + acc1) + (acc2 + acc3));
}

// 0-based column of entry k of a 1-based CRS
struct ColIx {
  const int32_t *__restrict__ col_idx;
  __device__ int64_t operator()(int64_t k) const { return (int64_t)col_idx[k] - 1; }
};

// position of entry k's column in its row block's staged copy of x
struct LocalIx {
  const uint16_t *__restrict__ lidx;
  __device__ int operator()(int64_t k) const { return lidx[k]; }
};

template <typename V, typename A, int SUB, typename XF>
__device__ inline A crs_row(const int64_t *__restrict__ row_ptr, const int32_t *__restrict__ col_idx,
                            const V *__restrict__ val, int64_t row, int sub, XF xf) {
  return crs_row_ix<V, A, SUB>(row_ptr, ColIx{col_idx}, val, row, sub, xf);
}

// the three row sums of node (slice s, lane l); xf(j) returns x_j
// Position of value q (= 3*row_dof + col_dof) of lane l inside a slot of 576 values: the
// first 8 in groups of VW = 16/sizeof(V) per lane (group g of all 64 lanes contiguous), the
// ninth after them.  Measured at 100^3 against nine 8-byte loads per slot (one 64-lane plane
// per value): y = A p 0.3835 -> 0.3688 ms, v = A e 0.397 -> 0.385 ms.
template <typename V>
__host__ __device__ constexpr int sell3_value_index(int q, int l) {
  return q < 8 ? (q / (16 / (int)sizeof(V))) * 64 * (16 / (int)sizeof(V)) + l * (16 / (int)sizeof(V)) +
                     q % (16 / (int)sizeof(V))
               : 512 + l;
}

template <typename V, typename A, typename XF>
__device__ inline void sell3_node(const int64_t *__restrict__ soff, const int32_t *__restrict__ width,
                                  const int32_t *__restrict__ col, const V *__restrict__ val, int64_t s,
                                  int lane, XF xf, A &y0, A &y1, A &y2) {
  const int64_t so = soff[s];
  const int w = width[s];
  A a0 = 0, a1 = 0, a2 = 0;
  constexpr int VW = 16 / sizeof(V);  // values per 16-byte group
  typedef V VV __attribute__((ext_vector_type(VW)));
  const int32_t *cp = col + so * 64 + lane;
  const V *vs = val + so * 64 * 9;  // slot 0 of the slice (576 values per slot, 16-byte aligned)
#pragma unroll 2
  for (int t = 0; t < w; ++t) {
    const int32_t c = ldnt(cp + (int64_t)t * 64);
    const V *v = vs + (int64_t)t * 64 * 9;
    V q[9];
    const VV *g = reinterpret_cast<const VV *>(v) + lane;
#pragma unroll
    for (int p = 0; p < 8 / VW; ++p) {
      const VV pv = ldnt(g + 64 * p);
#pragma unroll
      for (int h = 0; h < VW; ++h) q[p * VW + h] = pv[h];
    }
    q[8] = ldnt(v + 512 + lane);
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

// x staged in LDS (index = position in the row block's list of distinct columns)
template <typename V>
struct LdsX {
  const V *x;
  __device__ V operator()(int j) const { return x[j]; }
};

// ---- CRS rows with x staged in LDS per row block ------------------------------------
// A row block is rpb consecutive rows.  Its entries reference few distinct columns (hex8
// rows: ~12% of the entries), so the block gathers each distinct x entry once into LDS and
// its rows then read only streams (values, 16-bit positions) and LDS instead of one
// dependent col_idx -> x gather per entry.  ucnt[b] < 0: block b was not staged (above the
// limits) and gathers through col_idx.
constexpr int ST_UMAX = 2048;  // distinct columns of a staged block (16 KB of FP64 LDS)
constexpr int ST_RPS = 4;      // rows per subgroup in a row block (fewer on small matrices)

struct Stage {
  const uint16_t *lidx;  // per entry: position of its column in the block's list
  const int32_t *ucnt;   // per block: distinct columns, -1 = not staged
  const int64_t *ubase;  // per block: offset of its list in ucol
  const int32_t *ucol;   // the lists: sorted 0-based columns
  int64_t nb;
  int rpb;
};

// Rows of block b: op(row, sum) by one lane per row; sums bitwise those of crs_row (xf's values
// are staged unchanged).  Every thread of the block calls this (it holds two block barriers).
template <typename V, typename A, int SUB, int NT, typename XF, typename OP>
__device__ inline void staged_block(const int64_t *__restrict__ row_ptr, const int32_t *__restrict__ col_idx,
                                    const V *__restrict__ val, const Stage &st, int64_t n, int64_t b, A *sx, XF xf,
                                    OP op) {
  const int U = st.ucnt[b];
  if (U > 0) {
    const int32_t *uc = st.ucol + st.ubase[b];
    for (int i = threadIdx.x; i < U; i += NT) sx[i] = (A)xf((int64_t)uc[i]);
  }
  __syncthreads();
  const int sg = threadIdx.x / SUB, sub = threadIdx.x % SUB;
  const int64_t r0 = b * st.rpb;
  const int64_t r1 = r0 + st.rpb < n ? r0 + st.rpb : n;
  for (int64_t row = r0 + sg; row < r1; row += NT / SUB) {  // whole subgroups
    const A s = U >= 0 ? crs_row_ix<V, A, SUB>(row_ptr, LocalIx{st.lidx}, val, row, sub, LdsX<A>{sx})
                       : crs_row<V, A, SUB>(row_ptr, col_idx, val, row, sub, xf);
    if (sub == 0) op(row, s);
  }
  __syncthreads();  // the next block rewrites sx
}

}  // namespace pfdev
