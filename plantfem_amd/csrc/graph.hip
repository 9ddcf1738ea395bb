// graph.hip -- CRS sparsity of a finite-element mesh in plantFEM's first-appearance
// column order (ZeroMatrix_as_CRS_FEMDomain, FEMDomainClass.f90:10665-10838).
//
// The reference builds, per node, the concatenation of the local node lists of its
// incident elements taken in ascending element id (:10680-10699), drops later repeats
// (:10702-10734) and expands every node entry to a DOF x DOF block (:10780-10836).
// Here the same order is produced in parallel:
//   1. count incident elements per node (atomics), scan to offsets;
//   2. scatter element ids into per-node lists (atomics -> unordered), then insertion-sort
//      each short list so it is ascending again;
//   3. one thread per node walks its candidate sequence twice: first to count the
//      distinct nodes, then (after a scan) to write them in first-appearance order;
//   4. DOF expansion writes row_ptr per node and col_idx per (node, neighbour) block.
#include "pf_internal.hpp"

namespace {

__global__ void k_check_nodes(const int32_t *__restrict__ en, int64_t total, int64_t nn,
                              int32_t *__restrict__ bad) {
  for (int64_t i = blockIdx.x * (int64_t)blockDim.x + threadIdx.x; i < total;
       i += (int64_t)gridDim.x * blockDim.x) {
    int32_t v = en[i];
    if (v < 1 || (int64_t)v > nn) atomicOr(bad, 1);
  }
}

__global__ void k_count_incident(const int32_t *__restrict__ en, int64_t total,
                                 int32_t *__restrict__ cnt) {
  for (int64_t i = blockIdx.x * (int64_t)blockDim.x + threadIdx.x; i < total;
       i += (int64_t)gridDim.x * blockDim.x)
    atomicAdd(&cnt[en[i] - 1], 1);
}

__global__ void k_fill_incident(const int32_t *__restrict__ en, int64_t ne, int64_t total,
                                const int64_t *__restrict__ eoff, int32_t *__restrict__ fill,
                                int32_t *__restrict__ elist) {
  for (int64_t i = blockIdx.x * (int64_t)blockDim.x + threadIdx.x; i < total;
       i += (int64_t)gridDim.x * blockDim.x) {
    int32_t node = en[i] - 1;
    int32_t e = (int32_t)(i % ne);
    int32_t slot = atomicAdd(&fill[node], 1);
    elist[eoff[node] + slot] = e;
  }
}

// ascending element ids per node (lists are short: 8 for interior hex nodes)
__global__ void k_sort_incident(int64_t nn, const int64_t *__restrict__ eoff,
                                int32_t *__restrict__ elist) {
  int64_t i = blockIdx.x * (int64_t)blockDim.x + threadIdx.x;
  if (i >= nn) return;
  const int64_t s0 = eoff[i], s1 = eoff[i + 1];
  for (int64_t a = s0 + 1; a < s1; ++a) {
    int32_t v = elist[a];
    int64_t b = a - 1;
    while (b >= s0 && elist[b] > v) {
      elist[b + 1] = elist[b];
      --b;
    }
    elist[b + 1] = v;
  }
}

// Candidate j of node i is local node (j % nne) of the (j / nne)-th incident element.
// Pass 0 counts first occurrences; pass 1 writes them (0-based node ids) at noff[i].
// One wavefront per node: lane j holds candidate j (coalesced over the node's element list),
// is a first occurrence when no lower lane holds the same node (64 shuffles), and the first
// occurrences are compacted in lane order by a ballot -- the order of the serial scan.  A
// node with more than 64 candidates is scanned serially by lane 0.
template <int PASS>
__global__ void __launch_bounds__(256) k_node_neighbours(int64_t nn, int64_t ne, int nne,
                                                         const int32_t *__restrict__ en,
                                                         const int64_t *__restrict__ eoff,
                                                         const int32_t *__restrict__ elist,
                                                         int32_t *__restrict__ deg, const int64_t *__restrict__ noff,
                                                         int32_t *__restrict__ ncol) {
  const int64_t i = (blockIdx.x * (int64_t)blockDim.x + threadIdx.x) >> 6;
  const int lane = threadIdx.x & 63;
  if (i >= nn) return;  // whole wavefronts
  const int64_t s0 = eoff[i];
  const int64_t m = (eoff[i + 1] - s0) * nne;
  if (m <= 64) {
    const int32_t cj = lane < m ? en[(int64_t)(lane % nne) * ne + elist[s0 + lane / nne]] : 0;
    bool first = lane < m;
    for (int k = 0; k < 63; ++k) {
      const int32_t v = __shfl(cj, k, 64);
      if (k < lane && v == cj) first = false;
    }
    const uint64_t mask = __ballot(first);
    if (PASS == 0) {
      if (lane == 0) deg[i] = (int32_t)__popcll(mask);
    } else if (first) {
      ncol[noff[i] + __popcll(mask & ((1ull << lane) - 1ull))] = cj - 1;
    }
    return;
  }
  if (lane != 0) return;
  int32_t count = 0;
  int32_t *out = nullptr;
  if (PASS == 1) out = ncol + noff[i];
  for (int64_t j = 0; j < m; ++j) {
    const int32_t cj = en[(int64_t)(j % nne) * ne + elist[s0 + j / nne]];
    bool dup = false;
    if (PASS == 0) {
      for (int64_t k = 0; k < j; ++k)
        if (en[(int64_t)(k % nne) * ne + elist[s0 + k / nne]] == cj) {
          dup = true;
          break;
        }
    } else {
      for (int32_t k = 0; k < count; ++k)
        if (out[k] == cj - 1) {
          dup = true;
          break;
        }
    }
    if (!dup) {
      if (PASS == 1) out[count] = cj - 1;
      ++count;
    }
  }
  if (PASS == 0) deg[i] = count;
}

__global__ void k_max_deg(int64_t nn, const int32_t *__restrict__ deg, int32_t *__restrict__ mx) {
  int64_t i = blockIdx.x * (int64_t)blockDim.x + threadIdx.x;
  if (i < nn) atomicMax(mx, deg[i]);
}

// row_ptr (1-based) of the DOF-expanded CRS: row dof*i+d starts at
// 1 + dof^2*noff[i] + d*dof*deg[i]
__global__ void k_expand_rowptr(int64_t nn, int dof, const int64_t *__restrict__ noff,
                                const int32_t *__restrict__ deg, int64_t *__restrict__ row_ptr) {
  int64_t i = blockIdx.x * (int64_t)blockDim.x + threadIdx.x;
  if (i >= nn) return;
  const int64_t base = 1 + (int64_t)dof * dof * noff[i];
  for (int d = 0; d < dof; ++d) row_ptr[(int64_t)dof * i + d] = base + (int64_t)d * dof * deg[i];
  if (i == nn - 1) row_ptr[(int64_t)dof * nn] = 1 + (int64_t)dof * dof * noff[nn];
}

// one thread per (node, neighbour) block: writes its dof x dof column ids in every row of
// the node.  The owning node is found by binary search of noff.
__global__ void k_expand_cols(int64_t nn, int64_t nblocks, int dof, const int64_t *__restrict__ noff,
                              const int32_t *__restrict__ deg, const int32_t *__restrict__ ncol,
                              int32_t *__restrict__ col_idx) {
  for (int64_t t = blockIdx.x * (int64_t)blockDim.x + threadIdx.x; t < nblocks;
       t += (int64_t)gridDim.x * blockDim.x) {
    int64_t lo = 0, hi = nn;  // largest i with noff[i] <= t
    while (hi - lo > 1) {
      int64_t mid = (lo + hi) >> 1;
      if (noff[mid] <= t) lo = mid; else hi = mid;
    }
    const int64_t i = lo;
    const int64_t c = t - noff[i];
    const int64_t dg = deg[i];
    const int64_t base = (int64_t)dof * dof * noff[i];
    const int32_t colnode = ncol[t];
    for (int d = 0; d < dof; ++d)
      for (int k = 0; k < dof; ++k)
        col_idx[base + d * dof * dg + dof * c + k] = (int32_t)((int64_t)dof * colnode + k + 1);
  }
}

}  // namespace

int pf_graph_check_mesh(pf_ctx *ctx, const NodeGraph &g) {
  const int64_t total = g.ne * g.nne;
  if (total == 0) return PF_OK;
  DevBuf<int32_t> bad;
  PF_CHECK(bad.ensure(1));
  PF_HIP(hipMemsetAsync(bad.p, 0, sizeof(int32_t), ctx->stream));
  k_check_nodes<<<pf_blocks(total, 256, 16384), 256, 0, ctx->stream>>>(g.elemnod.p, total, g.nn, bad.p);
  int32_t h = 0;
  PF_HIP(hipMemcpyAsync(&h, bad.p, sizeof(int32_t), hipMemcpyDeviceToHost, ctx->stream));
  PF_HIP(hipStreamSynchronize(ctx->stream));
  if (h) {
    pf_set_error("ElemNod holds a node id outside 1..nn");
    return PF_INVALID;
  }
  return PF_OK;
}

int pf_graph_node_elems(pf_ctx *ctx, NodeGraph &g) {
  const int64_t total = g.ne * g.nne;
  if (total >= (int64_t)INT32_MAX || g.ne >= (int64_t)INT32_MAX || g.nn >= (int64_t)INT32_MAX) {
    pf_set_error("mesh too large for 32-bit element/node ids");
    return PF_INVALID;
  }
  PF_CHECK(pf_graph_check_mesh(ctx, g));
  DevBuf<int32_t> cnt;
  DevBuf<int64_t> tmp;
  PF_CHECK(cnt.ensure((size_t)g.nn + 1));
  PF_CHECK(g.eoff.ensure((size_t)g.nn + 1));
  PF_CHECK(g.elist.ensure((size_t)total + 1));
  PF_HIP(hipMemsetAsync(cnt.p, 0, sizeof(int32_t) * (g.nn + 1), ctx->stream));
  if (total > 0)
    k_count_incident<<<pf_blocks(total, 256, 16384), 256, 0, ctx->stream>>>(g.elemnod.p, total, cnt.p);
  PF_CHECK(pf_scan_i32_to_i64(ctx, cnt.p, g.nn, g.eoff.p, tmp));
  PF_HIP(hipMemsetAsync(cnt.p, 0, sizeof(int32_t) * (g.nn + 1), ctx->stream));
  if (total > 0) {
    k_fill_incident<<<pf_blocks(total, 256, 16384), 256, 0, ctx->stream>>>(g.elemnod.p, g.ne, total,
                                                                         g.eoff.p, cnt.p, g.elist.p);
    k_sort_incident<<<pf_blocks(g.nn, 256), 256, 0, ctx->stream>>>(g.nn, g.eoff.p, g.elist.p);
  }
  PF_HIP(hipGetLastError());
  PF_HIP(hipStreamSynchronize(ctx->stream));  // cnt/tmp are freed on return
  g.built = true;
  return PF_OK;
}

int pf_graph_node_cols(pf_ctx *ctx, NodeGraph &g) {
  if (!g.built) PF_CHECK(pf_graph_node_elems(ctx, g));
  DevBuf<int64_t> tmp;
  DevBuf<int32_t> mx;
  PF_CHECK(g.deg.ensure((size_t)g.nn + 1));
  PF_CHECK(g.noff.ensure((size_t)g.nn + 1));
  PF_CHECK(mx.ensure(1));
  if (g.nn > 0)
    k_node_neighbours<0><<<pf_blocks(g.nn * 64, 256, INT32_MAX), 256, 0, ctx->stream>>>(
        g.nn, g.ne, g.nne, g.elemnod.p, g.eoff.p, g.elist.p, g.deg.p, nullptr, nullptr);
  PF_CHECK(pf_scan_i32_to_i64(ctx, g.deg.p, g.nn, g.noff.p, tmp));
  int64_t nb = 0;
  PF_CHECK(pf_read_i64(ctx, g.noff.p + g.nn, &nb));
  PF_CHECK(g.ncol.ensure((size_t)nb + 1));
  PF_HIP(hipMemsetAsync(mx.p, 0, sizeof(int32_t), ctx->stream));
  if (g.nn > 0) {
    k_node_neighbours<1><<<pf_blocks(g.nn * 64, 256, INT32_MAX), 256, 0, ctx->stream>>>(
        g.nn, g.ne, g.nne, g.elemnod.p, g.eoff.p, g.elist.p, g.deg.p, g.noff.p, g.ncol.p);
    k_max_deg<<<pf_blocks(g.nn, 256), 256, 0, ctx->stream>>>(g.nn, g.deg.p, mx.p);
  }
  int32_t hmx = 0;
  PF_HIP(hipMemcpyAsync(&hmx, mx.p, sizeof(int32_t), hipMemcpyDeviceToHost, ctx->stream));
  PF_HIP(hipStreamSynchronize(ctx->stream));
  PF_HIP(hipGetLastError());
  g.nblocks = nb;
  g.maxdeg = hmx;
  g.cols_built = true;
  return PF_OK;
}

int pf_graph_expand(pf_ctx *ctx, const NodeGraph &g, int dof, int64_t *row_ptr, int32_t *col_idx) {
  if (!g.cols_built) {
    pf_set_error("graph expand before node columns");
    return PF_INVALID;
  }
  if (g.nn == 0) {
    int64_t one = 1;
    PF_HIP(hipMemcpyAsync(row_ptr, &one, sizeof(int64_t), hipMemcpyHostToDevice, ctx->stream));
    PF_HIP(hipStreamSynchronize(ctx->stream));
    return PF_OK;
  }
  k_expand_rowptr<<<pf_blocks(g.nn, 256), 256, 0, ctx->stream>>>(g.nn, dof, g.noff.p, g.deg.p, row_ptr);
  if (g.nblocks > 0 && col_idx)
    k_expand_cols<<<pf_blocks(g.nblocks, 256, 65536), 256, 0, ctx->stream>>>(
        g.nn, g.nblocks, dof, g.noff.p, g.deg.p, g.ncol.p, col_idx);
  PF_HIP(hipGetLastError());
  return PF_OK;
}

// ---- multi-domain pattern (setCRS with several domains, FEMSolverClass.f90:466-676) -------
// The reference concatenates the domains' ZeroMatrix patterns (block diagonal), adds every
// node pair of every overset connect (the 16 nodes of the own and the paired element) to a
// COO that drops repeats, and heap-sorts each row.  Here the (row node, column node) pairs of
// both sources become 64-bit keys, one radix sort and one unique pass leave the node
// neighbour lists sorted and distinct, and the usual DOF expansion writes the CRS.
#include <hipcub/hipcub.hpp>

namespace {

__global__ void k_keys_graph(int64_t nn, const int64_t *__restrict__ noff, const int32_t *__restrict__ ncol,
                             uint64_t *__restrict__ keys) {
  const int64_t i = blockIdx.x * (int64_t)blockDim.x + threadIdx.x;
  if (i >= nn) return;
  for (int64_t q = noff[i]; q < noff[i + 1]; ++q) keys[q] = ((uint64_t)i << 32) | (uint32_t)ncol[q];
}

// 1-based node of slot j of connect c: GPP slots 0-7 own element, 8-15 paired element; P2P
// (gp == 0) slot 0 the own node my[c] (0-based), 1-8 the paired element, 9-15 repeat slot 0
__device__ inline int32_t connect_node(int64_t ne, const int32_t *__restrict__ en, int32_t my, int32_t pair,
                                       bool p2p, int j) {
  if (p2p) return j >= 1 && j <= 8 ? en[(int64_t)(j - 1) * ne + pair] : my + 1;
  return j < 8 ? en[(int64_t)j * ne + my] : en[(int64_t)(j - 8) * ne + pair];
}

// connect c couples element my[c] and element pair[c] (0-based) of one mesh: 16 x 16 pairs
__global__ void k_keys_connect(int64_t ncon, int64_t ne, const int32_t *__restrict__ en,
                               const int32_t *__restrict__ my, const int32_t *__restrict__ pair,
                               const int32_t *__restrict__ gp, uint64_t *__restrict__ keys) {
  const int64_t t = blockIdx.x * (int64_t)blockDim.x + threadIdx.x;
  if (t >= ncon * 256) return;
  const int64_t c = t >> 8;
  const int j = (int)((t >> 4) & 15), k = (int)(t & 15);
  const bool p2p = gp && gp[c] == 0;
  const int32_t rj = connect_node(ne, en, my[c], pair[c], p2p, j);
  const int32_t ck = connect_node(ne, en, my[c], pair[c], p2p, k);
  keys[t] = ((uint64_t)(uint32_t)(rj - 1) << 32) | (uint32_t)(ck - 1);
}

__global__ void k_keys_count(int64_t u, const uint64_t *__restrict__ keys, int32_t *__restrict__ deg) {
  const int64_t t = blockIdx.x * (int64_t)blockDim.x + threadIdx.x;
  if (t >= u) return;
  atomicAdd(&deg[keys[t] >> 32], 1);
}

__global__ void k_keys_cols(int64_t u, const uint64_t *__restrict__ keys, int32_t *__restrict__ ncol) {
  const int64_t t = blockIdx.x * (int64_t)blockDim.x + threadIdx.x;
  if (t >= u) return;
  ncol[t] = (int32_t)(keys[t] & 0xffffffffu);
}

}  // namespace

int pf_graph_merged_cols(pf_ctx *ctx, NodeGraph &g, int64_t ncon, const int32_t *my, const int32_t *pair,
                         const int32_t *gp, NodeGraph &out) {
  if (g.nne != 8) {
    pf_set_error("overset connects need an 8-node mesh");
    return PF_INVALID;
  }
  if (!g.cols_built) PF_CHECK(pf_graph_node_cols(ctx, g));
  const int64_t total = g.nblocks + ncon * 256;
  out.nn = g.nn;
  out.ne = g.ne;
  out.nne = g.nne;
  PF_CHECK(out.deg.ensure((size_t)g.nn + 1));
  PF_CHECK(out.noff.ensure((size_t)g.nn + 1));
  DevBuf<uint64_t> ka, kb;
  DevBuf<int64_t> tmp, nsel;
  DevBuf<unsigned char> cub;
  PF_CHECK(ka.ensure((size_t)total + 1));
  PF_CHECK(kb.ensure((size_t)total + 1));
  PF_CHECK(nsel.ensure(1));
  if (g.nn > 0) k_keys_graph<<<pf_blocks(g.nn, 256, INT32_MAX), 256, 0, ctx->stream>>>(g.nn, g.noff.p, g.ncol.p, ka.p);
  if (ncon > 0)
    k_keys_connect<<<pf_blocks(ncon * 256, 256, INT32_MAX), 256, 0, ctx->stream>>>(ncon, g.ne, g.elemnod.p, my,
                                                                                  pair, gp, ka.p + g.nblocks);
  PF_HIP(hipGetLastError());
  int end_bit = 33;
  while (end_bit < 64 && ((int64_t)1 << (end_bit - 32)) < g.nn) ++end_bit;
  size_t b1 = 0, b2 = 0;
  PF_HIP(hipcub::DeviceRadixSort::SortKeys(nullptr, b1, ka.p, kb.p, total, 0, end_bit, ctx->stream));
  PF_HIP(hipcub::DeviceSelect::Unique(nullptr, b2, kb.p, ka.p, nsel.p, total, ctx->stream));
  PF_CHECK(cub.ensure(std::max(b1, b2) + 1));
  PF_HIP(hipcub::DeviceRadixSort::SortKeys(cub.p, b1, ka.p, kb.p, total, 0, end_bit, ctx->stream));
  PF_HIP(hipcub::DeviceSelect::Unique(cub.p, b2, kb.p, ka.p, nsel.p, total, ctx->stream));
  int64_t u = 0;
  PF_CHECK(pf_read_i64(ctx, nsel.p, &u));
  PF_HIP(hipMemsetAsync(out.deg.p, 0, sizeof(int32_t) * (g.nn + 1), ctx->stream));
  PF_CHECK(out.ncol.ensure((size_t)u + 1));
  if (u > 0) {
    k_keys_count<<<pf_blocks(u, 256, INT32_MAX), 256, 0, ctx->stream>>>(u, ka.p, out.deg.p);
    k_keys_cols<<<pf_blocks(u, 256, INT32_MAX), 256, 0, ctx->stream>>>(u, ka.p, out.ncol.p);
  }
  PF_HIP(hipGetLastError());
  PF_CHECK(pf_scan_i32_to_i64(ctx, out.deg.p, g.nn, out.noff.p, tmp));
  PF_HIP(hipStreamSynchronize(ctx->stream));  // scratch buffers are freed on return
  out.nblocks = u;
  out.cols_built = true;
  return PF_OK;
}

int pf_sort_u64(pf_ctx *ctx, uint64_t *keys, uint64_t *tmp, int64_t n, int end_bit, bool *in_tmp) {
  *in_tmp = false;
  if (n <= 1) return PF_OK;
  size_t bytes = 0;
  PF_HIP(hipcub::DeviceRadixSort::SortKeys(nullptr, bytes, keys, tmp, n, 0, end_bit, ctx->stream));
  DevBuf<unsigned char> cub;
  PF_CHECK(cub.ensure(bytes + 1));
  PF_HIP(hipcub::DeviceRadixSort::SortKeys(cub.p, bytes, keys, tmp, n, 0, end_bit, ctx->stream));
  PF_HIP(hipStreamSynchronize(ctx->stream));  // cub scratch is freed on return
  *in_tmp = true;
  return PF_OK;
}
