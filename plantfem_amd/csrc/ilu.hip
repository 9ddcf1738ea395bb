// ilu.hip -- the ILU(0) preconditioner of bicgstab_CRS_ILU (LinearSolverClass.f90:4104-4229)
// on the device: the factor as ILUCRS(fill_in_order=0) builds it (SparseClass.f90:1629-1700)
// and its application ILU_matvecCRS (SparseClass.f90:1795-1845).
//
// Factor.  Rows i = 2..n are processed in order; for every stored column k < i of row i, in
// STORED order (ZeroMatrix stores columns in first-appearance order, the multi-domain setCRS
// sorts them), A_ik = a_ik / d_k replaces a_ik and every stored a_ij with j > k gets
// - A_ik * get(k, j) added (get reads 0 for a column row k does not store); d_i = get(i,i)
// afterwards.  Row i needs the final rows k < i it stores, so rows are scheduled by level
// (level = 1 + the largest level of the rows it depends on) and one kernel launch runs one
// level, a wavefront per row group.  Inside a row the k-steps run in the reference's order
// with the same roundings (separate multiply and add), so the factor is bitwise the
// reference's.  Groups: with the DOF=3 node-block structure (the three rows of a node store
// the same columns, in full node triples) a group is one node's three rows, processed in
// order by one wavefront; otherwise a group is one row.  Grouping only changes the schedule.
//
// Apply.  ILU_matvecCRS computes y = v, forward substitution with the unit lower triangle
// (stored columns < i), then y_i / d_i (1 where d_i = 0).  Its closing backward loop stores
// into new_vector(i) inside the column loop and then restores the value from before the
// loop, so the strict upper triangle is never applied: y = D^-1 L^-1 v.  The sweep runs
// level by level like the factor; the lower entries of every row are stored once more,
// compactly in level order (columns in earlier groups), together with the within-group
// entries and the diagonal, so one level reads one contiguous stretch of the factor.
#include <algorithm>
#include <cstring>
#include <vector>

#include "spmv_dev.hpp"

namespace {

constexpr int FT = 64;     // threads of a factor block: one wavefront per row group
constexpr int MAXR = 512;  // longest row the factor holds in LDS
constexpr int ST = 256;    // threads of a sweep block
constexpr int SSUB = 16;   // lanes per row group in the sweep

// first stored position of column c (1-based) in an LDS row, -1 when absent (wave-parallel)
__device__ inline int lds_find(const int32_t *cols, int len, int32_t c) {
  int best = MAXR;
  for (int q = threadIdx.x; q < len; q += FT)
    if (cols[q] == c && q < best) best = q;
#pragma unroll
  for (int off = 32; off > 0; off >>= 1) best = min(best, __shfl_xor(best, off, 64));
  return best == MAXR ? -1 : best;
}

// column -> position map of one row in LDS (open addressing, HT slots, load <= 1/2)
constexpr int HT = 2 * MAXR;
__device__ inline uint32_t hslot(int32_t c) { return ((uint32_t)c * 0x9E3779B1u) >> (32 - 10); }
static_assert(HT == 1024, "hslot takes 10 bits");

__device__ inline int hfind(const int32_t *hkey, const int16_t *hpos, int32_t c) {
  uint32_t h = hslot(c);
  for (;;) {
    const int32_t k = hkey[h];
    if (k == c) return hpos[h];
    if (k == -1) return -1;
    h = (h + 1) & (HT - 1);
  }
}

// err bits: 1 row longer than MAXR, 2 a row stores a column twice
template <int G>
__global__ void __launch_bounds__(FT) k_ilu_level(const int64_t *__restrict__ row_ptr,
                                                  const int32_t *__restrict__ col_idx, double *__restrict__ f,
                                                  double *__restrict__ dv, const int32_t *__restrict__ grp,
                                                  int64_t p0, int32_t *__restrict__ err) {
  __shared__ int32_t ci[G][MAXR];
  __shared__ double vi[G][MAXR];
  __shared__ int32_t ck[MAXR];
  __shared__ double vk[MAXR];
  __shared__ int32_t lenv[G];
  __shared__ double dloc[G];
  __shared__ int32_t hkey[HT];
  __shared__ int16_t hpos[HT];
  const int lane = threadIdx.x;
  const int64_t g = grp[p0 + blockIdx.x];
  for (int a = 0; a < G; ++a) {
    const int64_t i = (int64_t)G * g + a;  // 0-based row
    const int64_t ra = row_ptr[i] - 1;
    const int len = (int)(row_ptr[i + 1] - 1 - ra);
    if (len > MAXR) {
      if (lane == 0) atomicOr(err, 1);
      return;
    }
    for (int q = lane; q < len; q += FT) {
      ci[a][q] = col_idx[ra + q];
      vi[a][q] = f[ra + q];
    }
    if (lane == 0) lenv[a] = len;
    __syncthreads();
    bool dup = false;  // the in-place updates below assume one entry per column
    for (int q = lane; q < len; q += FT)
      for (int u = 0; u < q; ++u) dup |= ci[a][u] == ci[a][q];
    if (__any(dup)) {
      if (lane == 0) atomicOr(err, 2);
      return;
    }
    for (int h = lane; h < HT; h += FT) hkey[h] = -1;
    __syncthreads();
    for (int q = lane; q < len; q += FT) {  // distinct columns: every insert claims a new slot
      const int32_t c = ci[a][q];
      uint32_t h = hslot(c);
      while (atomicCAS(&hkey[h], -1, c) != -1) h = (h + 1) & (HT - 1);
      hpos[h] = (int16_t)q;
    }
    __syncthreads();
    for (int kk = 0; kk < len; ++kk) {
      const int32_t k = ci[a][kk];  // 1-based; the same value in every lane
      if ((int64_t)k >= i + 1) continue;
      const int64_t kr = k - 1;
      const int64_t kg = kr / G;
      const bool local = kg == g;  // an earlier row of this group, final in LDS
      const double dk = local ? dloc[kr - (int64_t)G * g] : dv[kr];
      const double aik = vi[a][kk] / dk;
      const int32_t *cp;
      const double *vp;
      int lenk;
      if (local) {
        const int b = (int)(kr - (int64_t)G * g);
        cp = ci[b];
        vp = vi[b];
        lenk = lenv[b];
      } else {
        const int64_t rk = row_ptr[kr] - 1;
        lenk = (int)(row_ptr[kr + 1] - 1 - rk);
        if (lenk > MAXR) {
          if (lane == 0) atomicOr(err, 1);
          return;
        }
        for (int q = lane; q < lenk; q += FT) {
          ck[q] = col_idx[rk + q];
          vk[q] = f[rk + q];
        }
        cp = ck;
        vp = vk;
      }
      __syncthreads();
      if (lane == 0) vi[a][kk] = aik;  // update(i,k,A_ik); column k itself is not touched below
      if (isfinite(aik)) {
        // add(i,j, -A_ik*get(k,j)) for the stored j > k of row i: a column row k does not
        // store adds -(A_ik*0) = -0, which leaves every value as it is, so only the entries of
        // row k are visited, each finding its column in row i through the hash
        for (int m = lane; m < lenk; m += FT) {
          const int32_t j = cp[m];
          if (j < k + 1) continue;
          const int q = hfind(hkey, hpos, j);
          if (q >= 0) vi[a][q] = __dadd_rn(vi[a][q], -__dmul_rn(aik, vp[m]));
        }
      } else {  // -(inf*0) = NaN reaches every stored j > k, as in the reference
        for (int jj = lane; jj < len; jj += FT) {
          const int32_t j = ci[a][jj];
          if (j < k + 1) continue;
          double gv = 0.0;  // get(k,j)
          for (int q = 0; q < lenk; ++q)
            if (cp[q] == j) {
              gv = vp[q];
              break;
            }
          vi[a][jj] = __dadd_rn(vi[a][jj], -__dmul_rn(aik, gv));
        }
      }
      __syncthreads();
    }
    const int qd = lds_find(ci[a], len, (int32_t)(i + 1));
    const double di = qd < 0 ? 0.0 : vi[a][qd];
    for (int q = lane; q < len; q += FT) f[ra + q] = vi[a][q];
    if (lane == 0) {
      dloc[a] = di;
      dv[i] = di;
    }
    __syncthreads();
  }
}

// lower groups of group g: count (fill = false) or write them
__global__ void k_ilu_deps(int64_t ngroups, int G, const int64_t *__restrict__ row_ptr,
                           const int32_t *__restrict__ col_idx, const int64_t *__restrict__ off,
                           int32_t *__restrict__ cnt, int32_t *__restrict__ dep) {
  const int64_t g = blockIdx.x * (int64_t)blockDim.x + threadIdx.x;
  if (g >= ngroups) return;
  const int64_t i = (int64_t)G * g;  // the rows of a group store the same node columns
  int32_t c = 0;
  int64_t o = off ? off[g] : 0;
  for (int64_t q = row_ptr[i] - 1; q < row_ptr[i + 1] - 1; ++q) {
    const int64_t col = (int64_t)col_idx[q] - 1;
    if (col % G != 0) continue;
    const int64_t h = col / G;
    if (h >= g) continue;
    if (off) dep[o++] = (int32_t)h;
    ++c;
  }
  if (!off) cnt[g] = c;
}

// lower entries (columns in earlier groups) per row in level order: count or fill
template <int G>
__global__ void k_ilu_lower(int64_t ngroups, const int32_t *__restrict__ grp, const int64_t *__restrict__ row_ptr,
                            const int32_t *__restrict__ col_idx, const double *__restrict__ f,
                            const double *__restrict__ dv, const int64_t *__restrict__ lptr,
                            int32_t *__restrict__ cnt, int32_t *__restrict__ lcol, double *__restrict__ lval,
                            double *__restrict__ wv, double *__restrict__ dd) {
  const int64_t t = blockIdx.x * (int64_t)blockDim.x + threadIdx.x;
  if (t >= ngroups * G) return;
  const int64_t pos = t / G;
  const int a = (int)(t % G);
  const int64_t g = grp[pos];
  const int64_t i = (int64_t)G * g + a;
  const int64_t lo = (int64_t)G * g;  // first column (0-based) of the group
  int32_t c = 0;
  int64_t o = lptr ? lptr[t] : 0;
  for (int64_t q = row_ptr[i] - 1; q < row_ptr[i + 1] - 1; ++q) {
    const int64_t col = (int64_t)col_idx[q] - 1;
    if (col < lo) {
      if (lptr) {
        lcol[o] = (int32_t)col;
        lval[o] = f[q];
        ++o;
      }
      ++c;
    } else if (lptr && G == 3 && col < i) {  // within-group lower entry (a, b), b < a
      const int b = (int)(col - lo);
      wv[pos * 3 + (a == 1 ? 0 : 1 + b)] = f[q];
    }
  }
  if (!lptr) {
    cnt[t] = c;
  } else {
    const double d = dv[i];
    dd[t] = d == 0.0 ? 1.0 : d;
  }
}

// diagCRS (SparseClass.f90:1614-1624): first stored diagonal entry, 0 when absent
__global__ void k_diag_rows(int64_t n, const int64_t *__restrict__ row_ptr, const int32_t *__restrict__ col_idx,
                            const double *__restrict__ f, double *__restrict__ dv) {
  const int64_t i = blockIdx.x * (int64_t)blockDim.x + threadIdx.x;
  if (i >= n) return;
  double d = 0.0;
  for (int64_t q = row_ptr[i] - 1; q < row_ptr[i + 1] - 1; ++q)
    if (col_idx[q] == i + 1) {
      d = f[q];
      break;
    }
  dv[i] = d;
}

__global__ void k_copy_f64(int64_t n, const double *__restrict__ a, double *__restrict__ b) {
  for (int64_t i = blockIdx.x * (int64_t)blockDim.x + threadIdx.x; i < n; i += (int64_t)gridDim.x * blockDim.x)
    b[i] = a[i];
}

// one level of y = D^-1 L^-1 src.  MODE 0: src = v; 1: p = r + beta (p - omega y) is formed
// and stored (the BiCGSTAB direction update); 2: e = r - alpha y is formed and stored.
template <int G, int MODE>
__global__ void __launch_bounds__(ST) k_ilu_sweep(int64_t p0, int64_t p1, const int32_t *__restrict__ grp,
                                                  const int64_t *__restrict__ lptr, const int32_t *__restrict__ lcol,
                                                  const double *__restrict__ lval, const double *__restrict__ wv,
                                                  const double *__restrict__ dd, IluSweepArgs sa,
                                                  double *__restrict__ t) {
  const int64_t pos = p0 + (blockIdx.x * (int64_t)ST + threadIdx.x) / SSUB;
  const int sub = threadIdx.x % SSUB;
  if (pos >= p1) return;  // whole subgroups leave together
  // the stop flag is read with the other loads and tested before anything is stored
  const int32_t done = sa.done ? *sa.done : 0;
  const int64_t g = grp[pos];
  double s[G];
#pragma unroll
  for (int a = 0; a < G; ++a) {
    double acc = 0.0;
    const int64_t q1 = lptr[pos * G + a + 1];
    for (int64_t q = lptr[pos * G + a] + sub; q < q1; q += SSUB) acc += lval[q] * t[lcol[q]];
    s[a] = pfdev::subgroup_sum<double, SSUB>(acc);
  }
  if (sub != 0 || done) return;
  double tv[G];
#pragma unroll
  for (int a = 0; a < G; ++a) {
    const int64_t i = (int64_t)G * g + a;
    double src;
    if (MODE == 0) {
      src = sa.v[i];
    } else if (MODE == 1) {
      src = sa.r[i] + *sa.beta * (sa.pe[i] - *sa.omega * sa.y[i]);
      sa.pe[i] = src;
    } else {
      src = sa.r[i] - *sa.alpha * sa.y[i];
      sa.pe[i] = src;
    }
    double x = src - s[a];
    if (G == 3) {
      if (a == 1) x -= wv[pos * 3] * tv[0];
      if (a == 2) x -= wv[pos * 3 + 1] * tv[0] + wv[pos * 3 + 2] * tv[1];
    }
    tv[a] = x;
    t[i] = x;
    sa.out[i] = x / dd[pos * G + a];
  }
}

}  // namespace

int pf_ilu_factor(pf_ctx *ctx, int64_t n, const int64_t *row_ptr, const int32_t *col_idx, const double *val,
                  int64_t nnz, int G, IluFactor &F, bool factor) {
  F.ready = false;
  if (G != 1 && G != 3) {
    pf_set_error("internal: ILU group size %d", G);
    return PF_INVALID;
  }
  if (G == 3 && n % 3 != 0) G = 1;
  F.n = n;
  F.nnz = nnz;
  F.G = G;
  F.ngroups = n / G;
  const int64_t ng = F.ngroups;
  hipStream_t st = ctx->stream;
  PF_CHECK(F.f.ensure((size_t)nnz + 1));
  PF_CHECK(F.dv.ensure((size_t)n + 1));
  PF_CHECK(F.t.ensure((size_t)n + 1));
  PF_CHECK(F.err.ensure(1));
  if (n <= 0) {
    F.lev.assign(1, 0);
    F.ready = true;
    return PF_OK;
  }
  // level schedule of the groups (host: a longest-path pass over the lower-group lists)
  DevBuf<int32_t> cnt, dep;
  DevBuf<int64_t> off, tmp;
  PF_CHECK(cnt.ensure((size_t)ng + 1));
  PF_CHECK(off.ensure((size_t)ng + 1));
  k_ilu_deps<<<pf_blocks(ng, 256, INT32_MAX), 256, 0, st>>>(ng, G, row_ptr, col_idx, nullptr, cnt.p, nullptr);
  PF_HIP(hipGetLastError());
  PF_CHECK(pf_scan_i32_to_i64(ctx, cnt.p, ng, off.p, tmp));
  int64_t ndep = 0;
  PF_CHECK(pf_read_i64(ctx, off.p + ng, &ndep));
  PF_CHECK(dep.ensure((size_t)ndep + 1));
  k_ilu_deps<<<pf_blocks(ng, 256, INT32_MAX), 256, 0, st>>>(ng, G, row_ptr, col_idx, off.p, nullptr, dep.p);
  PF_HIP(hipGetLastError());
  std::vector<int64_t> h_off((size_t)ng + 1);
  std::vector<int32_t> h_dep((size_t)ndep + 1);
  PF_HIP(hipMemcpyAsync(h_off.data(), off.p, sizeof(int64_t) * (ng + 1), hipMemcpyDeviceToHost, st));
  PF_HIP(hipMemcpyAsync(h_dep.data(), dep.p, sizeof(int32_t) * (ndep + 1), hipMemcpyDeviceToHost, st));
  PF_HIP(hipStreamSynchronize(st));
  std::vector<int32_t> lev((size_t)ng);
  int32_t nlev = 0;
  for (int64_t g = 0; g < ng; ++g) {
    int32_t l = 0;
    for (int64_t q = h_off[g]; q < h_off[g + 1]; ++q) l = std::max(l, lev[h_dep[q]] + 1);
    lev[g] = l;
    nlev = std::max(nlev, l + 1);
  }
  F.lev.assign((size_t)nlev + 1, 0);
  for (int64_t g = 0; g < ng; ++g) ++F.lev[lev[g] + 1];
  for (int32_t l = 0; l < nlev; ++l) F.lev[l + 1] += F.lev[l];
  std::vector<int32_t> h_grp((size_t)ng);
  {
    std::vector<int64_t> fill(F.lev.begin(), F.lev.end() - 1);
    for (int64_t g = 0; g < ng; ++g) h_grp[fill[lev[g]]++] = (int32_t)g;
  }
  PF_CHECK(F.grp.ensure((size_t)ng + 1));
  PF_HIP(hipMemcpyAsync(F.grp.p, h_grp.data(), sizeof(int32_t) * ng, hipMemcpyHostToDevice, st));

  // factor, level by level (factor = false: val already holds a factor; only its diagonal
  // is read here)
  k_copy_f64<<<pf_blocks(nnz, 256, 65536), 256, 0, st>>>(nnz, val, F.f.p);
  PF_HIP(hipMemsetAsync(F.err.p, 0, sizeof(int32_t), st));
  if (!factor) k_diag_rows<<<pf_blocks(n, 256, INT32_MAX), 256, 0, st>>>(n, row_ptr, col_idx, F.f.p, F.dv.p);
  for (int32_t l = 0; factor && l < nlev; ++l) {
    const int64_t c = F.lev[l + 1] - F.lev[l];
    if (G == 3)
      k_ilu_level<3><<<(unsigned)c, FT, 0, st>>>(row_ptr, col_idx, F.f.p, F.dv.p, F.grp.p, F.lev[l], F.err.p);
    else
      k_ilu_level<1><<<(unsigned)c, FT, 0, st>>>(row_ptr, col_idx, F.f.p, F.dv.p, F.grp.p, F.lev[l], F.err.p);
  }
  PF_HIP(hipGetLastError());
  int32_t err = 0;
  PF_HIP(hipMemcpyAsync(&err, F.err.p, sizeof(int32_t), hipMemcpyDeviceToHost, st));
  PF_HIP(hipStreamSynchronize(st));
  if (err & 1) {
    pf_set_error("ILU(0): a CRS row holds more than %d entries", MAXR);
    return PF_INVALID;
  }
  if (err & 2) {
    pf_set_error("ILU(0): a CRS row stores a column twice");
    return PF_INVALID;
  }

  // compact lower triangle in level order for the sweeps
  const int64_t nr = ng * G;
  PF_CHECK(cnt.ensure((size_t)nr + 1));
  PF_CHECK(F.lptr.ensure((size_t)nr + 1));
  PF_CHECK(F.dd.ensure((size_t)nr + 1));
  PF_CHECK(F.wv.ensure((size_t)(G == 3 ? ng * 3 : 1)));
  const unsigned bl = pf_blocks(nr, 256, INT32_MAX);
  if (G == 3)
    k_ilu_lower<3><<<bl, 256, 0, st>>>(ng, F.grp.p, row_ptr, col_idx, F.f.p, F.dv.p, nullptr, cnt.p, nullptr,
                                       nullptr, nullptr, nullptr);
  else
    k_ilu_lower<1><<<bl, 256, 0, st>>>(ng, F.grp.p, row_ptr, col_idx, F.f.p, F.dv.p, nullptr, cnt.p, nullptr,
                                       nullptr, nullptr, nullptr);
  PF_HIP(hipGetLastError());
  PF_CHECK(pf_scan_i32_to_i64(ctx, cnt.p, nr, F.lptr.p, tmp));
  PF_CHECK(pf_read_i64(ctx, F.lptr.p + nr, &F.lnnz));
  PF_CHECK(F.lcol.ensure((size_t)F.lnnz + 1));
  PF_CHECK(F.lval.ensure((size_t)F.lnnz + 1));
  if (G == 3)
    k_ilu_lower<3><<<bl, 256, 0, st>>>(ng, F.grp.p, row_ptr, col_idx, F.f.p, F.dv.p, F.lptr.p, nullptr, F.lcol.p,
                                       F.lval.p, F.wv.p, F.dd.p);
  else
    k_ilu_lower<1><<<bl, 256, 0, st>>>(ng, F.grp.p, row_ptr, col_idx, F.f.p, F.dv.p, F.lptr.p, nullptr, F.lcol.p,
                                       F.lval.p, F.wv.p, F.dd.p);
  PF_HIP(hipGetLastError());
  PF_HIP(hipStreamSynchronize(st));
  F.ready = true;
  return PF_OK;
}

int pf_ilu_sweep_enqueue(pf_ctx *ctx, IluFactor &F, int mode, const IluSweepArgs &a) {
  if (!F.ready) {
    pf_set_error("internal: ILU sweep without a factor");
    return PF_INVALID;
  }
  if (mode < 0 || mode > 2 || !a.out || (mode == 0 && !a.v) || (mode > 0 && (!a.r || !a.pe || !a.y)) ||
      (mode == 1 && (!a.beta || !a.omega)) || (mode == 2 && !a.alpha)) {
    pf_set_error("internal: ILU sweep arguments");
    return PF_INVALID;
  }
  hipStream_t st = ctx->stream;
  const int nlev = (int)F.lev.size() - 1;
  for (int l = 0; l < nlev; ++l) {
    const int64_t p0 = F.lev[l], p1 = F.lev[l + 1];
    const unsigned blocks = pf_blocks((p1 - p0) * SSUB, ST, INT32_MAX);
#define PF_SWEEP(GV, MV)                                                                                 \
  k_ilu_sweep<GV, MV><<<blocks, ST, 0, st>>>(p0, p1, F.grp.p, F.lptr.p, F.lcol.p, F.lval.p, F.wv.p, F.dd.p, a, \
                                              F.t.p)
    if (F.G == 3) {
      if (mode == 0) PF_SWEEP(3, 0);
      else if (mode == 1) PF_SWEEP(3, 1);
      else PF_SWEEP(3, 2);
    } else {
      if (mode == 0) PF_SWEEP(1, 0);
      else if (mode == 1) PF_SWEEP(1, 1);
      else PF_SWEEP(1, 2);
    }
#undef PF_SWEEP
  }
  PF_HIP(hipGetLastError());
  return PF_OK;
}

// JacobiPreconditionerCRS (FEMSolverClass.f90:2508-2529): every row whose diagonal (the
// last stored entry of the column, getCRSval :2488-2503) is nonzero is multiplied by its
// reciprocal, the right-hand side entry too
namespace {
// one SUB-lane subgroup per row: the lanes search the row for its diagonal (the one lane that
// finds it contributes A_kk, the others 0, so the subgroup sum is A_kk exactly), then scale
// their entries; lane 0 scales the right-hand side
template <int SUB>
__global__ void __launch_bounds__(256) k_jacobi_rows(int64_t n, const int64_t *__restrict__ row_ptr,
                                                     const int32_t *__restrict__ col_idx, double *__restrict__ val,
                                                     double *__restrict__ rhs) {
  const int64_t g = blockIdx.x * (int64_t)blockDim.x + threadIdx.x;
  const int64_t i = g / SUB;
  const int lane = (int)(g % SUB);
  const bool live = i < n;
  int64_t a = 0, b = 0;
  if (live) {
    a = row_ptr[i] - 1;
    b = row_ptr[i + 1] - 1;
  }
  double akk = 0.0;
  for (int64_t k = a + lane; k < b; k += SUB)
    if (col_idx[k] == i + 1) akk = val[k];
#pragma unroll
  for (int o = SUB / 2; o > 0; o >>= 1) akk += __shfl_xor(akk, o, SUB);
  if (!live || akk == 0.0) return;
  const double inv = 1.0 / akk;
  if (lane == 0 && rhs) rhs[i] = rhs[i] * inv;
  for (int64_t k = a + lane; k < b; k += SUB) val[k] = val[k] * inv;
}
}  // namespace

int pf_jacobi_rows(pf_ctx *ctx, int64_t n, const int64_t *row_ptr, const int32_t *col_idx, double *val,
                   double *rhs) {
  if (n <= 0) return PF_OK;
  k_jacobi_rows<16><<<pf_blocks(n * 16, 256, INT32_MAX), 256, 0, ctx->stream>>>(n, row_ptr, col_idx, val, rhs);
  PF_HIP(hipGetLastError());
  return PF_OK;
}
