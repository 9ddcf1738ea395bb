// eig.hip -- eigFEMSolver (FEMSolverClass.f90:1384-1541) on the device: A u = lambda B u for
// the matrices kept as "A" (stiffness) and "B" (mass) by keepThisMatrixAs, with the DOFs given
// to fix_eig removed from both (reduce_crs_matrix) and put back as zeros in the eigenvectors.
//
//  * every eigenpair (the reference's default: LAPACK DSPGVD on the packed upper triangles of
//    the reduced matrices): the reduced matrices are scattered into dense column-major arrays
//    on the device and rocSOLVER's dsygvd (itype 1, eigenvectors, upper) solves them; the
//    eigenvectors come back B-orthonormal, eigenvalues ascending, as DSPGVD returns them.
//  * the nev lowest pairs (the reference's use_LOBPCG option): block LOBPCG in the full DOF
//    space with the fixed rows held at zero, the CRS SpMV for A and B, a Jacobi (diag A)
//    preconditioner, Rayleigh-Ritz on [X W P] with rocBLAS Gram products; the projected
//    pencil (at most 3 nev square, nev <= 64) is solved on the host by Cholesky of the B-Gram
//    matrix and cyclic Jacobi rotations (a few microseconds; a device solver for matrices
//    this small costs milliseconds of launches); W and P are scaled to unit B-norm first.
#include <rocblas/rocblas.h>
#include <rocsolver/rocsolver.h>

#include <algorithm>
#include <cmath>
#include <cstring>
#include <vector>

#include "pf_internal.hpp"

#define PF_RB(call)                                                                 \
  do {                                                                              \
    rocblas_status s_ = (call);                                                     \
    if (s_ != rocblas_status_success) {                                             \
      pf_set_error("%s:%d %s -> rocblas status %d", __FILE__, __LINE__, #call, (int)s_); \
      return PF_ERR_HIP;                                                            \
    }                                                                               \
  } while (0)

namespace {

// dense reduced matrix, column-major nf x nf: D[red(j)*nf + red(i)] = a_ij for free i, j
__global__ void k_dense_scatter(int64_t n, const int64_t *__restrict__ rp, const int32_t *__restrict__ ci,
                                const double *__restrict__ val, const int32_t *__restrict__ red, int64_t nf,
                                double *__restrict__ D) {
  const int64_t i = blockIdx.x * (int64_t)blockDim.x + threadIdx.x;
  if (i >= n) return;
  const int32_t ir = red[i];
  if (ir < 0) return;
  for (int64_t k = rp[i] - 1; k < rp[i + 1] - 1; ++k) {
    const int32_t jr = red[ci[k] - 1];
    if (jr >= 0) D[(int64_t)jr * nf + ir] += val[k];
  }
}

// out (n x cols) = reduced vectors Z (nf x cols) put back at the free DOFs, 0 at the fixed ones
__global__ void k_expand(int64_t n, int64_t nf, int64_t cols, const int32_t *__restrict__ red,
                         const double *__restrict__ Z, double *__restrict__ out) {
  for (int64_t t = blockIdx.x * (int64_t)blockDim.x + threadIdx.x; t < n * cols;
       t += (int64_t)gridDim.x * blockDim.x) {
    const int64_t c = t / n, i = t - c * n;
    const int32_t ir = red[i];
    out[t] = ir >= 0 ? Z[c * nf + ir] : 0.0;
  }
}

// deterministic start block: uniform(-0.5, 0.5) from a hash of (row, column), 0 at fixed rows
__global__ void k_init_block(int64_t n, int64_t m, const int32_t *__restrict__ red, double *__restrict__ X) {
  for (int64_t t = blockIdx.x * (int64_t)blockDim.x + threadIdx.x; t < n * m; t += (int64_t)gridDim.x * blockDim.x) {
    const int64_t i = t % n;
    uint64_t h = (uint64_t)t * 0x9E3779B97F4A7C15ull + 0x632BE59BD9B4E019ull;
    h ^= h >> 31;
    h *= 0xBF58476D1CE4E5B9ull;
    h ^= h >> 29;
    X[t] = red[i] >= 0 ? (double)(h >> 11) * (1.0 / 9007199254740992.0) - 0.5 : 0.0;
  }
}

// W = dinv (AX - lambda BX), 0 at fixed rows; R (residual) kept for the norms
__global__ void k_residual_block(int64_t n, int64_t m, const double *__restrict__ AX, const double *__restrict__ BX,
                                 const double *__restrict__ lam, const double *__restrict__ dinv,
                                 const int32_t *__restrict__ red, double *__restrict__ R, double *__restrict__ W) {
  for (int64_t t = blockIdx.x * (int64_t)blockDim.x + threadIdx.x; t < n * m; t += (int64_t)gridDim.x * blockDim.x) {
    const int64_t c = t / n, i = t - c * n;
    const double r = red[i] >= 0 ? AX[t] - lam[c] * BX[t] : 0.0;
    R[t] = r;
    W[t] = dinv[i] * r;
  }
}

__global__ void k_dinv(int64_t n, const int64_t *__restrict__ rp, const int32_t *__restrict__ ci,
                       const double *__restrict__ val, const int32_t *__restrict__ red, double *__restrict__ dinv) {
  const int64_t i = blockIdx.x * (int64_t)blockDim.x + threadIdx.x;
  if (i >= n) return;
  double d = 0.0;
  for (int64_t k = rp[i] - 1; k < rp[i + 1] - 1; ++k)
    if (ci[k] == i + 1) {
      d = val[k];
      break;
    }
  dinv[i] = (red[i] >= 0 && d != 0.0) ? 1.0 / fabs(d) : 0.0;
}

// columns c of S, AS, BS (n rows) scaled by s[c]
__global__ void k_scale_cols(int64_t n, int64_t c0, int64_t cols, const double *__restrict__ s, double *__restrict__ S,
                             double *__restrict__ AS, double *__restrict__ BS) {
  for (int64_t t = blockIdx.x * (int64_t)blockDim.x + threadIdx.x; t < n * cols; t += (int64_t)gridDim.x * blockDim.x) {
    const int64_t c = t / n;
    const int64_t g = c0 * n + t;
    const double f = s[c];
    S[g] *= f;
    AS[g] *= f;
    BS[g] *= f;
  }
}

__global__ void k_copy(int64_t count, const double *__restrict__ a, double *__restrict__ b) {
  for (int64_t t = blockIdx.x * (int64_t)blockDim.x + threadIdx.x; t < count; t += (int64_t)gridDim.x * blockDim.x)
    b[t] = a[t];
}

// Rayleigh-Ritz update of the search block, in place: for M = S, AS, BS (leading dimension
// n, k columns [X | W | P]) and the lowest m Ritz vectors C (k x m, leading dimension ldc),
// X_M = M C and, with k > m, P_M = [W P]_M C[m:k, :] into columns 2m.. of M.  One thread per
// row reads the row's k entries before it writes the row's new ones.
template <int MAXM>
__global__ void __launch_bounds__(256) k_rr_update(int64_t n, int k, int m, const double *__restrict__ C, int ldc,
                                                   double *S, double *AS, double *BS) {
  for (int64_t i = blockIdx.x * (int64_t)blockDim.x + threadIdx.x; i < n; i += (int64_t)gridDim.x * blockDim.x) {
    for (int w = 0; w < 3; ++w) {
      double *M = w == 0 ? S : (w == 1 ? AS : BS);
      double xo[MAXM], po[MAXM];
#pragma unroll
      for (int j = 0; j < MAXM; ++j) xo[j] = po[j] = 0.0;
      for (int c = 0; c < k; ++c) {
        const double v = M[i + (int64_t)c * n];
        if (c < m) {
#pragma unroll
          for (int j = 0; j < MAXM; ++j)
            if (j < m) xo[j] += v * C[c + j * ldc];
        } else {
#pragma unroll
          for (int j = 0; j < MAXM; ++j)
            if (j < m) po[j] += v * C[c + j * ldc];
        }
      }
#pragma unroll
      for (int j = 0; j < MAXM; ++j)
        if (j < m) {
          M[i + (int64_t)j * n] = xo[j] + po[j];
          if (k > m) M[i + (int64_t)(2 * m + j) * n] = po[j];
        }
    }
  }
}

// out[c] = sum_i X[i, c] Y[i, c] in two steps: CD_NB blocks per column (blockIdx.y = column)
// leave one partial each, with four loads in flight per thread; then one thread per column
// adds its CD_NB partials in order
constexpr int CD_NB = 64;
__global__ void __launch_bounds__(256) k_coldots_part(int64_t n, const double *__restrict__ X,
                                                      const double *__restrict__ Y, double *__restrict__ part) {
  __shared__ double sh[4];
  const int64_t c = blockIdx.y;
  const double *x = X + c * n, *y = Y + c * n;
  const int64_t stride = (int64_t)CD_NB * 256;
  double a0 = 0.0, a1 = 0.0, a2 = 0.0, a3 = 0.0;
  int64_t i = blockIdx.x * 256 + threadIdx.x;
  for (; i + 3 * stride < n; i += 4 * stride) {
    a0 += x[i] * y[i];
    a1 += x[i + stride] * y[i + stride];
    a2 += x[i + 2 * stride] * y[i + 2 * stride];
    a3 += x[i + 3 * stride] * y[i + 3 * stride];
  }
  for (; i < n; i += stride) a0 += x[i] * y[i];
  double acc = (a0 + a1) + (a2 + a3);
  for (int off = 32; off > 0; off >>= 1) acc += __shfl_xor(acc, off, 64);
  if ((threadIdx.x & 63) == 0) sh[threadIdx.x >> 6] = acc;
  __syncthreads();
  if (threadIdx.x == 0) part[c * CD_NB + blockIdx.x] = ((sh[0] + sh[1]) + sh[2]) + sh[3];
}

__global__ void k_coldots_sum(int cols, const double *__restrict__ part, double *__restrict__ out) {
  const int c = blockIdx.x * blockDim.x + threadIdx.x;
  if (c >= cols) return;
  double t = 0.0;
  for (int b = 0; b < CD_NB; ++b) t += part[(int64_t)c * CD_NB + b];
  out[c] = t;
}

// partial Gram matrices G = S^T T over row tiles (S, T column-major n x k, k <= GK): 32 rows
// of both in LDS per step, thread p accumulates pairs p, p + 1024, ...; part[block][j*k + i]
constexpr int GK = 96, GR = 32, GPER = (GK * GK + 1023) / 1024;
__global__ void __launch_bounds__(1024) k_gram(int64_t n, int k, const double *__restrict__ S,
                                               const double *__restrict__ T, double *__restrict__ part) {
  __shared__ double sS[GR][GK];
  __shared__ double sT[GR][GK];
  const int npair = k * k;
  double acc[GPER];
#pragma unroll
  for (int q = 0; q < GPER; ++q) acc[q] = 0.0;
  for (int64_t r0 = (int64_t)blockIdx.x * GR; r0 < n; r0 += (int64_t)gridDim.x * GR) {
    for (int e = threadIdx.x; e < GR * k; e += 1024) {
      const int rr = e % GR, cc = e / GR;
      const int64_t r = r0 + rr;
      sS[rr][cc] = r < n ? S[(int64_t)cc * n + r] : 0.0;
      sT[rr][cc] = r < n ? T[(int64_t)cc * n + r] : 0.0;
    }
    __syncthreads();
#pragma unroll
    for (int q = 0; q < GPER; ++q) {
      const int p = threadIdx.x + q * 1024;
      if (p < npair) {
        const int i = p % k, j = p / k;
        double t = 0.0;
#pragma unroll 8
        for (int rr = 0; rr < GR; ++rr) t += sS[rr][i] * sT[rr][j];
        acc[q] += t;
      }
    }
    __syncthreads();
  }
#pragma unroll
  for (int q = 0; q < GPER; ++q) {
    const int p = threadIdx.x + q * 1024;
    if (p < npair) part[(int64_t)p * gridDim.x + blockIdx.x] = acc[q];
  }
}

// G[p] = sum over the blocks of k_gram of their partials of entry p (part[p * nblocks + b]):
// one wavefront per entry, lanes over the blocks, then a fixed shuffle tree
__global__ void __launch_bounds__(256) k_gram_sum(int npair, int nblocks, const double *__restrict__ part,
                                                  double *__restrict__ G) {
  const int p = (int)((blockIdx.x * (int64_t)blockDim.x + threadIdx.x) >> 6);
  const int lane = threadIdx.x & 63;
  if (p >= npair) return;  // whole wavefronts
  double t = 0.0;
  for (int b = lane; b < nblocks; b += 64) t += part[(int64_t)p * nblocks + b];
#pragma unroll
  for (int off = 32; off > 0; off >>= 1) t += __shfl_xor(t, off, 64);
  if (lane == 0) G[p] = t;
}

// symmetric-definite pencil GA z = theta GB z (k x k, column-major, upper parts used) on the
// host: GB = L L^T, C = L^-1 GA L^-T, cyclic Jacobi on C, z = L^-T y; ascending theta,
// Z column-major k x k.  Returns false when GB is not positive definite.
bool small_pencil(int k, const std::vector<double> &GA, const std::vector<double> &GB, std::vector<double> &theta,
                  std::vector<double> &Z) {
  std::vector<double> Lm((size_t)k * k, 0.0), C((size_t)k * k), V((size_t)k * k, 0.0);
  auto a = [&](int i, int j) { return GA[(size_t)std::max(i, j) * k + std::min(i, j)]; };  // upper of column-major
  auto b = [&](int i, int j) { return GB[(size_t)std::max(i, j) * k + std::min(i, j)]; };
  for (int j = 0; j < k; ++j) {  // Cholesky, L row-major Lm[i*k + j]
    double d = b(j, j);
    for (int q = 0; q < j; ++q) d -= Lm[(size_t)j * k + q] * Lm[(size_t)j * k + q];
    if (!(d > 0.0)) return false;
    const double l = std::sqrt(d);
    Lm[(size_t)j * k + j] = l;
    for (int i = j + 1; i < k; ++i) {
      double t = b(i, j);
      for (int q = 0; q < j; ++q) t -= Lm[(size_t)i * k + q] * Lm[(size_t)j * k + q];
      Lm[(size_t)i * k + j] = t / l;
    }
  }
  // T = L^-1 GA (forward solves on the columns), C = T L^-T
  std::vector<double> T((size_t)k * k);
  for (int c = 0; c < k; ++c)
    for (int i = 0; i < k; ++i) {
      double t = a(i, c);
      for (int q = 0; q < i; ++q) t -= Lm[(size_t)i * k + q] * T[(size_t)q * k + c];
      T[(size_t)i * k + c] = t / Lm[(size_t)i * k + i];
    }
  for (int r = 0; r < k; ++r)
    for (int j = 0; j < k; ++j) {
      double t = T[(size_t)r * k + j];
      for (int q = 0; q < j; ++q) t -= C[(size_t)r * k + q] * Lm[(size_t)j * k + q];
      C[(size_t)r * k + j] = t / Lm[(size_t)j * k + j];
    }
  for (int i = 0; i < k; ++i)
    for (int j = i + 1; j < k; ++j) C[(size_t)i * k + j] = C[(size_t)j * k + i] = 0.5 * (C[(size_t)i * k + j] + C[(size_t)j * k + i]);
  for (int i = 0; i < k; ++i) V[(size_t)i * k + i] = 1.0;
  for (int sweep = 0; sweep < 100; ++sweep) {
    double off = 0.0, tot = 0.0;
    for (int i = 0; i < k; ++i)
      for (int j = 0; j < k; ++j) {
        tot += C[(size_t)i * k + j] * C[(size_t)i * k + j];
        if (i != j) off += C[(size_t)i * k + j] * C[(size_t)i * k + j];
      }
    if (off <= 1e-30 * tot) break;
    for (int p = 0; p < k - 1; ++p)
      for (int q = p + 1; q < k; ++q) {
        const double apq = C[(size_t)p * k + q];
        if (apq == 0.0) continue;
        const double tau = (C[(size_t)q * k + q] - C[(size_t)p * k + p]) / (2.0 * apq);
        const double t = (tau >= 0 ? 1.0 : -1.0) / (std::fabs(tau) + std::sqrt(1.0 + tau * tau));
        const double cs = 1.0 / std::sqrt(1.0 + t * t), sn = t * cs;
        for (int r = 0; r < k; ++r) {  // columns p, q
          const double xp = C[(size_t)r * k + p], xq = C[(size_t)r * k + q];
          C[(size_t)r * k + p] = cs * xp - sn * xq;
          C[(size_t)r * k + q] = sn * xp + cs * xq;
        }
        for (int r = 0; r < k; ++r) {  // rows p, q
          const double xp = C[(size_t)p * k + r], xq = C[(size_t)q * k + r];
          C[(size_t)p * k + r] = cs * xp - sn * xq;
          C[(size_t)q * k + r] = sn * xp + cs * xq;
        }
        for (int r = 0; r < k; ++r) {
          const double vp = V[(size_t)r * k + p], vq = V[(size_t)r * k + q];
          V[(size_t)r * k + p] = cs * vp - sn * vq;
          V[(size_t)r * k + q] = sn * vp + cs * vq;
        }
      }
  }
  std::vector<int> ord(k);
  for (int i = 0; i < k; ++i) ord[i] = i;
  std::sort(ord.begin(), ord.end(), [&](int x, int y) { return C[(size_t)x * k + x] < C[(size_t)y * k + y]; });
  theta.assign(k, 0.0);
  Z.assign((size_t)k * k, 0.0);
  for (int c = 0; c < k; ++c) {
    const int o = ord[c];
    theta[c] = C[(size_t)o * k + o];
    for (int i = k - 1; i >= 0; --i) {  // z = L^-T y (back substitution)
      double t = V[(size_t)i * k + o];
      for (int q = i + 1; q < k; ++q) t -= Lm[(size_t)q * k + i] * Z[(size_t)c * k + q];
      Z[(size_t)c * k + i] = t / Lm[(size_t)i * k + i];
    }
  }
  return true;
}

struct Handle {
  rocblas_handle h = nullptr;
  ~Handle() {
    if (h) (void)rocblas_destroy_handle(h);
  }
};

unsigned grid_for(int64_t work) { return pf_blocks(work, 256, 65536); }

}  // namespace

int pf_eig_device(pf_ctx *ctx, int64_t n, const int64_t *row_ptr, const int32_t *col_idx, const double *A,
                  const double *B, int64_t nnz, const std::vector<int32_t> &fixed_ids, int32_t nev, double tol,
                  int32_t maxit, double *values, double *vectors, int32_t *count, EigResult &res) {
  res = EigResult();
  hipStream_t st = ctx->stream;
  // reduced index of every DOF (-1: removed)
  std::vector<int32_t> h_red((size_t)n, 0);
  for (int32_t id : fixed_ids)
    if (id >= 1 && id <= n) h_red[id - 1] = -1;
  int64_t nf = 0;
  for (int64_t i = 0; i < n; ++i)
    if (h_red[i] >= 0) h_red[i] = (int32_t)nf++;
  DevBuf<int32_t> red;
  PF_CHECK(red.ensure((size_t)n + 1));
  PF_HIP(hipMemcpyAsync(red.p, h_red.data(), sizeof(int32_t) * n, hipMemcpyHostToDevice, st));
  Handle H;
  PF_RB(rocblas_create_handle(&H.h));
  PF_RB(rocblas_set_stream(H.h, st));
  PF_HIP(hipEventRecord(ctx->ev0, st));
  const double avg_row = n > 0 ? (double)nnz / (double)n : 1.0;

  if (nev <= 0) {  // ---------------------------------------------------- every eigenpair
    if (nf > 16384) {
      pf_set_error("eig: %lld free DOFs is too many for the dense solve (at most 16384); ask for nev pairs",
                   (long long)nf);
      return PF_INVALID;
    }
    *count = (int32_t)nf;
    if (nf == 0) return PF_OK;
    DevBuf<double> Da, Db, W, E;
    DevBuf<rocblas_int> info;
    PF_CHECK(Da.ensure((size_t)(nf * nf)));
    PF_CHECK(Db.ensure((size_t)(nf * nf)));
    PF_CHECK(W.ensure((size_t)nf));
    PF_CHECK(E.ensure((size_t)nf));
    PF_CHECK(info.ensure(1));
    PF_HIP(hipMemsetAsync(Da.p, 0, sizeof(double) * nf * nf, st));
    PF_HIP(hipMemsetAsync(Db.p, 0, sizeof(double) * nf * nf, st));
    k_dense_scatter<<<pf_blocks(n, 256, INT32_MAX), 256, 0, st>>>(n, row_ptr, col_idx, A, red.p, nf, Da.p);
    k_dense_scatter<<<pf_blocks(n, 256, INT32_MAX), 256, 0, st>>>(n, row_ptr, col_idx, B, red.p, nf, Db.p);
    PF_HIP(hipGetLastError());
    PF_RB(rocsolver_dsygvd(H.h, rocblas_eform_ax, rocblas_evect_original, rocblas_fill_upper, (rocblas_int)nf, Da.p,
                           (rocblas_int)nf, Db.p, (rocblas_int)nf, W.p, E.p, info.p));
    rocblas_int hinfo = 0;
    PF_HIP(hipMemcpyAsync(&hinfo, info.p, sizeof(hinfo), hipMemcpyDeviceToHost, st));
    if (values) PF_HIP(hipMemcpyAsync(values, W.p, sizeof(double) * nf, hipMemcpyDeviceToHost, st));
    if (vectors) {
      DevBuf<double> out;
      PF_CHECK(out.ensure((size_t)(n * nf)));
      k_expand<<<grid_for(n * nf), 256, 0, st>>>(n, nf, nf, red.p, Da.p, out.p);
      PF_HIP(hipGetLastError());
      PF_HIP(hipMemcpyAsync(vectors, out.p, sizeof(double) * n * nf, hipMemcpyDeviceToHost, st));
      PF_HIP(hipStreamSynchronize(st));
    }
    PF_HIP(hipStreamSynchronize(st));
    if (hinfo != 0) {
      pf_set_error("eig: dsygvd info %d (B not positive definite on the free DOFs, or no convergence)", (int)hinfo);
      res.status = PF_BREAKDOWN_C2C3;
      return res.status;
    }
    res.status = PF_OK;
    res.nconv = (int32_t)nf;
    PF_HIP(hipEventRecord(ctx->ev1, st));
    PF_HIP(hipEventSynchronize(ctx->ev1));
    float ms = 0.f;
    PF_HIP(hipEventElapsedTime(&ms, ctx->ev0, ctx->ev1));
    res.ms = ms;
    return PF_OK;
  }

  // ---------------------------------------------------------------------------- LOBPCG
  const int64_t m = nev;
  if (m > nf || 3 * m > GK) {
    pf_set_error("eig: nev = %d must be at most the free DOF count (%lld) and %d", nev, (long long)nf, GK / 3);
    return PF_INVALID;
  }
  *count = nev;
  const int64_t K = 3 * m;
  DevBuf<double> S, AS, BS, R, dinv, GA, GB, th, E, sc;
  DevBuf<rocblas_int> info;
  PF_CHECK(S.ensure((size_t)(n * K)));
  PF_CHECK(AS.ensure((size_t)(n * K)));
  PF_CHECK(BS.ensure((size_t)(n * K)));
  PF_CHECK(R.ensure((size_t)(n * m)));
  PF_CHECK(dinv.ensure((size_t)n));
  PF_CHECK(GA.ensure((size_t)(K * K)));
  PF_CHECK(GB.ensure((size_t)(K * K)));
  PF_CHECK(th.ensure((size_t)K));
  PF_CHECK(E.ensure((size_t)K));
  PF_CHECK(sc.ensure((size_t)K));
  PF_CHECK(info.ensure(1));
  // one workgroup per CU for the partial Gram matrices (a wavefront per entry sums them)
  const unsigned gram_blocks = (unsigned)std::min<int64_t>((n + GR - 1) / GR, (int64_t)std::max(ctx->num_cus, 1));
  DevBuf<double> gpart;
  PF_CHECK(gpart.ensure((size_t)gram_blocks * K * K + 1));
  double *X = S.p, *Wb = S.p + n * m;  // P: columns 2m .. 3m
  double *AX = AS.p, *AW = AS.p + n * m;
  double *BX = BS.p, *BW = BS.p + n * m;
  k_dinv<<<pf_blocks(n, 256, INT32_MAX), 256, 0, st>>>(n, row_ptr, col_idx, A, red.p, dinv.p);
  k_init_block<<<grid_for(n * m), 256, 0, st>>>(n, m, red.p, X);
  PF_HIP(hipGetLastError());
  // every column of a block in one launch (the matrix streamed once, not once per column)
  auto spmv_block = [&](const double *M, const double *in, double *out, int64_t cols) -> int {
    return pf_spmm_crs_launch(ctx, n, cols, row_ptr, col_idx, M, in, out, avg_row);
  };
  // out[0 .. cols) = column dots of X and Y (column-major, leading dimension n)
  DevBuf<double> cdpart;
  PF_CHECK(cdpart.ensure((size_t)K * CD_NB));
  auto coldots = [&](const double *Xc, const double *Yc, int64_t cols, double *out) -> int {
    k_coldots_part<<<dim3(CD_NB, (unsigned)cols), 256, 0, st>>>(n, Xc, Yc, cdpart.p);
    k_coldots_sum<<<(unsigned)((cols + 63) / 64), 64, 0, st>>>((int)cols, cdpart.p, out);
    PF_HIP(hipGetLastError());
    return PF_OK;
  };
  // Rayleigh-Ritz on the first k columns of S; keeps the lowest m in X (and, with k > m, the
  // W/P part of the combination in P).  Returns 1 when dsygvd fails.
  std::vector<double> h_th((size_t)K);
  auto rayleigh_ritz = [&](int64_t k, int *failed) -> int {
    // Gram matrices S^T (A S) and S^T (B S) (a transposed tall-skinny product, which the
    // library GEMM runs on a handful of workgroups: 21 ms for k = 30, n = 134k; this: one pass)
    k_gram<<<gram_blocks, 1024, 0, st>>>(n, (int)k, S.p, AS.p, gpart.p);
    k_gram_sum<<<(unsigned)((k * k * 64 + 255) / 256), 256, 0, st>>>((int)(k * k), (int)gram_blocks, gpart.p, GA.p);
    k_gram<<<gram_blocks, 1024, 0, st>>>(n, (int)k, S.p, BS.p, gpart.p);
    k_gram_sum<<<(unsigned)((k * k * 64 + 255) / 256), 256, 0, st>>>((int)(k * k), (int)gram_blocks, gpart.p, GB.p);
    PF_HIP(hipGetLastError());
    std::vector<double> hA((size_t)k * k), hB((size_t)k * k), hth, hZ;
    PF_HIP(hipMemcpyAsync(hA.data(), GA.p, sizeof(double) * k * k, hipMemcpyDeviceToHost, st));
    PF_HIP(hipMemcpyAsync(hB.data(), GB.p, sizeof(double) * k * k, hipMemcpyDeviceToHost, st));
    PF_HIP(hipStreamSynchronize(st));
    *failed = !small_pencil((int)k, hA, hB, hth, hZ);
    if (*failed) return PF_OK;
    for (int64_t c = 0; c < k; ++c) h_th[c] = hth[c];
    PF_HIP(hipMemcpyAsync(GA.p, hZ.data(), sizeof(double) * k * k, hipMemcpyHostToDevice, st));
    // X = S C, AX = AS C, BX = BS C for the lowest m vectors C = GA[:, 0:m], and with k > m
    // P = [W P] C[m:k, :] (the same for AS, BS): one pass over the block, in place
    const unsigned rg = pf_blocks(n, 256, (int64_t)ctx->num_cus * 8);
    if (m <= 8)
      k_rr_update<8><<<rg, 256, 0, st>>>(n, (int)k, (int)m, GA.p, (int)k, S.p, AS.p, BS.p);
    else if (m <= 16)
      k_rr_update<16><<<rg, 256, 0, st>>>(n, (int)k, (int)m, GA.p, (int)k, S.p, AS.p, BS.p);
    else
      k_rr_update<32><<<rg, 256, 0, st>>>(n, (int)k, (int)m, GA.p, (int)k, S.p, AS.p, BS.p);
    PF_HIP(hipGetLastError());
    return PF_OK;
  };

  PF_CHECK(spmv_block(A, X, AX, m));
  PF_CHECK(spmv_block(B, X, BX, m));
  int failed = 0;
  PF_CHECK(rayleigh_ritz(m, &failed));
  if (failed) {
    pf_set_error("eig: B is not positive definite on the start block");
    res.status = PF_BREAKDOWN_C2C3;
    return res.status;
  }
  bool have_p = false;
  std::vector<double> rn((size_t)m), bn((size_t)m), wn((size_t)(2 * m));
  int it = 0;
  res.status = PF_NOT_CONVERGED;
  for (it = 1; it <= maxit; ++it) {
    PF_HIP(hipMemcpyAsync(th.p, h_th.data(), sizeof(double) * m, hipMemcpyHostToDevice, st));
    k_residual_block<<<grid_for(n * m), 256, 0, st>>>(n, m, AX, BX, th.p, dinv.p, red.p, R.p, Wb);
    PF_HIP(hipGetLastError());
    PF_CHECK(coldots(R.p, R.p, m, sc.p));
    PF_CHECK(coldots(BX, BX, m, sc.p + m));
    PF_HIP(hipGetLastError());
    std::vector<double> nrm((size_t)(2 * m));
    PF_HIP(hipMemcpyAsync(nrm.data(), sc.p, sizeof(double) * 2 * m, hipMemcpyDeviceToHost, st));
    PF_HIP(hipStreamSynchronize(st));
    for (int64_t c = 0; c < m; ++c) {
      rn[c] = std::sqrt(nrm[c]);
      bn[c] = std::sqrt(nrm[m + c]);
    }
    // |A x - l B x| relative to the block's largest |l| |B x| (rigid-body modes have l = 0);
    // pairs below tol are soft-locked: their W and P columns leave the search space, which
    // keeps the projected B definite once some pairs have converged
    double lmax = 0.0, worst = 0.0;
    for (int64_t c = 0; c < m; ++c) lmax = std::max(lmax, std::fabs(h_th[c]));
    std::vector<int64_t> act;
    for (int64_t c = 0; c < m; ++c) {
      const double rc = rn[c] / std::max(lmax * bn[c], 1e-300);
      worst = std::max(worst, rc);
      if (!(rc < tol)) act.push_back(c);
    }
    res.max_res = worst;
    if (act.empty()) {
      res.status = PF_OK;
      break;
    }
    const int64_t na = (int64_t)act.size();
    // compact [X | W_act | P_act]: every copy moves a column to a lower index, in order
    for (int64_t q = 0; q < na; ++q)
      if (act[q] != q) k_copy<<<grid_for(n), 256, 0, st>>>(n, Wb + act[q] * n, Wb + q * n);
    if (have_p)
      for (int64_t q = 0; q < na; ++q) {
        const int64_t src = 2 * m + act[q], dst = m + na + q;
        if (src != dst) {
          k_copy<<<grid_for(n), 256, 0, st>>>(n, S.p + src * n, S.p + dst * n);
          k_copy<<<grid_for(n), 256, 0, st>>>(n, AS.p + src * n, AS.p + dst * n);
          k_copy<<<grid_for(n), 256, 0, st>>>(n, BS.p + src * n, BS.p + dst * n);
        }
      }
    PF_HIP(hipGetLastError());
    PF_CHECK(spmv_block(A, Wb, AW, na));
    PF_CHECK(spmv_block(B, Wb, BW, na));
    // unit B-norm columns of W (and P)
    const int64_t ncols = have_p ? 2 * na : na;
    PF_CHECK(coldots(S.p + m * n, BS.p + m * n, ncols, sc.p));
    PF_HIP(hipGetLastError());
    PF_HIP(hipMemcpyAsync(wn.data(), sc.p, sizeof(double) * ncols, hipMemcpyDeviceToHost, st));
    PF_HIP(hipStreamSynchronize(st));
    std::vector<double> f((size_t)ncols);
    for (int64_t c = 0; c < ncols; ++c) f[c] = wn[c] > 0.0 ? 1.0 / std::sqrt(wn[c]) : 0.0;
    PF_HIP(hipMemcpyAsync(sc.p, f.data(), sizeof(double) * ncols, hipMemcpyHostToDevice, st));
    k_scale_cols<<<grid_for(n * ncols), 256, 0, st>>>(n, m, ncols, sc.p, S.p, AS.p, BS.p);
    PF_HIP(hipGetLastError());
    PF_CHECK(rayleigh_ritz(m + ncols, &failed));
    if (failed && have_p) PF_CHECK(rayleigh_ritz(m + na, &failed));  // projected B lost definiteness: drop P
    if (failed) {
      pf_set_error("eig: LOBPCG Rayleigh-Ritz failed (projected B not positive definite) at iteration %d", it);
      res.status = PF_BREAKDOWN_C2C3;
      break;
    }
    have_p = true;
  }
  res.iterations = std::min(it, maxit);
  res.nconv = res.status == PF_OK ? nev : 0;
  if (values) std::memcpy(values, h_th.data(), sizeof(double) * m);
  if (vectors) PF_HIP(hipMemcpyAsync(vectors, X, sizeof(double) * n * m, hipMemcpyDeviceToHost, st));
  PF_HIP(hipEventRecord(ctx->ev1, st));
  PF_HIP(hipEventSynchronize(ctx->ev1));
  float ms = 0.f;
  PF_HIP(hipEventElapsedTime(&ms, ctx->ev0, ctx->ev1));
  res.ms = ms;
  if (res.status == PF_NOT_CONVERGED)
    pf_set_error("eig: LOBPCG not converged after %d iterations (worst relative residual %.3e)", res.iterations,
                 res.max_res);
  return PF_OK;
}
