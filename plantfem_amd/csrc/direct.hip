// direct.hip -- direct solve of the eliminated system: banded Cholesky (L L^T) per mesh.
//
// The reference's solve() has a direct branch, solve(algorithm="GaussJordan")
// (FEMSolverClass.f90:1795-1798) -> gauss_jordan_crs (FEMSolverClass.f90:2413-2460), whose body
// stops with "Under construction" after the first pivot test ("A is singular" on a zero pivot,
// :2455).  This file gives that branch a working direct method for the symmetric positive
// definite systems the elimination produces (FEMSolverClass.f90:1707-1756 keeps K symmetric:
// fixed rows and columns are zeroed with a unit diagonal).
//
// Per mesh (one mesh = the whole system, or every mesh of a block-diagonal batch):
//  * ordering (host, once per sparsity pattern): the node (DOF=3 block) or DOF graph is
//    renumbered by the candidate of smallest bandwidth among the natural order, reverse
//    Cuthill-McKee and lexicographic coordinate sorts along each axis;
//  * band storage (LAPACK dpbtrf 'L' layout): entry (r, c), 0 <= r - c <= b, of the permuted
//    matrix at band[c*ld + r - c], ld >= b + 1; the order is padded to a multiple of 32 with
//    unit diagonal rows;
//  * factorization: one workgroup per mesh from a largest-first queue (persistent kernel),
//    left-looking by blocks of 32 columns: the block's (b+32) x 32 rows are updated by the
//    previous b columns with FP64 MFMA 16x16x4 tiles, the two 16 x 16 diagonal tiles are
//    factored with cross-lane operations and the tiles below are multiplied by their inverse
//    transposes;
//  * triangular solves: blocks of 16, the diagonal solve in one wave, the off-diagonal
//    update with one thread per row (forward) or column (backward).
// Fixed-precision iterative refinement on the true residual b - A x (the CRS values, in
// FP64) follows as requested.
#include <algorithm>
#include <atomic>
#include <chrono>
#include <cmath>
#include <cstring>
#include <numeric>
#include <thread>
#include <vector>

#include "pf_internal.hpp"

namespace {

typedef double d4 __attribute__((ext_vector_type(4)));

constexpr int DBT = 512;   // threads per workgroup of the factor/solve kernel
constexpr int MAXT = 2;    // row tiles a wave holds at once (each 16 rows x 32 columns)

struct DMesh {
  int64_t d0;    // first DOF of the mesh in the system
  int64_t boff;  // first band entry
  int64_t woff;  // first entry of the mesh's permutation and work vector
  int32_t n;     // DOFs
  int32_t npad;  // n rounded up to a multiple of 32
  int32_t b;     // half bandwidth of the permuted matrix (DOFs)
  int32_t ld;    // leading dimension of the band (>= b + 1, even)
};

struct DArgs {
  const DMesh *mesh;
  const int32_t *order;  // queue order (largest first)
  int32_t *queue;
  int32_t nmesh;
  int32_t mode;          // 0: factor, then solve x = A^-1 rhs; 1: solve only, x += A^-1 rhs
  double *band;
  const int32_t *perm;   // band index q of mesh m -> local DOF (perm[woff + q], q < n)
  double *w;             // work vector (npad per mesh)
  const double *rhs;
  double *x;
  int32_t *status;       // per mesh
  int64_t *t_raw;        // per mesh: wall-clock start, end of the factor+solve
};

__device__ inline double shfl(double v, int src) { return __shfl(v, src, 64); }

// The wave index within the workgroup as a scalar: the compiler cannot tell that
// threadIdx.x / 64 is the same in every lane of a wave, and without this every tile index,
// chunk bound and loop counter derived from it would be kept per lane, with masked branches
// and full waits for outstanding loads where they join.
__device__ inline int wave_id() { return __builtin_amdgcn_readfirstlane((int)(threadIdx.x >> 6)); }

// The lane index, re-read inside a loop so that the compiler rebuilds the lane masks derived
// from it there instead of hoisting hundreds of them out of the loop and spilling them.
__device__ inline int lane_id() {
  int l = (int)(threadIdx.x & 63);
  asm volatile("" : "+v"(l));
  return l;
}

// Cholesky of the 16 x 16 diagonal tile held in the FP64 MFMA accumulator layout (lane l,
// element v: row 4*v + l/16, column l%16).  Lower part becomes L11; bad is set on a
// non-positive pivot.
__device__ inline void chol16(d4 &a, int q, int jl, bool &bad) {
#pragma unroll
  for (int k = 0; k < 16; ++k) {
    const double piv = shfl(a[k >> 2], (k & 3) * 16 + k);
    bad = bad || !(piv > 0.0);
    const double s = sqrt(piv);
    const double rs = 1.0 / s;
    if (jl == k) {
#pragma unroll
      for (int v = 0; v < 4; ++v) {
        const int i = 4 * v + q;
        if (i > k) a[v] *= rs;
        else if (i == k) a[v] = s;
      }
    }
    // L[i][k] of this lane's rows (same lane group, column k) and L[jl][k]
    double lik[4];
#pragma unroll
    for (int v = 0; v < 4; ++v) lik[v] = shfl(a[v], 16 * q + k);
    const int src = 16 * (jl & 3) + k;
    const double e0 = shfl(a[0], src), e1 = shfl(a[1], src), e2 = shfl(a[2], src), e3 = shfl(a[3], src);
    const int sel = jl >> 2;
    const double ljk = sel == 0 ? e0 : (sel == 1 ? e1 : (sel == 2 ? e2 : e3));
    if (jl > k) {
#pragma unroll
      for (int v = 0; v < 4; ++v)
        if (4 * v + q >= jl) a[v] -= lik[v] * ljk;
    }
  }
}

// acc += A B for the 16x4 fragment A (lane l holds A[l%16][l/16]) and the 4x16 fragment B
// (lane l holds B[l/16][l%16]); accumulator layout as chol16 (the register layout of
// v_mfma_f64_16x16x4f64).  MF: the matrix core; else the
// same product from cross-lane reads and FMAs.
template <bool MF>
__device__ inline void tile_fma(d4 &acc, double af, double bf, int q, int jl) {
  if constexpr (MF) {
    acc = __builtin_amdgcn_mfma_f64_16x16x4f64(af, bf, acc, 0, 0, 0);
  } else {
#pragma unroll
    for (int k = 0; k < 4; ++k) {
      const double bk = shfl(bf, 16 * k + jl);
#pragma unroll
      for (int v = 0; v < 4; ++v) acc[v] += shfl(af, 16 * k + 4 * v + q) * bk;
    }
  }
}

// Geometry of one block of 32 columns j0 .. j0+31 and of a wave's tiles in one pass: row
// tiles t (rows R = j0 + 16 t .. R+15, t < T) of the rows j0 .. j0+31+b, column tiles
// ct = 0 (columns j0 .. j0+15) and ct = 1 (j0+16 .. j0+31); the wave holds tiles
// tb, tb + W, ..., tb + (MAXT-1) W.
struct BlockGeom {
  int j0, b, npad, T, tb, W;
  int64_t ld;
};

// Fragments of U consecutive 4-column chunks c0, c0+4, ... of the previous columns: A
// fragments L[R + l%16][c + l/16] of the wave's tiles and B fragments L[j0 + 16 ct + l%16]
// [c + l/16] of both column tiles.  This is the form for the first chunks of a tile (where
// some lanes' entries lie outside the band) and the last rows: every load has an index
// clamped into the band, so the loads are unconditional and stay in flight together, and the
// entries outside the band are replaced by zero where they are used.  Chunks below cend.
template <int U>
struct Frags {
  double b[U][2];
  double a[U][MAXT];
};

template <int U>
__device__ inline void load_frags(Frags<U> &f, const double *__restrict__ band, const BlockGeom &g, int c0,
                                  int cend, int q, int jl) {
#pragma unroll
  for (int u = 0; u < U; ++u) {
    const int c = c0 + 4 * u;
    if (c >= cend) continue;  // wave-uniform
    const int64_t ck = c + q;
#pragma unroll
    for (int ct = 0; ct < 2; ++ct) {
      const int db = g.j0 + 16 * ct + jl - (int)ck;  // >= 1
      f.b[u][ct] = band[ck * g.ld + min(db, g.b)];
    }
#pragma unroll
    for (int i = 0; i < MAXT; ++i) {
      const int da = g.j0 + 16 * (g.tb + g.W * i) + jl - (int)ck;  // >= 1
      f.a[u][i] = band[ck * g.ld + min(da, g.b)];
    }
  }
}

// acc[i][ct] += A (-B_ct) over the group's chunks (tile 0 has no column tile 1)
template <bool MF, int U>
__device__ inline void mfma_frags(const Frags<U> &f, d4 (*acc)[2], const BlockGeom &g, int c0, int cend, int q,
                                  int jl) {
#pragma unroll
  for (int u = 0; u < U; ++u) {
    const int c = c0 + 4 * u;
    if (c >= cend) continue;  // wave-uniform
    const int ck = c + q;
    double bn[2];
#pragma unroll
    for (int ct = 0; ct < 2; ++ct) bn[ct] = (g.j0 + 16 * ct + jl - ck <= g.b) ? -f.b[u][ct] : 0.0;
#pragma unroll
    for (int i = 0; i < MAXT; ++i) {
      const int t = g.tb + g.W * i;
      const int R = g.j0 + 16 * t;
      if (t < g.T && c + 3 >= R - g.b) {  // wave-uniform
        const int ra = R + jl;
        const double af = (ra < g.npad && ra - ck <= g.b) ? f.a[u][i] : 0.0;
        tile_fma<MF>(acc[i][0], af, bn[0], q, jl);
        if (t >= 1) tile_fma<MF>(acc[i][1], af, bn[1], q, jl);
      }
    }
  }
}

// The steady part of a pass: chunks cs <= c < j0, where every lane's entry of each of the
// wave's nt existing tiles and of both column tiles lies inside the band and the tiles' rows
// end before the padding.  Branch-free: per tile a pointer that advances by 4 (ld - 1)
// doubles per chunk, two groups of chunks in flight, each A fragment used by both column
// tiles.
template <bool MF, int U>
__device__ inline void steady_update(d4 (*acc)[2], const double *__restrict__ band, const BlockGeom &g, int cs,
                                     int nt, int q, int jl) {
  const int64_t step = g.ld - 1;
  const double *pa[MAXT];
#pragma unroll
  for (int i = 0; i < MAXT; ++i) pa[i] = band + (int64_t)(cs + q) * step + g.j0 + 16 * (g.tb + g.W * i) + jl;
  const double *pb = band + (int64_t)(cs + q) * step + g.j0 + jl;
  double a0[U][MAXT], a1[U][MAXT], b0[U][2], b1[U][2];
  auto load = [&](double (&a)[U][MAXT], double (&bb)[U][2], int c) {
    const int64_t off = (int64_t)(c - cs) * step;
#pragma unroll
    for (int u = 0; u < U; ++u) {
      if (c + 4 * u >= g.j0) continue;  // wave-uniform
#pragma unroll
      for (int i = 0; i < MAXT; ++i)
        if (i < nt) a[u][i] = pa[i][off + (int64_t)(4 * u) * step];
#pragma unroll
      for (int ct = 0; ct < 2; ++ct) bb[u][ct] = pb[off + (int64_t)(4 * u) * step + 16 * ct];
    }
  };
  auto prod = [&](const double (&a)[U][MAXT], const double (&bb)[U][2], int c) {
#pragma unroll
    for (int u = 0; u < U; ++u) {
      if (c + 4 * u >= g.j0) continue;  // wave-uniform
      const double bn0 = -bb[u][0], bn1 = -bb[u][1];
#pragma unroll
      for (int i = 0; i < MAXT; ++i)
        if (i < nt) {
          tile_fma<MF>(acc[i][0], a[u][i], bn0, q, jl);
          if (g.tb + g.W * i >= 1) tile_fma<MF>(acc[i][1], a[u][i], bn1, q, jl);
        }
    }
  };
  int c = cs;
  if (c < g.j0) load(a0, b0, c);
  for (; c < g.j0; c += 8 * U) {
    const int c1 = c + 4 * U, c2 = c1 + 4 * U;
    if (c1 < g.j0) load(a1, b1, c1);
    prod(a0, b0, c);
    if (c2 < g.j0) load(a0, b0, c2);
    if (c1 < g.j0) prod(a1, b1, c1);
  }
}

// Diagonal tile of column tile ct (held by the calling wave in accumulator layout): Cholesky
// factor L_d to the band (columns col0 .. col0+15) and LDS sL, then L_d^-T (the solution X of
// X L_d^T = I, column by column across the lanes) to sLi; *sflag on a non-positive pivot.
template <bool MF>
__device__ inline void diag_block(d4 &a, double *__restrict__ band, int64_t ld, int b, int col0, double *sL,
                                  double *sLi, int32_t *sflag) {
  const int ln = lane_id();
  const int q = ln >> 4, jl = ln & 15;
  bool bad = false;
  chol16(a, q, jl, bad);
#pragma unroll
  for (int v = 0; v < 4; ++v) {
    const int i = 4 * v + q;
    sL[i * 16 + jl] = a[v];
    const int d = i - jl;
    if (d >= 0 && d <= b) band[(int64_t)(col0 + jl) * ld + d] = a[v];
  }
  double Lrow[16];  // L_d[jl][k]
#pragma unroll
  for (int k = 0; k < 16; ++k) Lrow[k] = sL[jl * 16 + k];
  d4 x;
#pragma unroll
  for (int v = 0; v < 4; ++v) x[v] = (4 * v + q == jl) ? 1.0 : 0.0;
#pragma unroll
  for (int k = 0; k < 16; ++k) {
    if (jl == k) {
      const double rk = 1.0 / Lrow[k];
#pragma unroll
      for (int v = 0; v < 4; ++v) x[v] *= rk;
    }
    double xk[4];
#pragma unroll
    for (int v = 0; v < 4; ++v) xk[v] = shfl(x[v], 16 * q + k);
    if (jl > k) {
#pragma unroll
      for (int v = 0; v < 4; ++v) x[v] -= xk[v] * Lrow[k];
    }
  }
#pragma unroll
  for (int v = 0; v < 4; ++v) sLi[(4 * v + q) * 16 + jl] = x[v];
  if (bad && ln == 0) *sflag = 1;
}

// x = c m^T for 16 x 16 tiles: c in accumulator layout (through the wave's LDS tile myC into
// A fragments), m given by its B fragments mb[cq] = m[jl][4 cq + q]... i.e. B[kk][n] =
// m^T[4 cq + kk][n]; sign -1 subtracts the product from acc instead
template <bool MF>
__device__ inline void tile_times(d4 &acc, const d4 &c, const double *mb, double *myC) {
  const int ln = lane_id();
  const int q = ln >> 4, jl = ln & 15;
#pragma unroll
  for (int v = 0; v < 4; ++v) myC[(4 * v + q) * 16 + jl] = c[v];
  double af[4];
#pragma unroll
  for (int cq = 0; cq < 4; ++cq) af[cq] = myC[jl * 16 + 4 * cq + q];
#pragma unroll
  for (int cq = 0; cq < 4; ++cq) tile_fma<MF>(acc, af[cq], mb[cq], q, jl);
}

// store a tile (accumulator layout) of column tile starting at col0 into the band
__device__ inline void store_tile(const d4 &x, double *__restrict__ band, int64_t ld, int b, int npad, int R,
                                  int col0) {
  const int ln = lane_id();
  const int q = ln >> 4, jl = ln & 15;
#pragma unroll
  for (int v = 0; v < 4; ++v) {
    const int r = R + 4 * v + q;
    const int d = r - (col0 + jl);
    if (r < npad && d >= 0 && d <= b) band[(int64_t)(col0 + jl) * ld + d] = x[v];
  }
}

// Left-looking blocked band Cholesky of one mesh by the whole workgroup, blocks of 32 columns.
// Per block: every wave updates its tiles of rows j0 .. j0+31+b in both column tiles with the
// previous b columns (MFMA; each A fragment feeds both column tiles; loads of the next group
// of chunks overlap the products of the current one).  Then wave 0 factors the first
// diagonal tile (L00, and L00^-T); every wave forms its column-0 tiles L(t,0) = C L00^-T;
// the owner of tile 1 publishes L10; every wave subtracts L(t,0) L10^T from its column-1
// tiles; wave 1 factors the second diagonal tile (L11, L11^-T); every wave forms
// L(t,1) = C L11^-T.  Four barriers per block.
template <int BT, bool MF>
__device__ int factor_mesh(const DMesh &M, double *__restrict__ band, double *sL, double *sLi, double *sC,
                           int32_t *sflag) {
  constexpr int W = BT / 64;
  constexpr int U = 2;   // chunks per group of the steady part
  constexpr int UR = 2;  // chunks per group of the ramp
  const int wv = wave_id();
  const int b = M.b, npad = M.npad;
  const int64_t ld = M.ld;
  double *myC = sC + 256 * wv;
  double *sLi0 = sLi, *sL10 = sLi + 256, *sLi1 = sLi + 512;
  if (threadIdx.x == 0) *sflag = 0;
  __syncthreads();
  for (int j0 = 0; j0 < npad; j0 += 32) {
    const int lane = lane_id();
    const int q = lane >> 4, jl = lane & 15;
    const int rows = min(32 + b, npad - j0);  // rows j0 .. j0+rows-1 hold the block's nonzeros
    const int T = (rows + 15) >> 4;           // >= 2
    const int P = (T + W * MAXT - 1) / (W * MAXT);
    double mb0[4], mb10[4], mb1[4];  // B fragments of L00^-T, -L10^T, L11^-T (set in pass 0)
    for (int p = 0; p < P; ++p) {
      BlockGeom g{j0, b, npad, T, wv + W * MAXT * p, W, ld};
      d4 acc[MAXT][2];
#pragma unroll
      for (int i = 0; i < MAXT; ++i) {
        const int t = g.tb + W * i;
#pragma unroll
        for (int ct = 0; ct < 2; ++ct) {
          const int64_t cc = j0 + 16 * ct + jl;
#pragma unroll
          for (int v = 0; v < 4; ++v) {
            const int r = j0 + 16 * t + 4 * v + q;
            acc[i][ct][v] = band[cc * ld + min(max(r - (int)cc, 0), b)];  // clamped: inside the band
          }
#pragma unroll
          for (int v = 0; v < 4; ++v) {
            const int r = j0 + 16 * t + 4 * v + q;
            const int d = r - (int)cc;
            if (!(t < T && r < npad && d >= 0 && d <= b)) acc[i][ct][v] = 0.0;
          }
        }
      }
      if (g.tb < T) {
        // A[block rows, j0..j0+31] -= L[block rows, c] L[j0..j0+31, c]^T, c < j0: a ramp with
        // predicates up to cs, then the branch-free steady part
        const int clo = max(0, j0 + 16 * g.tb - b) & ~3;
        const int nt = min(MAXT, (T - 1 - g.tb) / W + 1);  // existing tiles of this wave
        const int Rl = j0 + 16 * (g.tb + W * (nt - 1));
        const int cs = Rl + 16 <= npad ? min(j0, max(clo, (max(Rl + 15, j0 + 31) - b + 3) & ~3)) : j0;
        for (int c = clo; c < cs; c += 4 * UR) {
          Frags<UR> f;
          load_frags<UR>(f, band, g, c, cs, q, jl);
          mfma_frags<MF, UR>(f, acc, g, c, cs, q, jl);
        }
        steady_update<MF, U>(acc, band, g, cs, nt, q, jl);
      }
      if (p == 0) {
        if (wv == 0) diag_block<MF>(acc[0][0], band, ld, b, j0, sL, sLi0, sflag);
        __syncthreads();
        if (*sflag) return PF_NOT_SPD;
#pragma unroll
        for (int cq = 0; cq < 4; ++cq) mb0[cq] = sLi0[(4 * cq + q) * 16 + jl];
      }
      // column tile 0 below the diagonal: L(t,0) = C L00^-T; tile 1 publishes L10
#pragma unroll
      for (int i = 0; i < MAXT; ++i) {
        const int t = g.tb + W * i;
        if (t >= 1 && t < T) {
          d4 x = {0.0, 0.0, 0.0, 0.0};
          tile_times<MF>(x, acc[i][0], mb0, myC);
          acc[i][0] = x;
          store_tile(x, band, ld, b, npad, j0 + 16 * t, j0);
          if (t == 1) {
#pragma unroll
            for (int v = 0; v < 4; ++v) sL10[(4 * v + q) * 16 + jl] = x[v];
          }
        }
      }
      if (p == 0) {
        __syncthreads();
        // B fragment of -L10^T: B[kk][n] = -L10[n][4 cq + kk]
#pragma unroll
        for (int cq = 0; cq < 4; ++cq) mb10[cq] = -sL10[jl * 16 + 4 * cq + q];
      }
      // column tile 1: C(t,1) -= L(t,0) L10^T
#pragma unroll
      for (int i = 0; i < MAXT; ++i) {
        const int t = g.tb + W * i;
        if (t >= 1 && t < T) tile_times<MF>(acc[i][1], acc[i][0], mb10, myC);
      }
      if (p == 0) {
        if (wv == 1) diag_block<MF>(acc[0][1], band, ld, b, j0 + 16, sL, sLi1, sflag);
        __syncthreads();
        if (*sflag) return PF_NOT_SPD;
#pragma unroll
        for (int cq = 0; cq < 4; ++cq) mb1[cq] = sLi1[(4 * cq + q) * 16 + jl];
      }
      // column tile 1 below its diagonal: L(t,1) = C L11^-T
#pragma unroll
      for (int i = 0; i < MAXT; ++i) {
        const int t = g.tb + W * i;
        if (t >= 2 && t < T) {
          d4 x = {0.0, 0.0, 0.0, 0.0};
          tile_times<MF>(x, acc[i][1], mb1, myC);
          store_tile(x, band, ld, b, npad, j0 + 16 * t, j0 + 16);
        }
      }
    }
    __syncthreads();
  }
  return PF_OK;
}

// L L^T z = w in place (w: npad entries of the mesh's work vector).  The 16 x 16 diagonal
// solves run in wave 0 with the block's L entries loaded before the serial steps.
template <int BT>
__device__ void band_solve(const DMesh &M, const double *__restrict__ band, double *__restrict__ w, double *sY) {
  const int wv = wave_id();
  const int b = M.b, npad = M.npad;
  const int64_t ld = M.ld;
  // forward: L y = w
  for (int j0 = 0; j0 < npad; j0 += 16) {
    if (wv == 0) {
      const int lane = lane_id();
      const int jr = lane & 15;
      double Lr[16];  // row jr of the diagonal block: L[j0+jr][j0+k]
#pragma unroll
      for (int k = 0; k < 16; ++k) Lr[k] = band[(int64_t)(j0 + k) * ld + min(max(jr - k, 0), b)];
#pragma unroll
      for (int k = 0; k < 16; ++k)
        if (!(k <= jr && jr - k <= b)) Lr[k] = 0.0;
      double acc = lane < 16 ? w[j0 + lane] : 0.0;
#pragma unroll
      for (int k = 0; k < 16; ++k) {
        if (lane == k) acc /= Lr[k];
        const double yk = shfl(acc, k);
        if (lane > k && lane < 16) acc -= Lr[k] * yk;
      }
      if (lane < 16) {
        w[j0 + lane] = acc;
        sY[lane] = acc;
      }
    }
    __syncthreads();
    const int rend = min(npad, j0 + 16 + b);
    for (int r = j0 + 16 + (int)threadIdx.x; r < rend; r += BT) {
      double l[16];
#pragma unroll
      for (int j = 0; j < 16; ++j) l[j] = band[(int64_t)(j0 + j) * ld + min(r - j0 - j, b)];  // d >= 1
      double s = 0.0;
#pragma unroll
      for (int j = 0; j < 16; ++j)
        if (r - j0 - j <= b) s += l[j] * sY[j];
      w[r] -= s;
    }
    __syncthreads();
  }
  // backward: L^T z = y
  for (int j0 = npad - 16; j0 >= 0; j0 -= 16) {
    if (wv == 0) {
      const int lane = lane_id();
      const int jr = lane & 15;
      double Lc[16];  // column jr of the diagonal block: L[j0+k][j0+jr]
#pragma unroll
      for (int k = 0; k < 16; ++k) Lc[k] = band[(int64_t)(j0 + jr) * ld + min(max(k - jr, 0), b)];
#pragma unroll
      for (int k = 0; k < 16; ++k)
        if (!(k >= jr && k - jr <= b)) Lc[k] = 0.0;
      double acc = lane < 16 ? w[j0 + lane] : 0.0;
#pragma unroll
      for (int k = 15; k >= 0; --k) {
        if (lane == k) acc /= Lc[k];
        const double xk = shfl(acc, k);
        if (lane < k) acc -= Lc[k] * xk;
      }
      if (lane < 16) {
        w[j0 + lane] = acc;
        sY[lane] = acc;
      }
    }
    __syncthreads();
    const int ibeg = max(0, j0 - b);
    for (int i = ibeg + (int)threadIdx.x; i < j0; i += BT) {
      double l[16];
#pragma unroll
      for (int j = 0; j < 16; ++j) l[j] = band[(int64_t)i * ld + min(j0 + j - i, b)];  // d >= 1
      double s = 0.0;
#pragma unroll
      for (int j = 0; j < 16; ++j)
        if (j0 + j - i <= b) s += l[j] * sY[j];
      w[i] -= s;
    }
    __syncthreads();
  }
}

__global__ void k_direct_queue_reset(int32_t *q) {
  if (threadIdx.x == 0 && blockIdx.x == 0) q[0] = 0;
}

__global__ void k_flag_clear(int32_t *f) {
  if (threadIdx.x == 0 && blockIdx.x == 0) f[0] = 0;
}

template <int BT, bool MF>
__global__ void __launch_bounds__(BT) k_direct(DArgs a) {
  __shared__ double sL[256];
  __shared__ double sLi[768];
  __shared__ double sC[256 * (BT / 64)];
  __shared__ double sY[16];
  __shared__ int32_t next, sflag;
  for (;;) {
    __syncthreads();
    if (threadIdx.x == 0) next = atomicAdd(a.queue, 1);
    __syncthreads();
    const int32_t qi = next;
    if (qi >= a.nmesh) break;
    const int32_t m = a.order[qi];
    const DMesh M = a.mesh[m];
    const int64_t t0 = wall_clock64();
    double *band = a.band + M.boff;
    int st = a.mode == 0 ? factor_mesh<BT, MF>(M, band, sL, sLi, sC, &sflag) : a.status[m];
    double *w = a.w + M.woff;
    const int32_t *perm = a.perm + M.woff;
    if (st == PF_OK) {
      for (int q = threadIdx.x; q < M.npad; q += BT) w[q] = q < M.n ? a.rhs[M.d0 + perm[q]] : 0.0;
      __syncthreads();
      band_solve<BT>(M, band, w, sY);
      for (int q = threadIdx.x; q < M.n; q += BT) {
        const int64_t r = M.d0 + perm[q];
        a.x[r] = a.mode == 0 ? w[q] : a.x[r] + w[q];
      }
    } else if (a.mode == 0) {
      for (int q = threadIdx.x; q < M.n; q += BT) a.x[M.d0 + q] = 0.0;
    }
    if (threadIdx.x == 0 && a.mode == 0) {
      a.status[m] = st;
      a.t_raw[2 * m] = t0;
      a.t_raw[2 * m + 1] = wall_clock64();
    }
  }
}

__global__ void k_band_zero(double *__restrict__ band, int64_t count) {
  const int64_t stride = (int64_t)gridDim.x * blockDim.x;
  for (int64_t i = blockIdx.x * (int64_t)blockDim.x + threadIdx.x; i < count; i += stride) band[i] = 0.0;
}

// lower triangle of P A P^T into the band: one SUB-lane subgroup per row
template <int SUB>
__global__ void __launch_bounds__(256) k_band_scatter(int64_t n, const int64_t *__restrict__ row_ptr,
                                                      const int32_t *__restrict__ col_idx,
                                                      const double *__restrict__ val, const DMesh *__restrict__ mesh,
                                                      const int32_t *__restrict__ dmesh,
                                                      const int32_t *__restrict__ bidx, double *__restrict__ band,
                                                      int32_t *__restrict__ bad) {
  const int64_t g = blockIdx.x * (int64_t)blockDim.x + threadIdx.x;
  const int64_t r = g / SUB;
  const int lane = (int)(g % SUB);
  if (r >= n) return;
  const int32_t m = dmesh[r];
  const DMesh M = mesh[m];
  const int64_t pr = bidx[r];
  const int64_t e = row_ptr[r + 1] - 1;
  for (int64_t k = row_ptr[r] - 1 + lane; k < e; k += SUB) {
    const int64_t c = col_idx[k] - 1;
    const int64_t pc = bidx[c];
    // an entry outside the mesh or the analysed band means a stale analysis: flag, never write
    if (dmesh[c] != m || pr - pc > M.b || pc - pr > M.b) {
      atomicOr(bad, 1);
    } else if (pr >= pc) {
      band[M.boff + pc * M.ld + (pr - pc)] = val[k];
    }
  }
}

// unit diagonal of the padding rows n .. npad-1 of every mesh
__global__ void k_band_pad(int32_t nmesh, const DMesh *__restrict__ mesh, double *__restrict__ band) {
  const int32_t m = blockIdx.x;
  if (m >= nmesh) return;
  const DMesh M = mesh[m];
  for (int q = M.n + (int)threadIdx.x; q < M.npad; q += blockDim.x) band[M.boff + (int64_t)q * M.ld] = 1.0;
}

// r = rhs - y over every mesh's rows, and per mesh (|r|^2, |rhs|^2); one workgroup per mesh
__global__ void __launch_bounds__(256) k_direct_resid(int32_t nmesh, const DMesh *__restrict__ mesh,
                                                      const double *__restrict__ rhs, const double *__restrict__ y,
                                                      double *__restrict__ r, double *__restrict__ norms) {
  __shared__ double sh[2][4];
  const int32_t m = blockIdx.x;
  if (m >= nmesh) return;
  const DMesh M = mesh[m];
  double t0 = 0.0, t1 = 0.0;
  for (int64_t i = M.d0 + threadIdx.x; i < M.d0 + M.n; i += 256) {
    const double bi = rhs[i];
    const double ri = bi - y[i];
    r[i] = ri;
    t0 += ri * ri;
    t1 += bi * bi;
  }
  for (int o = 32; o > 0; o >>= 1) {
    t0 += __shfl_xor(t0, o, 64);
    t1 += __shfl_xor(t1, o, 64);
  }
  const int wv = wave_id();
  if ((threadIdx.x & 63) == 0) {
    sh[0][wv] = t0;
    sh[1][wv] = t1;
  }
  __syncthreads();
  if (threadIdx.x == 0) {
    norms[2 * m] = sh[0][0] + sh[0][1] + sh[0][2] + sh[0][3];
    norms[2 * m + 1] = sh[1][0] + sh[1][1] + sh[1][2] + sh[1][3];
  }
}

// node (DOF group) neighbours from a DOF=3 blocked CRS: every third column of row 3i
__global__ void k_group_cols(int64_t nn, const int64_t *__restrict__ row_ptr, const int32_t *__restrict__ col_idx,
                             const int64_t *__restrict__ goff, int32_t *__restrict__ gcol) {
  const int64_t i = blockIdx.x * (int64_t)blockDim.x + threadIdx.x;
  if (i >= nn) return;
  const int64_t r0 = row_ptr[3 * i] - 1;
  const int64_t m = goff[i + 1] - goff[i];
  for (int64_t t = 0; t < m; ++t) gcol[goff[i] + t] = (col_idx[r0 + 3 * t] - 1) / 3;
}

// ---------------------------------------------------------------------------------
// orderings (host)

struct LocalGraph {
  int64_t g0 = 0, ng = 0;                 // groups g0 .. g0+ng-1
  const int64_t *off = nullptr;           // global group adjacency offsets
  const int32_t *adj = nullptr;           // global group ids
};

int64_t bandwidth_of(const LocalGraph &G, const std::vector<int32_t> &pos) {
  int64_t bw = 0;
  for (int64_t i = 0; i < G.ng; ++i) {
    const int64_t pi = pos[i];
    for (int64_t k = G.off[G.g0 + i]; k < G.off[G.g0 + i + 1]; ++k) {
      const int64_t j = G.adj[k] - G.g0;
      if (j < 0 || j >= G.ng) continue;  // outside the mesh (not expected in a checked partition)
      const int64_t d = pi > pos[j] ? pi - pos[j] : pos[j] - pi;
      if (d > bw) bw = d;
    }
  }
  return bw;
}

// reverse Cuthill-McKee from a pseudo-peripheral start (George-Liu), components in turn
std::vector<int32_t> rcm_order(const LocalGraph &G) {
  const int64_t n = G.ng;
  std::vector<int32_t> deg(n);
  for (int64_t i = 0; i < n; ++i) deg[i] = (int32_t)(G.off[G.g0 + i + 1] - G.off[G.g0 + i]);
  std::vector<int32_t> order;
  order.reserve(n);
  std::vector<char> done(n, 0);
  std::vector<int32_t> level(n, -1), queue;
  queue.reserve(n);
  auto bfs_levels = [&](int32_t s, int32_t &last_first, int32_t &nlev) {
    queue.clear();
    queue.push_back(s);
    level[s] = 0;
    size_t h = 0;
    while (h < queue.size()) {
      const int32_t u = queue[h++];
      for (int64_t k = G.off[G.g0 + u]; k < G.off[G.g0 + u + 1]; ++k) {
        const int64_t v = G.adj[k] - G.g0;
        if (v < 0 || v >= n || done[v] || level[v] >= 0) continue;
        level[v] = level[u] + 1;
        queue.push_back((int32_t)v);
      }
    }
    nlev = level[queue.back()] + 1;
    last_first = queue.back();
    int32_t best = queue.back();
    for (size_t t = queue.size(); t-- > 0;) {
      const int32_t u = queue[t];
      if (level[u] != nlev - 1) break;
      if (deg[u] < deg[best]) best = u;
    }
    last_first = best;
    for (int32_t u : queue) level[u] = -1;
  };
  for (;;) {
    int32_t s = -1;
    for (int64_t i = 0; i < n; ++i)
      if (!done[i] && (s < 0 || deg[i] < deg[s])) s = (int32_t)i;
    if (s < 0) break;
    int32_t cand = s, nlev = 0, nlev2 = 0;
    bfs_levels(s, cand, nlev);
    for (int it = 0; it < 8; ++it) {
      int32_t c2 = cand;
      bfs_levels(cand, c2, nlev2);
      if (nlev2 <= nlev) break;
      s = cand;
      nlev = nlev2;
      cand = c2;
    }
    // Cuthill-McKee from s: neighbours by increasing degree
    const size_t first = order.size();
    order.push_back(s);
    done[s] = 1;
    std::vector<int32_t> nb;
    for (size_t h = first; h < order.size(); ++h) {
      const int32_t u = order[h];
      nb.clear();
      for (int64_t k = G.off[G.g0 + u]; k < G.off[G.g0 + u + 1]; ++k) {
        const int64_t v = G.adj[k] - G.g0;
        if (v < 0 || v >= n || done[v]) continue;
        done[v] = 1;
        nb.push_back((int32_t)v);
      }
      std::stable_sort(nb.begin(), nb.end(), [&](int32_t x, int32_t y) { return deg[x] < deg[y]; });
      order.insert(order.end(), nb.begin(), nb.end());
    }
  }
  std::reverse(order.begin(), order.end());
  return order;
}

struct MeshOrder {
  std::vector<int32_t> order;  // new position -> local group
  int64_t bw = 0;              // in groups
  int32_t kind = 0;            // 0 natural, 1 RCM, 2/3/4 coordinate sort with x/y/z primary
};

MeshOrder best_order(const LocalGraph &G, const double *coords, int64_t ncoord) {
  MeshOrder best;
  best.order.resize(G.ng);
  std::iota(best.order.begin(), best.order.end(), 0);
  std::vector<int32_t> pos(G.ng);
  auto eval = [&](const std::vector<int32_t> &ord) {
    for (int64_t k = 0; k < G.ng; ++k) pos[ord[k]] = (int32_t)k;
    return bandwidth_of(G, pos);
  };
  best.bw = eval(best.order);
  best.kind = 0;
  {
    std::vector<int32_t> o = rcm_order(G);
    const int64_t bw = eval(o);
    if (bw < best.bw) {
      best.bw = bw;
      best.order = std::move(o);
      best.kind = 1;
    }
  }
  if (coords) {
    double lo[3], hi[3];
    for (int d = 0; d < 3; ++d) {
      lo[d] = 1e300;
      hi[d] = -1e300;
      for (int64_t i = 0; i < G.ng; ++i) {
        const double v = coords[d * ncoord + G.g0 + i];
        lo[d] = std::min(lo[d], v);
        hi[d] = std::max(hi[d], v);
      }
    }
    for (int p = 0; p < 3; ++p) {
      int s = (p + 1) % 3, t = (p + 2) % 3;
      if (hi[t] - lo[t] > hi[s] - lo[s]) std::swap(s, t);
      std::vector<int32_t> o(G.ng);
      std::iota(o.begin(), o.end(), 0);
      auto key = [&](int32_t i, int d) { return coords[d * ncoord + G.g0 + i]; };
      std::stable_sort(o.begin(), o.end(), [&](int32_t x, int32_t y) {
        if (key(x, p) != key(y, p)) return key(x, p) < key(y, p);
        if (key(x, s) != key(y, s)) return key(x, s) < key(y, s);
        return key(x, t) < key(y, t);
      });
      const int64_t bw = eval(o);
      if (bw < best.bw) {
        best.bw = bw;
        best.order = std::move(o);
        best.kind = 2 + p;
      }
    }
  }
  return best;
}

}  // namespace

// ---------------------------------------------------------------------------------

struct DirectWork {
  // analysis (sparsity pattern and partition)
  uint64_t pattern_key = 0;
  std::vector<int64_t> dof_off;
  int32_t nmesh = 0;
  int32_t group = 1;
  int64_t n = 0;
  std::vector<DMesh> hmesh;
  std::vector<int32_t> kinds;
  std::vector<int32_t> hperm;  // band index -> local DOF, per mesh at woff
  DevBuf<DMesh> mesh;
  DevBuf<int32_t> perm, bidx, dmesh, order, queue, status, flag;
  DevBuf<int64_t> t_raw;
  DevBuf<double> band, w, r, y, norms;
  int64_t band_entries = 0, work_entries = 0;
  double analyse_ms = 0.0;
  // factor held in band: the matrix version it was computed from (0: none)
  uint64_t factor_version = 0;
  bool factor_ok = false;
};

void pf_direct_work_destroy(DirectWork *d) { delete d; }

void pf_direct_invalidate(DirectWork *d) {
  if (d) d->factor_version = 0;
}

int pf_direct_analyse(pf_ctx *ctx, DirectWork **pd, int64_t n, const int64_t *row_ptr, const int32_t *col_idx,
                      int32_t nmesh, const int64_t *dof_off, const double *coords_dev, int64_t ncoord,
                      uint64_t pattern_key) {
  if (nmesh < 1 || n < 1 || dof_off[0] != 0 || dof_off[nmesh] != n) {
    pf_set_error("direct: meshes must partition the %lld DOFs", (long long)n);
    return PF_INVALID;
  }
  for (int32_t m = 0; m < nmesh; ++m)
    if (dof_off[m + 1] <= dof_off[m]) {
      pf_set_error("direct: mesh %d is empty", m);
      return PF_INVALID;
    }
  if (!*pd) *pd = new DirectWork();
  DirectWork *D = *pd;
  if (D->pattern_key == pattern_key && D->nmesh == nmesh && D->n == n &&
      (int64_t)D->dof_off.size() == nmesh + 1 && std::equal(D->dof_off.begin(), D->dof_off.end(), dof_off))
    return PF_OK;
  const auto tstart = std::chrono::steady_clock::now();
  D->pattern_key = 0;
  D->factor_version = 0;
  D->nmesh = 0;
  PF_CHECK(D->flag.ensure(1));
  int blocked3 = 0;
  if (n % 3 == 0) PF_CHECK(pf_block_structure_check(ctx, n, row_ptr, col_idx, n / 3, D->flag.p, &blocked3));
  const int32_t g = blocked3 ? 3 : 1;
  for (int32_t m = 0; m <= nmesh; ++m)
    if (dof_off[m] % g != 0) {
      pf_set_error("direct: mesh offsets must be whole DOF=3 blocks");
      return PF_INVALID;
    }
  const int64_t ng = n / g;
  // group adjacency on the host
  std::vector<int64_t> hrp(n + 1);
  PF_HIP(hipMemcpyAsync(hrp.data(), row_ptr, sizeof(int64_t) * (n + 1), hipMemcpyDeviceToHost, ctx->stream));
  PF_HIP(hipStreamSynchronize(ctx->stream));
  std::vector<int64_t> goff(ng + 1, 0);
  std::vector<int32_t> gadj;
  if (g == 3) {
    for (int64_t i = 0; i < ng; ++i) goff[i + 1] = goff[i] + (hrp[3 * i + 1] - hrp[3 * i]) / 3;
    DevBuf<int64_t> doff;
    DevBuf<int32_t> dcol;
    PF_CHECK(doff.ensure(ng + 1));
    PF_CHECK(dcol.ensure(std::max<int64_t>(goff[ng], 1)));
    PF_HIP(hipMemcpyAsync(doff.p, goff.data(), sizeof(int64_t) * (ng + 1), hipMemcpyHostToDevice, ctx->stream));
    k_group_cols<<<pf_blocks(ng, 256, INT32_MAX), 256, 0, ctx->stream>>>(ng, row_ptr, col_idx, doff.p, dcol.p);
    PF_HIP(hipGetLastError());
    gadj.resize(goff[ng]);
    PF_HIP(hipMemcpyAsync(gadj.data(), dcol.p, sizeof(int32_t) * goff[ng], hipMemcpyDeviceToHost, ctx->stream));
    PF_HIP(hipStreamSynchronize(ctx->stream));
  } else {
    const int64_t nnz = hrp[n] - 1;
    gadj.resize(nnz);
    PF_HIP(hipMemcpyAsync(gadj.data(), col_idx, sizeof(int32_t) * nnz, hipMemcpyDeviceToHost, ctx->stream));
    PF_HIP(hipStreamSynchronize(ctx->stream));
    for (int64_t i = 0; i <= n; ++i) goff[i] = hrp[i] - 1;
    for (auto &c : gadj) c -= 1;
  }
  std::vector<double> coords;
  if (coords_dev && g == 3 && ncoord == ng) {
    coords.resize(3 * ncoord);
    PF_HIP(hipMemcpyAsync(coords.data(), coords_dev, sizeof(double) * 3 * ncoord, hipMemcpyDeviceToHost,
                          ctx->stream));
    PF_HIP(hipStreamSynchronize(ctx->stream));
  }
  // per-mesh orderings, meshes spread over host threads
  std::vector<MeshOrder> ord(nmesh);
  {
    unsigned nt = std::max(1u, std::min(16u, std::thread::hardware_concurrency()));
    nt = (unsigned)std::min<int64_t>(nt, nmesh);
    std::vector<std::thread> th;
    std::atomic<int32_t> next{0};
    for (unsigned t = 0; t < nt; ++t)
      th.emplace_back([&]() {
        for (;;) {
          const int32_t m = next.fetch_add(1);
          if (m >= nmesh) break;
          LocalGraph G;
          G.g0 = dof_off[m] / g;
          G.ng = (dof_off[m + 1] - dof_off[m]) / g;
          G.off = goff.data();
          G.adj = gadj.data();
          ord[m] = best_order(G, coords.empty() ? nullptr : coords.data(), ncoord);
        }
      });
    for (auto &t : th) t.join();
  }
  // band layout
  D->hmesh.assign(nmesh, DMesh{});
  D->kinds.assign(nmesh, 0);
  int64_t boff = 0, woff = 0;
  for (int32_t m = 0; m < nmesh; ++m) {
    DMesh &M = D->hmesh[m];
    const int64_t nm = dof_off[m + 1] - dof_off[m];
    if (nm > (int64_t)INT32_MAX - 16) {
      pf_set_error("direct: mesh %d has %lld DOFs, more than the band kernels index", m, (long long)nm);
      return PF_INVALID;
    }
    M.d0 = dof_off[m];
    M.n = (int32_t)nm;
    M.npad = (int32_t)((nm + 31) / 32 * 32);
    M.b = (int32_t)(g * ord[m].bw + g - 1);
    M.ld = (M.b + 2) & ~1;
    M.boff = boff;
    M.woff = woff;
    boff += (int64_t)M.npad * M.ld;
    woff += M.npad;
    D->kinds[m] = ord[m].kind;
  }
  size_t free_b = 0, total_b = 0;
  PF_HIP(hipMemGetInfo(&free_b, &total_b));
  const double need = 8.0 * (double)boff;
  const double have = (double)free_b + 8.0 * (double)D->band.n;
  if (need > have - 2.0e9) {
    pf_set_error("direct: the band factors need %.2f GB of device memory, %.2f GB are free", need * 1e-9,
                 have * 1e-9);
    return PF_ERR_NOMEM;
  }
  D->hperm.assign(woff, 0);
  std::vector<int32_t> hbidx(n), hdm(n);
  for (int32_t m = 0; m < nmesh; ++m) {
    const DMesh &M = D->hmesh[m];
    const std::vector<int32_t> &o = ord[m].order;
    for (int64_t k = 0; k < (int64_t)o.size(); ++k)
      for (int d = 0; d < g; ++d) {
        const int32_t q = (int32_t)(g * k + d);
        const int32_t loc = g * o[k] + d;
        D->hperm[M.woff + q] = loc;
        hbidx[M.d0 + loc] = q;
        hdm[M.d0 + loc] = m;
      }
  }
  std::vector<int32_t> qorder(nmesh);
  std::iota(qorder.begin(), qorder.end(), 0);
  std::stable_sort(qorder.begin(), qorder.end(), [&](int32_t x, int32_t y) {
    const DMesh &a = D->hmesh[x], &b = D->hmesh[y];
    return (double)a.npad * a.b * a.b > (double)b.npad * b.b * b.b;
  });
  PF_CHECK(D->band.ensure(std::max<int64_t>(boff, 1)));
  PF_CHECK(D->mesh.ensure(nmesh));
  PF_CHECK(D->perm.ensure(std::max<int64_t>(woff, 1)));
  PF_CHECK(D->w.ensure(std::max<int64_t>(woff, 1)));
  PF_CHECK(D->bidx.ensure(n));
  PF_CHECK(D->dmesh.ensure(n));
  PF_CHECK(D->order.ensure(nmesh));
  PF_CHECK(D->queue.ensure(1));
  PF_CHECK(D->status.ensure(nmesh));
  PF_CHECK(D->t_raw.ensure(2 * (size_t)nmesh));
  PF_CHECK(D->norms.ensure(2 * (size_t)nmesh));
  PF_CHECK(D->r.ensure(n + 1));
  PF_CHECK(D->y.ensure(n + 1));
  PF_HIP(hipMemcpyAsync(D->mesh.p, D->hmesh.data(), sizeof(DMesh) * nmesh, hipMemcpyHostToDevice, ctx->stream));
  PF_HIP(hipMemcpyAsync(D->perm.p, D->hperm.data(), sizeof(int32_t) * woff, hipMemcpyHostToDevice, ctx->stream));
  PF_HIP(hipMemcpyAsync(D->bidx.p, hbidx.data(), sizeof(int32_t) * n, hipMemcpyHostToDevice, ctx->stream));
  PF_HIP(hipMemcpyAsync(D->dmesh.p, hdm.data(), sizeof(int32_t) * n, hipMemcpyHostToDevice, ctx->stream));
  PF_HIP(hipMemcpyAsync(D->order.p, qorder.data(), sizeof(int32_t) * nmesh, hipMemcpyHostToDevice, ctx->stream));
  PF_HIP(hipStreamSynchronize(ctx->stream));
  D->band_entries = boff;
  D->work_entries = woff;
  D->group = g;
  D->n = n;
  D->dof_off.assign(dof_off, dof_off + nmesh + 1);
  D->nmesh = nmesh;
  D->pattern_key = pattern_key;
  D->analyse_ms =
      std::chrono::duration<double, std::milli>(std::chrono::steady_clock::now() - tstart).count();
  return PF_OK;
}

namespace {
int launch_direct(pf_ctx *ctx, DirectWork *D, int mode, const double *rhs, double *x, bool mfma) {
  DArgs a;
  a.mesh = D->mesh.p;
  a.order = D->order.p;
  a.queue = D->queue.p;
  a.nmesh = D->nmesh;
  a.mode = mode;
  a.band = D->band.p;
  a.perm = D->perm.p;
  a.w = D->w.p;
  a.rhs = rhs;
  a.x = x;
  a.status = D->status.p;
  a.t_raw = D->t_raw.p;
  // two workgroups per CU at most; no more workgroups than meshes
  const int grid = std::max(1, std::min(D->nmesh, 2 * ctx->num_cus));
  k_direct_queue_reset<<<1, 64, 0, ctx->stream>>>(D->queue.p);
  PF_HIP(hipGetLastError());
  if (mfma)
    k_direct<DBT, true><<<grid, DBT, 0, ctx->stream>>>(a);
  else
    k_direct<DBT, false><<<grid, DBT, 0, ctx->stream>>>(a);
  PF_HIP(hipGetLastError());
  return PF_OK;
}

// r = rhs - A x and the per-mesh norms
int residual(pf_ctx *ctx, DirectWork *D, int64_t n, const int64_t *row_ptr, const int32_t *col_idx,
             const double *val, double avg_row, const double *rhs, const double *x) {
  PF_CHECK(pf_spmv_crs_launch<double>(ctx, n, row_ptr, col_idx, val, x, D->y.p, avg_row));
  k_direct_resid<<<D->nmesh, 256, 0, ctx->stream>>>(D->nmesh, D->mesh.p, rhs, D->y.p, D->r.p, D->norms.p);
  PF_HIP(hipGetLastError());
  return PF_OK;
}
}  // namespace

int pf_direct_solve_device(pf_ctx *ctx, DirectWork *D, int64_t n, const int64_t *row_ptr, const int32_t *col_idx,
                           const double *val, int64_t nnz, const double *rhs, double *x, uint64_t mat_version,
                           int32_t refine, int32_t mfma, pf_batch_mesh_stats *mesh_stats, pf_stats *total) {
  if (!D || D->nmesh < 1 || D->n != n) {
    pf_set_error("direct: no analysis for this system");
    return PF_INVALID;
  }
  if (refine < 0 || refine > 100) {
    pf_set_error("direct: refinement steps must be in [0, 100]");
    return PF_INVALID;
  }
  const double avg_row = n > 0 ? (double)nnz / (double)n : 1.0;
  float setup_ms = 0.f, solve_ms = 0.f;
  const bool refactor = D->factor_version != mat_version || mat_version == 0;
  PF_HIP(hipEventRecord(ctx->ev0, ctx->stream));
  if (refactor) {
    D->factor_version = 0;
    k_band_zero<<<pf_blocks(D->band_entries, 256, 8 * 2048), 256, 0, ctx->stream>>>(D->band.p, D->band_entries);
    PF_HIP(hipGetLastError());
    k_flag_clear<<<1, 64, 0, ctx->stream>>>(D->flag.p);
    k_band_scatter<16><<<pf_blocks(n * 16, 256, INT32_MAX), 256, 0, ctx->stream>>>(
        n, row_ptr, col_idx, val, D->mesh.p, D->dmesh.p, D->bidx.p, D->band.p, D->flag.p);
    PF_HIP(hipGetLastError());
    int32_t bad = 0;
    PF_HIP(hipMemcpyAsync(&bad, D->flag.p, sizeof(int32_t), hipMemcpyDeviceToHost, ctx->stream));
    PF_HIP(hipStreamSynchronize(ctx->stream));
    if (bad) {
      pf_set_error("direct: the matrix has entries outside the analysed meshes or band (pattern changed?)");
      return PF_INVALID;
    }
    k_band_pad<<<D->nmesh, 64, 0, ctx->stream>>>(D->nmesh, D->mesh.p, D->band.p);
    PF_HIP(hipGetLastError());
  }
  PF_HIP(hipEventRecord(ctx->ev1, ctx->stream));
  PF_HIP(hipEventSynchronize(ctx->ev1));
  PF_HIP(hipEventElapsedTime(&setup_ms, ctx->ev0, ctx->ev1));
  PF_HIP(hipEventRecord(ctx->ev0, ctx->stream));
  if (refactor) {
    PF_CHECK(launch_direct(ctx, D, 0, rhs, x, mfma != 0));
  } else {
    // factor kept: solve only (x = 0 + A^-1 rhs)
    PF_HIP(hipMemsetAsync(x, 0, sizeof(double) * n, ctx->stream));
    PF_CHECK(launch_direct(ctx, D, 1, rhs, x, mfma != 0));
  }
  for (int32_t s = 0; s < refine; ++s) {
    PF_CHECK(residual(ctx, D, n, row_ptr, col_idx, val, avg_row, rhs, x));
    PF_CHECK(launch_direct(ctx, D, 1, D->r.p, x, mfma != 0));
  }
  PF_CHECK(residual(ctx, D, n, row_ptr, col_idx, val, avg_row, rhs, x));
  PF_HIP(hipEventRecord(ctx->ev1, ctx->stream));
  PF_HIP(hipEventSynchronize(ctx->ev1));
  PF_HIP(hipEventElapsedTime(&solve_ms, ctx->ev0, ctx->ev1));
  const int32_t nm = D->nmesh;
  std::vector<int32_t> st(nm);
  std::vector<int64_t> raw(2 * (size_t)nm);
  std::vector<double> nrm(2 * (size_t)nm);
  PF_HIP(hipMemcpyAsync(st.data(), D->status.p, sizeof(int32_t) * nm, hipMemcpyDeviceToHost, ctx->stream));
  PF_HIP(hipMemcpyAsync(raw.data(), D->t_raw.p, sizeof(int64_t) * 2 * nm, hipMemcpyDeviceToHost, ctx->stream));
  PF_HIP(hipMemcpyAsync(nrm.data(), D->norms.p, sizeof(double) * 2 * nm, hipMemcpyDeviceToHost, ctx->stream));
  PF_HIP(hipStreamSynchronize(ctx->stream));
  int rate_khz = 0;
  PF_HIP(hipDeviceGetAttribute(&rate_khz, hipDeviceAttributeWallClockRate, ctx->device));
  const double tick_ms = rate_khz > 0 ? 1.0 / (double)rate_khz : 0.0;
  int64_t tmin = raw[0];
  for (int32_t m = 0; m < nm; ++m) tmin = std::min(tmin, raw[2 * m]);
  int status = PF_OK;
  double worst = 0.0;
  bool all_ok = true;
  for (int32_t m = 0; m < nm; ++m) {
    const double relres = nrm[2 * m + 1] > 0.0 ? std::sqrt(nrm[2 * m] / nrm[2 * m + 1]) : std::sqrt(nrm[2 * m]);
    if (st[m] != PF_OK) all_ok = false;
    if (status == PF_OK && st[m] != PF_OK) status = st[m];
    if (st[m] == PF_OK) worst = std::max(worst, relres);
    if (mesh_stats) {
      pf_batch_mesh_stats &s = mesh_stats[m];
      std::memset(&s, 0, sizeof(s));
      s.status = st[m];
      s.iterations = st[m] == PF_OK ? 1 + refine : 0;
      s.replacements = st[m] == PF_OK ? refine : 0;
      s.workgroup = -1;
      s.true_relres = relres;
      s.start_ms = (double)(raw[2 * m] - tmin) * tick_ms;
      s.end_ms = (double)(raw[2 * m + 1] - tmin) * tick_ms;
    }
  }
  if (refactor && all_ok) D->factor_version = mat_version;
  if (total) {
    std::memset(total, 0, sizeof(*total));
    total->status = status;
    total->iterations = 1 + refine;
    total->outer_iterations = refine;
    total->true_relres = worst;
    total->solve_ms = solve_ms;
    total->setup_ms = setup_ms;
  }
  return status;
}

int pf_direct_info(DirectWork *D, int32_t mesh, int64_t *n, int32_t *bandwidth, int32_t *kind, int32_t *perm) {
  if (!D || D->nmesh < 1 || mesh < 0 || mesh >= D->nmesh) {
    pf_set_error("direct: no analysis, or mesh %d out of range", mesh);
    return PF_INVALID;
  }
  const DMesh &M = D->hmesh[mesh];
  if (n) *n = M.n;
  if (bandwidth) *bandwidth = M.b;
  if (kind) *kind = D->kinds[mesh];
  if (perm) std::memcpy(perm, D->hperm.data() + M.woff, sizeof(int32_t) * M.n);
  return PF_OK;
}

int pf_direct_factor_get(pf_ctx *ctx, DirectWork *D, int32_t mesh, double *ab) {
  if (!D || D->nmesh < 1 || mesh < 0 || mesh >= D->nmesh || D->factor_version == 0) {
    pf_set_error("direct: no factor held for mesh %d", mesh);
    return PF_INVALID;
  }
  const DMesh &M = D->hmesh[mesh];
  // LAPACK lower band layout with ldab = b + 1, columns 0 .. n-1
  PF_HIP(hipMemcpy2DAsync(ab, sizeof(double) * (M.b + 1), D->band.p + M.boff, sizeof(double) * M.ld,
                          sizeof(double) * (M.b + 1), M.n, hipMemcpyDeviceToHost, ctx->stream));
  PF_HIP(hipStreamSynchronize(ctx->stream));
  return PF_OK;
}
