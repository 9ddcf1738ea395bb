// assemble.hip -- hex8 element stiffness and its CRS assembly on the device.
//
// Element matrix (StiffnessMatrixFEMDomain, FEMDomainClass.f90:11457-11515):
//   Ke = sum_{gp=1..8} (B^T D) B * detJ,  2x2x2 Gauss points in the reference's table order
//   (ShapeFunctionClass.f90:503-529), unit weights, J = dN/dxi * X (:1476),
//   B from G = J^-T dN/dxi (reference, FEMDomainClass.f90:12289) or J^-1 dN/dxi (textbook).
// The kernels evaluate one 24-wide ROW of Ke per thread and keep the reference's
// operation order: (B^T D)(r,k) and ((B^T D) B)(r,c) are Fortran matmul sums in ascending
// k whose only nonzero terms are written out below in that order, accumulated over Gauss
// points as K = K + (..)*detJ.  This file is compiled with -ffp-contract=off, so every
// rounding matches a plain IEEE evaluation of the reference formulas.
//
// Assembly (StiffnessMatrix_as_CRS_FEMDomain, FEMDomainClass.f90:11193-11229) is a
// deterministic gather instead of the reference's OpenMP reduction: the element matrices are
// formed once into a scratch array, then one wavefront per node (k_gather_nodes) walks the
// node's incident elements in ascending id and adds their Ke rows into the node's three CRS
// rows, so every entry is summed in ascending element order with no atomics.
#include <algorithm>

#include "pf_internal.hpp"

#pragma clang fp contract(off)

namespace {

constexpr double GPV = 0.57735026918962576;
// local node sign pattern of the hex8 (ShapeFunctionClass.f90:781-788)
__device__ __constant__ double SX[8] = {-1, 1, 1, -1, -1, 1, 1, -1};
__device__ __constant__ double SY[8] = {-1, -1, 1, 1, -1, -1, 1, 1};
__device__ __constant__ double SZ[8] = {-1, -1, -1, -1, 1, 1, 1, 1};

// dN/dxi of node a at (x,y,z), ShapeFunctionClass.f90:1094-1119, evaluated as
// s*(0.125*(1 + t*u))*(1 + w*v) with the same grouping as the reference literals
__device__ inline void dN_node(int a, double x, double y, double z, double &d0, double &d1, double &d2) {
  const double sx = SX[a], sy = SY[a], sz = SZ[a];
  const double ty = 1.0 + sy * y, tz = 1.0 + sz * z, tx = 1.0 + sx * x;
  // +-(0.125*(1+-y))*(1+-z): the sign is applied to an exact product, so it is exact
  d0 = sx * ((0.125 * ty) * tz);
  d1 = sy * ((0.125 * tx) * tz);
  d2 = sz * ((0.125 * tx) * ty);
}

// det_mat_real64 cofactor expansion along column 1 (MathClass.f90:1460-1476)
__device__ inline double det3(const double J[3][3]) {
  double m1 = J[1][1] * J[2][2] - J[2][1] * J[1][2];
  double m2 = J[0][1] * J[2][2] - J[2][1] * J[0][2];
  double m3 = J[0][1] * J[1][2] - J[1][1] * J[0][2];
  double r = 0.0;
  r = r + J[0][0] * m1;
  r = r + (-J[1][0]) * m2;
  r = r + J[2][0] * m3;
  return r;
}

// inverse_rank_2 (MathClass.f90:1567-1598)
__device__ inline void inv3(const double A[3][3], double Ai[3][3]) {
  const double d1 = 1.0 / det3(A);
  Ai[0][0] = d1 * (A[1][1] * A[2][2] - A[1][2] * A[2][1]);
  Ai[0][1] = d1 * (A[0][2] * A[2][1] - A[0][1] * A[2][2]);
  Ai[0][2] = d1 * (A[0][1] * A[1][2] - A[0][2] * A[1][1]);
  Ai[1][0] = d1 * (A[1][2] * A[2][0] - A[1][0] * A[2][2]);
  Ai[1][1] = d1 * (A[0][0] * A[2][2] - A[0][2] * A[2][0]);
  Ai[1][2] = d1 * (A[0][2] * A[1][0] - A[0][0] * A[1][2]);
  Ai[2][0] = d1 * (A[1][0] * A[2][1] - A[1][1] * A[2][0]);
  Ai[2][1] = d1 * (A[0][1] * A[2][0] - A[0][0] * A[2][1]);
  Ai[2][2] = d1 * (A[0][0] * A[1][1] - A[0][1] * A[1][0]);
}

// Row r = 3a+d of Ke into acc[24] (acc[3b+d'] = Ke(3a+d, 3b+d')).
// D (FEMDomainClass.f90:11607-11636): mu = E/2/(1+v), lam = v*E/(1+v)/(1-2v).
__device__ inline void hex8_ke_row(const double (&X)[8][3], double E, double nu, int jmode, int a,
                                   int d, double (&acc)[24]) {
  const double mu = E / 2.0 / (1.0 + nu);
  const double lam = nu * E / (1.0 + nu) / (1.0 - 2.0 * nu);
  const double dg = 2.0 * mu + lam;  // D(1,1) = D(2,2) = D(3,3)
#pragma unroll 1
  for (int gp = 0; gp < 8; ++gp) {
    const double gx = (gp == 1 || gp == 2 || gp == 5 || gp == 6) ? GPV : -GPV;
    const double gy = (gp == 2 || gp == 3 || gp == 6 || gp == 7) ? GPV : -GPV;
    const double gz = (gp >= 4) ? GPV : -GPV;
    double dN[3][8];
#pragma unroll
    for (int b = 0; b < 8; ++b) dN_node(b, gx, gy, gz, dN[0][b], dN[1][b], dN[2][b]);
    double J[3][3];
#pragma unroll
    for (int i = 0; i < 3; ++i)
#pragma unroll
      for (int j = 0; j < 3; ++j) {
        double s = 0.0;
#pragma unroll
        for (int b = 0; b < 8; ++b) s = s + dN[i][b] * X[b][j];
        J[i][j] = s;
      }
    const double detJ = det3(J);
    double Ji[3][3];
    inv3(J, Ji);
    double G[3][8];
#pragma unroll
    for (int p = 0; p < 3; ++p)
#pragma unroll
      for (int b = 0; b < 8; ++b) {
        double s = 0.0;
        if (jmode == 0) {
#pragma unroll
          for (int i = 0; i < 3; ++i) s = s + dN[i][b] * Ji[i][p];
        } else {
#pragma unroll
          for (int i = 0; i < 3; ++i) s = s + Ji[p][i] * dN[i][b];
        }
        G[p][b] = s;
      }
    double ga0 = 0.0, ga1 = 0.0, ga2 = 0.0;
#pragma unroll
    for (int b = 0; b < 8; ++b)
      if (b == a) {
        ga0 = G[0][b];
        ga1 = G[1][b];
        ga2 = G[2][b];
      }
    // (B^T D)(r, 0..5) for r = 3a+d: single nonzero product each
    const double gd = (d == 0) ? ga0 : ((d == 1) ? ga1 : ga2);
    const double T0 = gd * ((d == 0) ? dg : lam);
    const double T1 = gd * ((d == 1) ? dg : lam);
    const double T2 = gd * ((d == 2) ? dg : lam);
    const double T3 = (d == 0) ? ga1 * mu : ((d == 1) ? ga0 * mu : 0.0);
    const double T4 = (d == 0) ? 0.0 : ((d == 1) ? ga2 * mu : ga1 * mu);
    const double T5 = (d == 0) ? ga2 * mu : ((d == 1) ? 0.0 : ga0 * mu);
#pragma unroll
    for (int b = 0; b < 8; ++b) {
      // ((B^T D) B)(r, 3b+d'), nonzero k in ascending order
      const double k0 = (T0 * G[0][b] + T3 * G[1][b]) + T5 * G[2][b];
      const double k1 = (T1 * G[1][b] + T3 * G[0][b]) + T4 * G[2][b];
      const double k2 = (T2 * G[2][b] + T4 * G[1][b]) + T5 * G[0][b];
      if (gp == 0) {
        acc[3 * b + 0] = k0 * detJ;
        acc[3 * b + 1] = k1 * detJ;
        acc[3 * b + 2] = k2 * detJ;
      } else {
        acc[3 * b + 0] = acc[3 * b + 0] + k0 * detJ;
        acc[3 * b + 1] = acc[3 * b + 1] + k1 * detJ;
        acc[3 * b + 2] = acc[3 * b + 2] + k2 * detJ;
      }
    }
  }
}

__device__ inline void load_element(int64_t nn, int64_t ne, const double *__restrict__ xyz,
                                    const int32_t *__restrict__ en, int64_t e, int32_t (&nodes)[8],
                                    double (&X)[8][3]) {
#pragma unroll
  for (int b = 0; b < 8; ++b) {
    nodes[b] = en[(int64_t)b * ne + e];
    const int64_t k = nodes[b] - 1;
    X[b][0] = xyz[k];
    X[b][1] = xyz[nn + k];
    X[b][2] = xyz[2 * nn + k];
  }
}

// one thread per CRS row r = 3i + d of a hex8 / DOF=3 mesh
__global__ void __launch_bounds__(128) k_assemble_rows(
    int64_t nn, int64_t ne, const double *__restrict__ xyz, const int32_t *__restrict__ en,
    const double *__restrict__ E, const double *__restrict__ nu, int jmode,
    const int64_t *__restrict__ eoff, const int32_t *__restrict__ elist,
    const int64_t *__restrict__ noff, const int32_t *__restrict__ ncol,
    const int64_t *__restrict__ row_ptr, double *__restrict__ val) {
  const int64_t r = blockIdx.x * (int64_t)blockDim.x + threadIdx.x;
  if (r >= 3 * nn) return;
  const int64_t i = r / 3;
  const int d = (int)(r - 3 * i);
  const int64_t base = row_ptr[r] - 1;
  const int64_t len = row_ptr[r + 1] - 1 - base;
  for (int64_t k = 0; k < len; ++k) val[base + k] = 0.0;
  const int64_t c0 = noff[i];
  const int32_t deg = (int32_t)(noff[i + 1] - c0);
  const int64_t s0 = eoff[i], s1 = eoff[i + 1];
  for (int64_t s = s0; s < s1; ++s) {
    const int64_t e = elist[s];
    if (s > s0 && elist[s - 1] == e) continue;  // node listed twice in one element
    int32_t nodes[8];
    double X[8][3];
    load_element(nn, ne, xyz, en, e, nodes, X);
    const double Ee = E[e], ve = nu[e];
    for (int a = 0; a < 8; ++a) {
      if (nodes[a] != (int32_t)(i + 1)) continue;
      double acc[24];
      hex8_ke_row(X, Ee, ve, jmode, a, d, acc);
#pragma unroll
      for (int b = 0; b < 8; ++b) {
        int32_t p = 0;
        const int32_t target = nodes[b] - 1;
        for (int32_t q = 0; q < deg; ++q)
          if (ncol[c0 + q] == target) {
            p = q;
            break;
          }
        double *dst = val + base + 3 * (int64_t)p;
        dst[0] = dst[0] + acc[3 * b + 0];
        dst[1] = dst[1] + acc[3 * b + 1];
        dst[2] = dst[2] + acc[3 * b + 2];
      }
    }
  }
}

// dense Ke of elements first..first+count-1: one thread per (element, row)
__global__ void __launch_bounds__(128) k_ke_batch(int64_t nn, int64_t ne, const double *__restrict__ xyz,
                                                  const int32_t *__restrict__ en,
                                                  const double *__restrict__ E,
                                                  const double *__restrict__ nu, int jmode,
                                                  int64_t first, int64_t count,
                                                  double *__restrict__ ke) {
  const int64_t t = blockIdx.x * (int64_t)blockDim.x + threadIdx.x;
  if (t >= count * 24) return;
  const int64_t q = t / 24;
  const int r = (int)(t - q * 24);
  const int64_t e = first + q;
  int32_t nodes[8];
  double X[8][3];
  load_element(nn, ne, xyz, en, e, nodes, X);
  double acc[24];
  hex8_ke_row(X, E[e], nu[e], jmode, r / 3, r % 3, acc);
  double *out = ke + q * 576;
#pragma unroll
  for (int c = 0; c < 24; ++c) out[c * 24 + r] = acc[c];
}

// ---- body force and stress recovery --------------------------------------------------
// J, detJ and G = (J^-T or J^-1) dN/dxi at Gauss point gp, in the operation order of
// hex8_ke_row (ShapeFunctionClass.f90:1476, FEMDomainClass.f90:12289)
__device__ inline double hex8_gp_G(const double (&X)[8][3], int gp, int jmode, double (&G)[3][8], bool want_G) {
  const double gx = (gp == 1 || gp == 2 || gp == 5 || gp == 6) ? GPV : -GPV;
  const double gy = (gp == 2 || gp == 3 || gp == 6 || gp == 7) ? GPV : -GPV;
  const double gz = (gp >= 4) ? GPV : -GPV;
  double dN[3][8];
#pragma unroll
  for (int b = 0; b < 8; ++b) dN_node(b, gx, gy, gz, dN[0][b], dN[1][b], dN[2][b]);
  double J[3][3];
#pragma unroll
  for (int i = 0; i < 3; ++i)
#pragma unroll
    for (int j = 0; j < 3; ++j) {
      double s = 0.0;
#pragma unroll
      for (int b = 0; b < 8; ++b) s = s + dN[i][b] * X[b][j];
      J[i][j] = s;
    }
  const double detJ = det3(J);
  if (want_G) {
    double Ji[3][3];
    inv3(J, Ji);
#pragma unroll
    for (int p = 0; p < 3; ++p)
#pragma unroll
      for (int b = 0; b < 8; ++b) {
        double s = 0.0;
        if (jmode == 0) {
#pragma unroll
          for (int i = 0; i < 3; ++i) s = s + dN[i][b] * Ji[i][p];
        } else {
#pragma unroll
          for (int i = 0; i < 3; ++i) s = s + Ji[p][i] * dN[i][b];
        }
        G[p][b] = s;
      }
  }
  return detJ;
}

// Ke of one element per wavefront with the geometry evaluated once per element: lanes 0..7
// evaluate J, det J and G = (J^-T or J^-1) dN/dxi at Gauss point `lane` (hex8_gp_G, the same
// operations as hex8_ke_row) into LDS; lane 2r+h (r < 24) then forms row r of Ke over the
// columns of nodes 4h..4h+3 with hex8_ke_row's products, sums and Gauss point order, so every
// entry is bitwise the one hex8_ke_row gives.  COLMAJOR: ke[q*576 + c*24 + r] (the reference's
// Fortran layout), else row-major ke[q*576 + r*24 + c].
template <bool COLMAJOR>
__global__ void __launch_bounds__(256) k_ke_elem(int64_t nn, int64_t ne, const double *__restrict__ xyz,
                                                 const int32_t *__restrict__ en, const double *__restrict__ E,
                                                 const double *__restrict__ nu, int jmode, int64_t first,
                                                 int64_t count, double *__restrict__ ke) {
  __shared__ double sG[4][8][25];  // per wavefront: 8 Gauss points x (G(3,8), det J)
  const int w = threadIdx.x >> 6, lane = threadIdx.x & 63;
  const int64_t q = (blockIdx.x * (int64_t)blockDim.x + threadIdx.x) >> 6;
  const bool live = q < count;
  const int64_t e = first + (live ? q : 0);
  if (live && lane < 8) {
    int32_t nodes[8];
    double X[8][3];
    load_element(nn, ne, xyz, en, e, nodes, X);
    double G[3][8];
    const double dj = hex8_gp_G(X, lane, jmode, G, true);
#pragma unroll
    for (int p = 0; p < 3; ++p)
#pragma unroll
      for (int b = 0; b < 8; ++b) sG[w][lane][p * 8 + b] = G[p][b];
    sG[w][lane][24] = dj;
  }
  __syncthreads();
  if (!live || lane >= 48) return;
  const int r = lane >> 1, h = lane & 1;
  const int a = r / 3, d = r - 3 * (r / 3);
  const double Ee = E[e], ve = nu[e];
  const double mu = Ee / 2.0 / (1.0 + ve);
  const double lam = ve * Ee / (1.0 + ve) / (1.0 - 2.0 * ve);
  const double dg = 2.0 * mu + lam;
  double acc[12];
#pragma unroll 1
  for (int gp = 0; gp < 8; ++gp) {
    const double *g = sG[w][gp];
    const double detJ = g[24];
    const double ga0 = g[a], ga1 = g[8 + a], ga2 = g[16 + a];
    const double gd = (d == 0) ? ga0 : ((d == 1) ? ga1 : ga2);
    const double T0 = gd * ((d == 0) ? dg : lam);
    const double T1 = gd * ((d == 1) ? dg : lam);
    const double T2 = gd * ((d == 2) ? dg : lam);
    const double T3 = (d == 0) ? ga1 * mu : ((d == 1) ? ga0 * mu : 0.0);
    const double T4 = (d == 0) ? 0.0 : ((d == 1) ? ga2 * mu : ga1 * mu);
    const double T5 = (d == 0) ? ga2 * mu : ((d == 1) ? 0.0 : ga0 * mu);
#pragma unroll
    for (int bb = 0; bb < 4; ++bb) {
      const int b = 4 * h + bb;
      const double G0 = g[b], G1 = g[8 + b], G2 = g[16 + b];
      const double k0 = (T0 * G0 + T3 * G1) + T5 * G2;
      const double k1 = (T1 * G1 + T3 * G0) + T4 * G2;
      const double k2 = (T2 * G2 + T4 * G1) + T5 * G0;
      if (gp == 0) {
        acc[3 * bb + 0] = k0 * detJ;
        acc[3 * bb + 1] = k1 * detJ;
        acc[3 * bb + 2] = k2 * detJ;
      } else {
        acc[3 * bb + 0] = acc[3 * bb + 0] + k0 * detJ;
        acc[3 * bb + 1] = acc[3 * bb + 1] + k1 * detJ;
        acc[3 * bb + 2] = acc[3 * bb + 2] + k2 * detJ;
      }
    }
  }
  double *out = ke + q * 576;
#pragma unroll
  for (int t = 0; t < 12; ++t) {
    const int c = 12 * h + t;
    out[COLMAJOR ? c * 24 + r : r * 24 + c] = acc[t];
  }
}

// ---- element matrices through the isotropic split (contraction on MFMA or on the VALU) ----
// For the isotropic D of DMatrix (lam, mu), B^T D B summed over the Gauss points splits
// per node pair (a, b) into
//   Ke(3a+d, 3b+e) = lam M_ab(d, e) + mu M_ab(e, d) + mu [d == e] (M_ab(0,0) + M_ab(1,1) + M_ab(2,2)),
//   M_ab(p, q) = sum_gp detJ G(p, a) G(q, b),
// so the only dense contraction is the 24 x 24 x 8 product M = Gt^T diag(detJ) Gt with
// Gt(gp, 8p + a) = G(p, a) at Gauss point gp.  The geometry (J, det J, G) is bitwise the one
// of hex8_gp_G; the sums are in another order than the reference's matmul chain, so the entries
// agree with k_ke_elem to rounding, not bitwise.  MF: the product on the FP64 matrix cores
// (v_mfma_f64_16x16x4f64, tiles (0,0), (0,1), (1,1) of the 32 x 32 padded M, two steps of 4
// Gauss points each, M(1,0) = M(0,1)^T through LDS), else each lane forms the 3 x 3 block M_ab
// of its node pair in registers.  One wavefront per 8 elements: lane 8el + gp evaluates the
// geometry of element el at Gauss point gp into the wave's LDS; then per element, lane
// 8a + b builds and stores the 3 x 3 block (a, b) of Ke (row-major, as k_ke_elem<false>).
typedef double ke_d4 __attribute__((ext_vector_type(4)));

__device__ inline void wave_lds_sync() {
  // LDS operations of one wavefront are performed in issue order; the fences keep the compiler
  // from moving LDS accesses across this point
  __builtin_amdgcn_fence(__ATOMIC_RELEASE, "wavefront");
  __builtin_amdgcn_wave_barrier();
  __builtin_amdgcn_fence(__ATOMIC_ACQUIRE, "wavefront");
}

constexpr int KI_W = 4;     // wavefronts per block
constexpr int KI_EL = 8;    // elements per wavefront
constexpr int KI_GS = 25;   // G(3,8) and det J per Gauss point
constexpr int KI_MS = 25;   // row stride of M in LDS

template <bool MF>
__global__ void __launch_bounds__(256) k_ke_iso(int64_t nn, int64_t ne, const double *__restrict__ xyz,
                                                const int32_t *__restrict__ en, const double *__restrict__ E,
                                                const double *__restrict__ nu, int jmode, int64_t first,
                                                int64_t count, double *__restrict__ ke) {
  __shared__ double sG[KI_W][KI_EL][8][KI_GS];
  __shared__ double sLM[KI_W][KI_EL][2];
  __shared__ double sM[MF ? KI_W : 1][MF ? 24 * KI_MS : 1];
  const int w = threadIdx.x >> 6, lane = threadIdx.x & 63;
  const int64_t q0 = ((blockIdx.x * (int64_t)blockDim.x + threadIdx.x) >> 6) * KI_EL;
  if (q0 >= count) return;  // whole wavefronts; no block-wide barrier follows
  const int nel = (int)(count - q0 < KI_EL ? count - q0 : KI_EL);
  {
    const int el = lane >> 3, gp = lane & 7;
    if (el < nel) {
      const int64_t e = first + q0 + el;
      int32_t nodes[8];
      double X[8][3];
      load_element(nn, ne, xyz, en, e, nodes, X);
      double G[3][8];
      const double dj = hex8_gp_G(X, gp, jmode, G, true);
      double *g = sG[w][el][gp];
#pragma unroll
      for (int p = 0; p < 3; ++p)
#pragma unroll
        for (int b = 0; b < 8; ++b) g[p * 8 + b] = G[p][b];
      g[24] = dj;
      if (gp == 0) {
        const double Ee = E[e], ve = nu[e];
        sLM[w][el][0] = ve * Ee / (1.0 + ve) / (1.0 - 2.0 * ve);
        sLM[w][el][1] = Ee / 2.0 / (1.0 + ve);
      }
    }
  }
  wave_lds_sync();
  const int a = lane >> 3, b = lane & 7;
  for (int el = 0; el < nel; ++el) {
    double m[3][3];
    if constexpr (MF) {
      const int li = lane & 15, lk = lane >> 4;
      ke_d4 t00 = {0.0, 0.0, 0.0, 0.0}, t01 = t00, t11 = t00;
#pragma unroll
      for (int s = 0; s < 2; ++s) {
        const double *g = sG[w][el][4 * s + lk];
        const double dj = g[24];
        const double g0 = g[li];
        const double g1 = li < 8 ? g[16 + (li & 7)] : 0.0;
        const double a0 = dj * g0, a1 = dj * g1;
        t00 = __builtin_amdgcn_mfma_f64_16x16x4f64(a0, g0, t00, 0, 0, 0);
        t01 = __builtin_amdgcn_mfma_f64_16x16x4f64(a0, g1, t01, 0, 0, 0);
        t11 = __builtin_amdgcn_mfma_f64_16x16x4f64(a1, g1, t11, 0, 0, 0);
      }
      // accumulator layout: lane l, element v holds row 4v + l/16, column l%16 of its tile
      double *M = sM[w];
#pragma unroll
      for (int v = 0; v < 4; ++v) {
        const int r = 4 * v + lk;
        M[r * KI_MS + li] = t00[v];
        if (li < 8) {
          M[r * KI_MS + 16 + li] = t01[v];
          M[(16 + li) * KI_MS + r] = t01[v];
          if (r < 8) M[(16 + r) * KI_MS + 16 + li] = t11[v];
        }
      }
      wave_lds_sync();
#pragma unroll
      for (int p = 0; p < 3; ++p)
#pragma unroll
        for (int q = 0; q < 3; ++q) m[p][q] = M[(8 * p + a) * KI_MS + 8 * q + b];
      wave_lds_sync();  // the next element rewrites M
    } else {
#pragma unroll
      for (int p = 0; p < 3; ++p)
#pragma unroll
        for (int q = 0; q < 3; ++q) m[p][q] = 0.0;
#pragma unroll
      for (int gp = 0; gp < 8; ++gp) {
        const double *g = sG[w][el][gp];
        const double dj = g[24];
        const double ga0 = dj * g[a], ga1 = dj * g[8 + a], ga2 = dj * g[16 + a];
        const double gb0 = g[b], gb1 = g[8 + b], gb2 = g[16 + b];
        m[0][0] = __builtin_fma(ga0, gb0, m[0][0]);
        m[0][1] = __builtin_fma(ga0, gb1, m[0][1]);
        m[0][2] = __builtin_fma(ga0, gb2, m[0][2]);
        m[1][0] = __builtin_fma(ga1, gb0, m[1][0]);
        m[1][1] = __builtin_fma(ga1, gb1, m[1][1]);
        m[1][2] = __builtin_fma(ga1, gb2, m[1][2]);
        m[2][0] = __builtin_fma(ga2, gb0, m[2][0]);
        m[2][1] = __builtin_fma(ga2, gb1, m[2][1]);
        m[2][2] = __builtin_fma(ga2, gb2, m[2][2]);
      }
    }
    const double lam = sLM[w][el][0], mu = sLM[w][el][1];
    const double tr = (m[0][0] + m[1][1]) + m[2][2];
    double *out = ke + (q0 + el) * 576 + (3 * a) * 24 + 3 * b;
#pragma unroll
    for (int d = 0; d < 3; ++d)
#pragma unroll
      for (int e = 0; e < 3; ++e) {
        const double sh = (d == e) ? m[e][d] + tr : m[e][d];
        out[d * 24 + e] = __builtin_fma(lam, m[d][e], mu * sh);
      }
  }
}

// CRS values of a hex8 / DOF=3 ZeroMatrix pattern (rows = node neighbour lists ncol expanded
// by DOF, FEMDomainClass.f90:10780-10836) from row-major element matrices: one
// SUB-lane subgroup per CRS row r = 3i + d, lanes over the row's entries; entry (r, 3j+d')
// sums Ke_e(3a+d, 3b+d') over the row node's incident elements e in ascending id, local nodes
// a (= i) and b (= j) ascending -- the order of the reference's element loop -- starting from
// 0.0, and is written once.
template <int SUB, int EPL>
__global__ void __launch_bounds__(256) k_gather_rows(int64_t nn, int64_t ne, const int32_t *__restrict__ en,
                                                     const int64_t *__restrict__ eoff,
                                                     const int32_t *__restrict__ elist,
                                                     const int64_t *__restrict__ noff,
                                                     const int32_t *__restrict__ ncol,
                                                     const int64_t *__restrict__ row_ptr,
                                                     const double *__restrict__ ke, double *__restrict__ val) {
  const int64_t g = blockIdx.x * (int64_t)blockDim.x + threadIdx.x;
  const int64_t r = g / SUB;
  const int lane = (int)(g % SUB);
  if (r >= 3 * nn) return;  // whole subgroups
  const int64_t i = r / 3;
  const int d = (int)(r - 3 * i);
  const int64_t base = row_ptr[r] - 1;
  const int64_t len = row_ptr[r + 1] - 1 - base;
  const int64_t s0 = eoff[i], s1 = eoff[i + 1];
  // up to SUB * EPL entries per pass (one pass for the <= 81 entries of a hex8 row): lane
  // entry t is q = q0 + t * SUB + lane, accumulated in registers over the element loop
  for (int64_t q0 = 0; q0 < len; q0 += SUB * EPL) {
    int32_t j1[EPL];
    int dc[EPL];
    double v[EPL];
#pragma unroll
    for (int t = 0; t < EPL; ++t) {
      const int64_t q = q0 + t * SUB + lane;
      // the row is the node's neighbour list expanded by DOF: entry q is DOF q % 3 of
      // neighbour q / 3 (node ids are >= 1, so 0 marks an entry past the row)
      j1[t] = q < len ? ncol[noff[i] + q / 3] + 1 : 0;
      dc[t] = (int)(q % 3);
      v[t] = 0.0;
    }
    for (int64_t s = s0; s < s1; ++s) {
      const int64_t e = elist[s];
      if (s > s0 && elist[s - 1] == e) continue;  // node listed twice in one element
      int32_t nodes[8];
#pragma unroll
      for (int b = 0; b < 8; ++b) nodes[b] = en[(int64_t)b * ne + e];
#pragma unroll 1
      for (int a = 0; a < 8; ++a) {
        if (nodes[a] != (int32_t)(i + 1)) continue;
        const double *kr = ke + e * 576 + (3 * a + d) * 24;
#pragma unroll
        for (int t = 0; t < EPL; ++t)
#pragma unroll
          for (int b = 0; b < 8; ++b)
            if (nodes[b] == j1[t]) v[t] = v[t] + kr[3 * b + dc[t]];
      }
    }
#pragma unroll
    for (int t = 0; t < EPL; ++t) {
      const int64_t q = q0 + t * SUB + lane;
      if (q < len) val[base + q] = v[t];
    }
  }
}

// The same sums per node instead of per row: one wavefront per node i gathers its three CRS
// rows (3 x 3 deg entries, lane l holding entries l, l + 64, ..., EPN of them), so the node's
// incident elements and their connectivity are loaded once (lane 8 el + b holds node b of
// element el of a chunk of 8, shared by shuffles) and the element-matrix loads of all three
// rows are in flight together.  Every entry still sums Ke_e(3a+d, 3b+d') over the incident
// elements in ascending id, local nodes a and b ascending, from 0.0: bitwise k_gather_rows.
// Needs 9 deg <= 64 EPN.
template <int EPN>
__global__ void __launch_bounds__(256) k_gather_nodes(int64_t nn, int64_t ne, const int32_t *__restrict__ en,
                                                      const int64_t *__restrict__ eoff,
                                                      const int32_t *__restrict__ elist,
                                                      const int64_t *__restrict__ noff,
                                                      const int32_t *__restrict__ ncol,
                                                      const int64_t *__restrict__ row_ptr,
                                                      const double *__restrict__ ke, double *__restrict__ val) {
  const int64_t i = (blockIdx.x * (int64_t)blockDim.x + threadIdx.x) >> 6;
  const int lane = threadIdx.x & 63;
  if (i >= nn) return;  // whole wavefronts
  const int64_t n0 = noff[i];
  const int len = 3 * (int)(noff[i + 1] - n0);  // entries of each of the node's rows
  const int64_t s0 = eoff[i], s1 = eoff[i + 1];
  int32_t jn[EPN];
  int dd[EPN], dq[EPN], dc[EPN];
  double v[EPN];
#pragma unroll
  for (int k = 0; k < EPN; ++k) {
    const int P = lane + 64 * k;
    const bool live = P < 3 * len;
    dd[k] = live ? P / len : 0;
    dq[k] = live ? P - dd[k] * len : 0;
    dc[k] = dq[k] % 3;
    jn[k] = live ? ncol[n0 + dq[k] / 3] + 1 : 0;  // node ids are >= 1: 0 matches no node
    v[k] = 0.0;
  }
  for (int64_t c0 = s0; c0 < s1; c0 += 8) {
    const int mc = (int)(s1 - c0 < 8 ? s1 - c0 : 8);
    int32_t el_id = 0, nv = 0;
    if (lane < 8 * mc) {
      el_id = elist[c0 + (lane >> 3)];
      nv = en[(int64_t)(lane & 7) * ne + el_id];
    }
    const int32_t prev0 = c0 > s0 ? elist[c0 - 1] : -1;  // the listing before this chunk
    const uint64_t self = __ballot(nv == (int32_t)(i + 1));  // (element, local node) pairs of node i
    for (int el = 0; el < mc; ++el) {
      const int32_t e = __shfl(el_id, 8 * el, 64);
      const int32_t ep = el > 0 ? __shfl(el_id, 8 * (el - 1), 64) : prev0;
      if (ep == e) continue;  // node listed twice in one element
      int32_t nodes[8];
#pragma unroll
      for (int b = 0; b < 8; ++b) nodes[b] = __shfl(nv, 8 * el + b, 64);
      const unsigned am = (unsigned)(self >> (8 * el)) & 0xffu;
      // each entry's local node b in this element (an element lists a node once but a
      // degenerate one may repeat it: then every match adds, in ascending b)
      int bk[EPN];
      bool multi = false;
#pragma unroll
      for (int k = 0; k < EPN; ++k) {
        int bb = -1, cnt = 0;
#pragma unroll
        for (int b = 0; b < 8; ++b)
          if (nodes[b] == jn[k]) {
            bb = bb < 0 ? b : bb;
            ++cnt;
          }
        bk[k] = bb;
        multi = multi || cnt > 1;
      }
#pragma unroll 1
      for (int a = 0; a < 8; ++a) {
        if (!((am >> a) & 1u)) continue;
        const double *kr = ke + (int64_t)e * 576 + (3 * a) * 24;
        if (!multi) {
          double ld[EPN];  // loads from in-bounds addresses, all in flight, then selected
#pragma unroll
          for (int k = 0; k < EPN; ++k) ld[k] = kr[dd[k] * 24 + 3 * (bk[k] < 0 ? 0 : bk[k]) + dc[k]];
#pragma unroll
          for (int k = 0; k < EPN; ++k)
            if (bk[k] >= 0) v[k] = v[k] + ld[k];
        } else {
#pragma unroll
          for (int k = 0; k < EPN; ++k)
#pragma unroll
            for (int b = 0; b < 8; ++b)
              if (nodes[b] == jn[k]) v[k] = v[k] + kr[dd[k] * 24 + 3 * b + dc[k]];
        }
      }
    }
  }
#pragma unroll
  for (int k = 0; k < EPN; ++k)
    if (lane + 64 * k < 3 * len) val[row_ptr[3 * i + dd[k]] - 1 + dq[k]] = v[k];
}

// N_a = 0.1250d0*(1 -+ xi)*(1 -+ eta)*(1 -+ zeta) at Gauss point gp (ShapeFunctionClass.f90:781-788)
__device__ inline double hex8_N(int a, int gp) {
  const double gx = (gp == 1 || gp == 2 || gp == 5 || gp == 6) ? GPV : -GPV;
  const double gy = (gp == 2 || gp == 3 || gp == 6 || gp == 7) ? GPV : -GPV;
  const double gz = (gp >= 4) ? GPV : -GPV;
  const double t0 = SX[a] > 0 ? 1.0 + gx : 1.0 - gx;
  const double t1 = SY[a] > 0 ? 1.0 + gy : 1.0 - gy;
  const double t2 = SZ[a] > 0 ? 1.0 + gz : 1.0 - gz;
  return ((0.125 * t0) * t1) * t2;
}

// MassVectorFEMDomain (FEMDomainClass.f90:10993-11073) entries of local node a:
// mv_k = rho * sum_gp (N_a accel_k) detJ (the other terms of matmul(Nmat, accel) are zeros)
__device__ inline void hex8_mass_node(const double (&X)[8][3], int a, double rho, double ax, double ay, double az,
                                      double &m0, double &m1, double &m2) {
  double G[3][8];
  m0 = 0.0;
  m1 = 0.0;
  m2 = 0.0;
#pragma unroll 1
  for (int gp = 0; gp < 8; ++gp) {
    const double detJ = hex8_gp_G(X, gp, 0, G, false);
    const double Na = hex8_N(a, gp);
    m0 = m0 + (Na * ax) * detJ;
    m1 = m1 + (Na * ay) * detJ;
    m2 = m2 + (Na * az) * detJ;
  }
  m0 = rho * m0;
  m1 = rho * m1;
  m2 = rho * m2;
}

// element mass vectors of elements first..first+count-1: one thread per (element, node)
__global__ void k_mass_vec_batch(int64_t nn, int64_t ne, const double *__restrict__ xyz,
                                 const int32_t *__restrict__ en, double rho, double ax, double ay, double az,
                                 int64_t first, int64_t count, double *__restrict__ out) {
  const int64_t t = blockIdx.x * (int64_t)blockDim.x + threadIdx.x;
  if (t >= count * 8) return;
  const int64_t q = t / 8;
  const int a = (int)(t - q * 8);
  int32_t nodes[8];
  double X[8][3];
  load_element(nn, ne, xyz, en, first + q, nodes, X);
  double m0, m1, m2;
  hex8_mass_node(X, a, rho, ax, ay, az, m0, m1, m2);
  out[q * 24 + 3 * a] = m0;
  out[q * 24 + 3 * a + 1] = m1;
  out[q * 24 + 3 * a + 2] = m2;
}

// MassMatrixFEMDomain (FEMDomainClass.f90:10874-10945), DOF = 3, row block of local node a:
// m_b = sum_gp (N_a N_b) detJ for the 8 local nodes b; the matrix entry (3a+d, 3b+k) is
// rho * m_b when d == k and rho * 0 otherwise (products with zeros in matmul(Nmat, Nmat^T))
__device__ inline void hex8_mass_row(const double (&X)[8][3], int a, double (&m)[8]) {
  double G[3][8];
#pragma unroll
  for (int b = 0; b < 8; ++b) m[b] = 0.0;
#pragma unroll 1
  for (int gp = 0; gp < 8; ++gp) {
    const double detJ = hex8_gp_G(X, gp, 0, G, false);
    const double Na = hex8_N(a, gp);
#pragma unroll
    for (int b = 0; b < 8; ++b) m[b] = m[b] + (Na * hex8_N(b, gp)) * detJ;
  }
}

// dense element mass matrices of elements first..first+count-1 (24x24 column-major): one
// thread per (element, local node a) writes the three rows of a
__global__ void k_mass_mat_batch(int64_t nn, int64_t ne, const double *__restrict__ xyz,
                                 const int32_t *__restrict__ en, const double *__restrict__ rho, int srho,
                                 int64_t first, int64_t count, double *__restrict__ out) {
  const int64_t t = blockIdx.x * (int64_t)blockDim.x + threadIdx.x;
  if (t >= count * 8) return;
  const int64_t q = t / 8;
  const int a = (int)(t - q * 8);
  const int64_t e = first + q;
  int32_t nodes[8];
  double X[8][3];
  load_element(nn, ne, xyz, en, e, nodes, X);
  double m[8];
  hex8_mass_row(X, a, m);
  const double r = rho[srho ? e : 0];
  double *o = out + q * 576;
  for (int b = 0; b < 8; ++b)
    for (int k = 0; k < 3; ++k)
      for (int d = 0; d < 3; ++d) o[(3 * b + k) * 24 + 3 * a + d] = d == k ? r * m[b] : r * 0.0;
}

// MassMatrix_as_CRS_FEMDomain (FEMDomainClass.f90:10364-10467): one thread per CRS row, the
// element contributions gathered in ascending element order
__global__ void __launch_bounds__(128) k_mass_rows(
    int64_t nn, int64_t ne, const double *__restrict__ xyz, const int32_t *__restrict__ en,
    const double *__restrict__ rho, int srho, const int64_t *__restrict__ eoff, const int32_t *__restrict__ elist,
    const int64_t *__restrict__ noff, const int32_t *__restrict__ ncol, const int64_t *__restrict__ row_ptr,
    double *__restrict__ val) {
  const int64_t r = blockIdx.x * (int64_t)blockDim.x + threadIdx.x;
  if (r >= 3 * nn) return;
  const int64_t i = r / 3;
  const int d = (int)(r - 3 * i);
  const int64_t base = row_ptr[r] - 1;
  const int64_t len = row_ptr[r + 1] - 1 - base;
  for (int64_t k = 0; k < len; ++k) val[base + k] = 0.0;
  const int64_t c0 = noff[i];
  const int32_t deg = (int32_t)(noff[i + 1] - c0);
  const int64_t s0 = eoff[i], s1 = eoff[i + 1];
  for (int64_t s = s0; s < s1; ++s) {
    const int64_t e = elist[s];
    if (s > s0 && elist[s - 1] == e) continue;  // node listed twice in one element
    int32_t nodes[8];
    double X[8][3];
    load_element(nn, ne, xyz, en, e, nodes, X);
    const double re = rho[srho ? e : 0];
    for (int a = 0; a < 8; ++a) {
      if (nodes[a] != (int32_t)(i + 1)) continue;
      double m[8];
      hex8_mass_row(X, a, m);
      for (int b = 0; b < 8; ++b) {
        int32_t p = 0;
        const int32_t target = nodes[b] - 1;
        for (int32_t q = 0; q < deg; ++q)
          if (ncol[c0 + q] == target) {
            p = q;
            break;
          }
        double *dst = val + base + 3 * (int64_t)p;
        dst[0] = dst[0] + (d == 0 ? re * m[b] : re * 0.0);
        dst[1] = dst[1] + (d == 1 ? re * m[b] : re * 0.0);
        dst[2] = dst[2] + (d == 2 ? re * m[b] : re * 0.0);
      }
    }
  }
}

// rhs += sum over elements (ascending id) of their mass vectors: one thread per node
__global__ void k_body_force(int64_t nn, int64_t ne, const double *__restrict__ xyz, const int32_t *__restrict__ en,
                             const int64_t *__restrict__ eoff, const int32_t *__restrict__ elist,
                             const double *__restrict__ rho, double rho_all, double ax, double ay, double az,
                             double *__restrict__ rhs) {
  const int64_t i = blockIdx.x * (int64_t)blockDim.x + threadIdx.x;
  if (i >= nn) return;
  double f0 = rhs[3 * i], f1 = rhs[3 * i + 1], f2 = rhs[3 * i + 2];
  const int64_t s0 = eoff[i], s1 = eoff[i + 1];
  for (int64_t s = s0; s < s1; ++s) {
    const int64_t e = elist[s];
    if (s > s0 && elist[s - 1] == e) continue;
    int32_t nodes[8];
    double X[8][3];
    load_element(nn, ne, xyz, en, e, nodes, X);
    const double r = rho ? rho[e] : rho_all;
    for (int a = 0; a < 8; ++a) {
      if (nodes[a] != (int32_t)(i + 1)) continue;
      double m0, m1, m2;
      hex8_mass_node(X, a, r, ax, ay, az, m0, m1, m2);
      f0 = f0 + m0;
      f1 = f1 + m1;
      f2 = f2 + m2;
    }
  }
  rhs[3 * i] = f0;
  rhs[3 * i + 1] = f1;
  rhs[3 * i + 2] = f2;
}

// StressMatrixFEMDomain without GaussPoint (FEMDomainClass.f90:11912-12032): the cell
// average over the 8 Gauss points of Stressvec = matmul(Dmat, matmul(Bmat, ElemDisp)),
// accumulated into the 3x3 tensor in the reference's order.  Sums skip the structural
// zeros of B and D (adding a signed zero does not change a sum that starts at +0).
// out[e*9 + 3j + i] = sigma(i+1, j+1); E, nu per element (stride 1) or one value (stride 0).
__global__ void __launch_bounds__(128) k_cell_stress(int64_t nn, int64_t ne, const double *__restrict__ xyz,
                                                     const int32_t *__restrict__ en, const double *__restrict__ disp,
                                                     const double *__restrict__ E, int sE,
                                                     const double *__restrict__ nu, int snu, int jmode,
                                                     double *__restrict__ out) {
  const int64_t e = blockIdx.x * (int64_t)blockDim.x + threadIdx.x;
  if (e >= ne) return;
  int32_t nodes[8];
  double X[8][3];
  load_element(nn, ne, xyz, en, e, nodes, X);
  double u[24];
#pragma unroll
  for (int a = 0; a < 8; ++a)
#pragma unroll
    for (int d = 0; d < 3; ++d) u[3 * a + d] = disp[3 * (int64_t)(nodes[a] - 1) + d];
  const double Ee = E[sE ? e : 0], ve = nu[snu ? e : 0];
  const double mu = Ee / 2.0 / (1.0 + ve);
  const double lam = ve * Ee / (1.0 + ve) / (1.0 - 2.0 * ve);
  const double dg = 2.0 * mu + lam;
  double S00 = 0.0, S11 = 0.0, S22 = 0.0, S01 = 0.0, S12 = 0.0, S02 = 0.0, S10 = 0.0, S21 = 0.0, S20 = 0.0;
#pragma unroll 1
  for (int gp = 0; gp < 8; ++gp) {
    double G[3][8];
    hex8_gp_G(X, gp, jmode, G, true);
    double e0 = 0.0, e1 = 0.0, e2 = 0.0, e3 = 0.0, e4 = 0.0, e5 = 0.0;
#pragma unroll
    for (int q = 0; q < 8; ++q) {
      e0 = e0 + G[0][q] * u[3 * q];
      e1 = e1 + G[1][q] * u[3 * q + 1];
      e2 = e2 + G[2][q] * u[3 * q + 2];
      e3 = e3 + G[1][q] * u[3 * q];
      e3 = e3 + G[0][q] * u[3 * q + 1];
      e4 = e4 + G[2][q] * u[3 * q + 1];
      e4 = e4 + G[1][q] * u[3 * q + 2];
      e5 = e5 + G[2][q] * u[3 * q];
      e5 = e5 + G[0][q] * u[3 * q + 2];
    }
    double v0 = 0.0, v1 = 0.0, v2 = 0.0;
    v0 = v0 + dg * e0;
    v0 = v0 + lam * e1;
    v0 = v0 + lam * e2;
    v1 = v1 + lam * e0;
    v1 = v1 + dg * e1;
    v1 = v1 + lam * e2;
    v2 = v2 + lam * e0;
    v2 = v2 + lam * e1;
    v2 = v2 + dg * e2;
    const double v3 = 0.0 + mu * e3, v4 = 0.0 + mu * e4, v5 = 0.0 + mu * e5;
    S00 = S00 + v0;
    S11 = S11 + v1;
    S22 = S22 + v2;
    S01 = S01 + v3;
    S12 = S12 + v4;
    S02 = S02 + v5;
    S10 = S10 + v3;
    S21 = S21 + v4;
    S20 = S20 + v5;
  }
  double *o = out + e * 9;
  o[0] = S00 / 8.0;
  o[1] = S10 / 8.0;
  o[2] = S20 / 8.0;
  o[3] = S01 / 8.0;
  o[4] = S11 / 8.0;
  o[5] = S21 / 8.0;
  o[6] = S02 / 8.0;
  o[7] = S12 / 8.0;
  o[8] = S22 / 8.0;
}

// ---- adjoint sensitivity: out[e] = sum_l lam_e(l)^T Ke(E_e) u_e(l) ----------------------
// Ke is linear in E, so this is d/d(log E_e) of sum_l lam(l)^T K u(l).  The element matrices
// of a chunk of elements come from k_ke_batch (same arithmetic as the assembly); here one
// wavefront per element stages its Ke in LDS and lane l contracts load case l of the
// interleaved vectors (dof i, case l at i*R + l).  Fixed DOFs (fixed != NULL) are dropped
// from both sides: the eliminated system's matrix is P K P + (I - P).
__global__ void __launch_bounds__(64) k_ke_contract(int64_t ne, const int32_t *__restrict__ en, int64_t first,
                                                    int64_t count, const double *__restrict__ ke, int R,
                                                    const double *__restrict__ U, const double *__restrict__ Lam,
                                                    const uint8_t *__restrict__ fixed, double *__restrict__ out) {
  __shared__ double K[24][25];
  __shared__ int64_t dofs[24];  // DOF of element row r, -1 when fixed
  const int lane = threadIdx.x;
  for (int64_t q = blockIdx.x; q < count; q += gridDim.x) {
    const int64_t e = first + q;
    const double *kq = ke + q * 576;  // column-major 24 x 24
    for (int t = lane; t < 576; t += 64) K[t % 24][t / 24] = kq[t];
    if (lane < 24) {
      const int64_t dof = 3 * (int64_t)(en[(int64_t)(lane / 3) * ne + e] - 1) + lane % 3;
      dofs[lane] = (fixed && fixed[dof]) ? -1 : dof;
    }
    __syncthreads();
    double s = 0.0;
    for (int l = lane; l < R; l += 64) {
      double u[24];
#pragma unroll
      for (int c = 0; c < 24; ++c) {
        const int64_t dof = dofs[c];
        u[c] = dof < 0 ? 0.0 : U[dof * R + l];
      }
#pragma unroll 1
      for (int r = 0; r < 24; ++r) {
        const int64_t dof = dofs[r];
        if (dof < 0) continue;
        double t = 0.0;
#pragma unroll
        for (int c = 0; c < 24; ++c) t += K[r][c] * u[c];
        s += Lam[dof * R + l] * t;
      }
    }
#pragma unroll
    for (int off = 32; off > 0; off >>= 1) s += __shfl_xor(s, off, 64);
    if (lane == 0) out[e] = s;
    __syncthreads();
  }
}

// ---- batched setMatrix / setVector: node -> (call index, local node) lists ---------
__global__ void k_batch_count(int64_t count, int nne, int64_t ne, const int32_t *__restrict__ en,
                              const int32_t *__restrict__ ids, int32_t *__restrict__ cnt) {
  const int64_t t = blockIdx.x * (int64_t)blockDim.x + threadIdx.x;
  if (t >= count * nne) return;
  const int64_t q = t / nne;
  const int a = (int)(t - q * nne);
  atomicAdd(&cnt[en[(int64_t)a * ne + (ids[q] - 1)] - 1], 1);
}

__global__ void k_batch_fill(int64_t count, int nne, int64_t ne, const int32_t *__restrict__ en,
                             const int32_t *__restrict__ ids, const int64_t *__restrict__ off,
                             int32_t *__restrict__ fill, int64_t *__restrict__ list) {
  const int64_t t = blockIdx.x * (int64_t)blockDim.x + threadIdx.x;
  if (t >= count * nne) return;
  const int64_t q = t / nne;
  const int a = (int)(t - q * nne);
  const int32_t node = en[(int64_t)a * ne + (ids[q] - 1)] - 1;
  const int32_t slot = atomicAdd(&fill[node], 1);
  list[off[node] + slot] = t;  // t = q*nne + a sorts by call order, then local node
}

__global__ void k_batch_sort(int64_t nn, const int64_t *__restrict__ off, int64_t *__restrict__ list) {
  const int64_t i = blockIdx.x * (int64_t)blockDim.x + threadIdx.x;
  if (i >= nn) return;
  const int64_t s0 = off[i], s1 = off[i + 1];
  for (int64_t a = s0 + 1; a < s1; ++a) {
    const int64_t v = list[a];
    int64_t b = a - 1;
    while (b >= s0 && list[b] > v) {
      list[b + 1] = list[b];
      --b;
    }
    list[b + 1] = v;
  }
}

// setValueFEMSolver matrix branch (FEMSolverClass.f90:1020-1057) for the rows of one
// node: linear search of the row for every column, add in call order.
__global__ void k_batch_add_mats(int64_t nn, int nne, int dof, int64_t ne,
                                 const int32_t *__restrict__ en, const int32_t *__restrict__ ids,
                                 const int64_t *__restrict__ off, const int64_t *__restrict__ list,
                                 const double *__restrict__ ke, const int64_t *__restrict__ row_ptr,
                                 const int32_t *__restrict__ col_idx, double *__restrict__ val,
                                 int32_t *__restrict__ err) {
  const int64_t r = blockIdx.x * (int64_t)blockDim.x + threadIdx.x;
  if (r >= (int64_t)dof * nn) return;
  const int64_t i = r / dof;
  const int j = (int)(r - i * dof);
  const int nk = nne * dof;
  const int64_t rb = row_ptr[r] - 1, re = row_ptr[r + 1] - 1;
  for (int64_t s = off[i]; s < off[i + 1]; ++s) {
    const int64_t q = list[s] / nne;
    const int a = (int)(list[s] - q * nne);
    const int64_t e = ids[q] - 1;
    const double *K = ke + q * (int64_t)nk * nk;
    for (int b = 0; b < nne; ++b)
      for (int jj = 0; jj < dof; ++jj) {
        const int32_t col = (int32_t)((int64_t)dof * (en[(int64_t)b * ne + e] - 1) + jj + 1);
        int64_t id = -1;
        for (int64_t k = rb; k < re; ++k)
          if (col_idx[k] == col) {
            id = k;
            break;
          }
        if (id < 0) {
          atomicOr(err, 1);
          continue;
        }
        val[id] = val[id] + K[(int64_t)(dof * b + jj) * nk + (dof * a + j)];
      }
  }
}

__global__ void k_batch_add_vecs(int64_t nn, int nne, int dof, const int64_t *__restrict__ off,
                                 const int64_t *__restrict__ list, const double *__restrict__ vec,
                                 double *__restrict__ rhs) {
  const int64_t r = blockIdx.x * (int64_t)blockDim.x + threadIdx.x;
  if (r >= (int64_t)dof * nn) return;
  const int64_t i = r / dof;
  const int j = (int)(r - i * dof);
  double acc = rhs[r];
  for (int64_t s = off[i]; s < off[i + 1]; ++s) {
    const int64_t q = list[s] / nne;
    const int a = (int)(list[s] - q * nne);
    acc = acc + vec[q * (int64_t)nne * dof + dof * a + j];
  }
  rhs[r] = acc;
}

// ---- traction (TractionVector_by_elemFEMDomain, FEMDomainClass.f90:17210-17292) ----
__device__ __constant__ int FACET[6][4] = {{4, 3, 2, 1}, {1, 2, 6, 5}, {2, 3, 7, 6},
                                           {3, 4, 8, 7}, {4, 1, 5, 8}, {5, 6, 7, 8}};

__device__ inline double dot3(const double *a, const double *b) {
  double s = 0.0;
  for (int i = 0; i < 3; ++i) s = s + a[i] * b[i];
  return s;
}

__device__ inline double tri_area(const double *A, const double *B, const double *C) {
  double ab[3], ac[3];
  for (int i = 0; i < 3; ++i) {
    ab[i] = B[i] - A[i];
    ac[i] = C[i] - A[i];
  }
  return 0.5 * sqrt(dot3(ab, ab) * dot3(ac, ac) - dot3(ab, ac) * dot3(ab, ac));
}

// one thread per node: sum, in ascending element then facet order, the quarter facet
// forces of every facet (fully inside the open box) that holds this node
// MULTI: node ranges node_off[m] .. node_off[m+1]-1 (0-based) each get their own box and
// pressure (the meshes of a batch); nodes outside every range are left alone
template <bool MULTI>
__global__ void k_traction(int64_t nn, int64_t ne, const double *__restrict__ xyz,
                           const int32_t *__restrict__ en, const int64_t *__restrict__ eoff,
                           const int32_t *__restrict__ elist, double bx0, double bx1, double by0,
                           double by1, double bz0, double bz1, double pressure_one, int32_t nbox,
                           const double *__restrict__ boxes, const double *__restrict__ pressures,
                           const int64_t *__restrict__ node_off, double *__restrict__ rhs) {
  const int64_t i = blockIdx.x * (int64_t)blockDim.x + threadIdx.x;
  if (i >= nn) return;
  double box[6] = {bx0, bx1, by0, by1, bz0, bz1};
  double pressure = pressure_one;
  if (MULTI) {
    if (i < node_off[0] || i >= node_off[nbox]) return;
    int32_t lo = 0, hi = nbox;  // largest m with node_off[m] <= i
    while (hi - lo > 1) {
      const int32_t mid = (lo + hi) >> 1;
      if (node_off[mid] <= i) lo = mid; else hi = mid;
    }
    for (int k = 0; k < 6; ++k) box[k] = boxes[6 * lo + k];
    pressure = pressures[lo];
  }
  double f0 = rhs[3 * i], f1 = rhs[3 * i + 1], f2 = rhs[3 * i + 2];
  const int64_t s0 = eoff[i], s1 = eoff[i + 1];
  for (int64_t s = s0; s < s1; ++s) {
    const int64_t e = elist[s];
    if (s > s0 && elist[s - 1] == e) continue;
    int32_t nodes[8];
    double P[8][3];
    load_element(nn, ne, xyz, en, e, nodes, P);
    int in[8];
    for (int a = 0; a < 8; ++a) {
      in[a] = 1;
      for (int dd = 0; dd < 3; ++dd)
        if (box[2 * dd] >= P[a][dd] || box[2 * dd + 1] <= P[a][dd]) in[a] = 0;
    }
    for (int f = 0; f < 6; ++f) {
      int cnt = 0;
      for (int j = 0; j < 4; ++j) cnt += in[FACET[f][j] - 1];
      if (cnt != 4) continue;
      const double *A = P[FACET[f][0] - 1], *B = P[FACET[f][1] - 1];
      const double *C = P[FACET[f][2] - 1], *Dq = P[FACET[f][3] - 1];
      const double area = tri_area(A, B, C) + tri_area(B, C, Dq);
      double u[3], v[3], n[3];
      for (int k = 0; k < 3; ++k) {
        u[k] = B[k] - A[k];
        v[k] = C[k] - A[k];
      }
      n[0] = u[1] * v[2] - u[2] * v[1];
      n[1] = u[2] * v[0] - u[0] * v[2];
      n[2] = u[0] * v[1] - u[1] * v[0];
      const double nrm = sqrt(dot3(n, n));
      const double fx = pressure * area * (n[0] / nrm);
      const double fy = pressure * area * (n[1] / nrm);
      const double fz = pressure * area * (n[2] / nrm);
      for (int j = 0; j < 4; ++j)
        if (nodes[FACET[f][j] - 1] == (int32_t)(i + 1)) {
          f0 = f0 + 0.25 * fx;
          f1 = f1 + 0.25 * fy;
          f2 = f2 + 0.25 * fz;
        }
    }
  }
  rhs[3 * i] = f0;
  rhs[3 * i + 1] = f1;
  rhs[3 * i + 2] = f2;
}

}  // namespace

int pf_assemble_hex8(pf_ctx *ctx, const NodeGraph &g, const double *E, const double *nu, int jmode,
                     const int64_t *row_ptr, double *val, int ke_kernel) {
  if (g.nne != 8 || !g.cols_built || !g.has_coords) {
    pf_set_error("hex8 assembly needs an 8-node mesh with coordinates and its graph built");
    return PF_INVALID;
  }
  if (ke_kernel < 0 || ke_kernel > 2) {
    pf_set_error("hex8 assembly: element kernel must be 0 (reference order), 1 (isotropic split, VALU) "
                 "or 2 (isotropic split, MFMA)");
    return PF_INVALID;
  }
  if (g.nn == 0) return PF_OK;
  // element matrices once per element (row-major scratch of 4.6 KB per element), then one
  // gather per CRS row; without room for the scratch, the one-pass row kernel
  if (g.ne > 0 && g.kebuf.ensure((size_t)g.ne * 576) == PF_OK) {
    const int64_t waves = (g.ne + KI_EL - 1) / KI_EL;
    if (ke_kernel == 0)
      k_ke_elem<false><<<pf_blocks(g.ne * 64, 256, INT32_MAX), 256, 0, ctx->stream>>>(
          g.nn, g.ne, g.nodcoord.p, g.elemnod.p, E, nu, jmode, 0, g.ne, g.kebuf.p);
    else if (ke_kernel == 1)
      k_ke_iso<false><<<pf_blocks(waves * 64, 64 * KI_W, INT32_MAX), 64 * KI_W, 0, ctx->stream>>>(
          g.nn, g.ne, g.nodcoord.p, g.elemnod.p, E, nu, jmode, 0, g.ne, g.kebuf.p);
    else
      k_ke_iso<true><<<pf_blocks(waves * 64, 64 * KI_W, INT32_MAX), 64 * KI_W, 0, ctx->stream>>>(
          g.nn, g.ne, g.nodcoord.p, g.elemnod.p, E, nu, jmode, 0, g.ne, g.kebuf.p);
    if (9 * g.maxdeg <= 64 * 4)  // a hex8 node has at most 27 neighbours
      k_gather_nodes<4><<<pf_blocks(g.nn * 64, 256, INT32_MAX), 256, 0, ctx->stream>>>(
          g.nn, g.ne, g.elemnod.p, g.eoff.p, g.elist.p, g.noff.p, g.ncol.p, row_ptr, g.kebuf.p, val);
    else
      k_gather_rows<16, 6><<<pf_blocks(3 * g.nn * 16, 256, INT32_MAX), 256, 0, ctx->stream>>>(
          g.nn, g.ne, g.elemnod.p, g.eoff.p, g.elist.p, g.noff.p, g.ncol.p, row_ptr, g.kebuf.p, val);
  } else {
    k_assemble_rows<<<pf_blocks(3 * g.nn, 128), 128, 0, ctx->stream>>>(
        g.nn, g.ne, g.nodcoord.p, g.elemnod.p, E, nu, jmode, g.eoff.p, g.elist.p, g.noff.p, g.ncol.p,
        row_ptr, val);
  }
  PF_HIP(hipGetLastError());
  return PF_OK;
}

int pf_hex8_ke_batch(pf_ctx *ctx, const NodeGraph &g, const double *E, const double *nu, int jmode,
                     int64_t first, int64_t count, double *ke) {
  if (g.nne != 8 || !g.has_coords) {
    pf_set_error("hex8 element stiffness needs nne = 8 and coordinates");
    return PF_INVALID;
  }
  if (count <= 0) return PF_OK;
  k_ke_elem<true><<<pf_blocks(count * 64, 256, INT32_MAX), 256, 0, ctx->stream>>>(g.nn, g.ne, g.nodcoord.p,
                                                                                 g.elemnod.p, E, nu, jmode, first,
                                                                                 count, ke);
  PF_HIP(hipGetLastError());
  return PF_OK;
}

static int batch_lists(pf_ctx *ctx, const NodeGraph &g, int64_t count, const int32_t *ids,
                       DevBuf<int64_t> &off, DevBuf<int64_t> &list) {
  DevBuf<int32_t> cnt;
  DevBuf<int64_t> tmp;
  const int64_t total = count * g.nne;
  PF_CHECK(cnt.ensure((size_t)g.nn + 1));
  PF_CHECK(off.ensure((size_t)g.nn + 1));
  PF_CHECK(list.ensure((size_t)total + 1));
  PF_HIP(hipMemsetAsync(cnt.p, 0, sizeof(int32_t) * (g.nn + 1), ctx->stream));
  k_batch_count<<<pf_blocks(total, 256), 256, 0, ctx->stream>>>(count, g.nne, g.ne, g.elemnod.p, ids, cnt.p);
  PF_CHECK(pf_scan_i32_to_i64(ctx, cnt.p, g.nn, off.p, tmp));
  PF_HIP(hipMemsetAsync(cnt.p, 0, sizeof(int32_t) * (g.nn + 1), ctx->stream));
  k_batch_fill<<<pf_blocks(total, 256), 256, 0, ctx->stream>>>(count, g.nne, g.ne, g.elemnod.p, ids,
                                                               off.p, cnt.p, list.p);
  k_batch_sort<<<pf_blocks(g.nn, 256), 256, 0, ctx->stream>>>(g.nn, off.p, list.p);
  PF_HIP(hipGetLastError());
  PF_HIP(hipStreamSynchronize(ctx->stream));
  return PF_OK;
}

int pf_add_element_mats(pf_ctx *ctx, const NodeGraph &g, int dof, int64_t count, const int32_t *ids,
                        const double *ke, const int64_t *row_ptr, const int32_t *col_idx, double *val,
                        int32_t *err) {
  if (count <= 0 || g.nn == 0) return PF_OK;
  DevBuf<int64_t> off, list;
  PF_CHECK(batch_lists(ctx, g, count, ids, off, list));
  const int64_t rows = (int64_t)dof * g.nn;
  k_batch_add_mats<<<pf_blocks(rows, 128), 128, 0, ctx->stream>>>(g.nn, g.nne, dof, g.ne, g.elemnod.p, ids,
                                                                  off.p, list.p, ke, row_ptr,
                                                                  col_idx, val, err);
  PF_HIP(hipGetLastError());
  PF_HIP(hipStreamSynchronize(ctx->stream));
  return PF_OK;
}

int pf_add_element_vecs(pf_ctx *ctx, const NodeGraph &g, int dof, int64_t count, const int32_t *ids,
                        const double *vec, double *rhs) {
  if (count <= 0 || g.nn == 0) return PF_OK;
  DevBuf<int64_t> off, list;
  PF_CHECK(batch_lists(ctx, g, count, ids, off, list));
  const int64_t rows = (int64_t)dof * g.nn;
  k_batch_add_vecs<<<pf_blocks(rows, 256), 256, 0, ctx->stream>>>(g.nn, g.nne, dof, off.p, list.p, vec, rhs);
  PF_HIP(hipGetLastError());
  PF_HIP(hipStreamSynchronize(ctx->stream));
  return PF_OK;
}

int pf_traction_box(pf_ctx *ctx, const NodeGraph &g, const double *box, double pressure, double *rhs) {
  if (g.nne != 8 || !g.built || !g.has_coords) {
    pf_set_error("traction needs an 8-node mesh with coordinates and its node->element lists");
    return PF_INVALID;
  }
  if (g.nn == 0) return PF_OK;
  k_traction<false><<<pf_blocks(g.nn, 128), 128, 0, ctx->stream>>>(
      g.nn, g.ne, g.nodcoord.p, g.elemnod.p, g.eoff.p, g.elist.p, box[0], box[1], box[2], box[3], box[4], box[5],
      pressure, 0, nullptr, nullptr, nullptr, rhs);
  PF_HIP(hipGetLastError());
  return PF_OK;
}

int pf_elem_bilinear(pf_ctx *ctx, const NodeGraph &g, const double *E, const double *nu, int jmode, int R,
                     const double *U, const double *Lam, const uint8_t *fixed, double *out) {
  if (g.nne != 8 || !g.has_coords) {
    pf_set_error("element sensitivities need an 8-node mesh with coordinates");
    return PF_INVALID;
  }
  if (g.ne == 0) return PF_OK;
  const int64_t chunk = std::min<int64_t>(g.ne, 32768);  // 151 MB of element matrices per chunk
  DevBuf<double> ke;
  PF_CHECK(ke.ensure((size_t)chunk * 576));
  for (int64_t first = 0; first < g.ne; first += chunk) {
    const int64_t count = std::min(chunk, g.ne - first);
    k_ke_elem<true><<<pf_blocks(count * 64, 256, INT32_MAX), 256, 0, ctx->stream>>>(
        g.nn, g.ne, g.nodcoord.p, g.elemnod.p, E, nu, jmode, first, count, ke.p);
    k_ke_contract<<<(unsigned)std::min<int64_t>(count, 65536), 64, 0, ctx->stream>>>(g.ne, g.elemnod.p, first, count,
                                                                                     ke.p, R, U, Lam, fixed, out);
    PF_HIP(hipGetLastError());
  }
  PF_HIP(hipStreamSynchronize(ctx->stream));
  return PF_OK;
}

int pf_hex8_mass_batch(pf_ctx *ctx, const NodeGraph &g, double rho, const double *accel, int64_t first,
                       int64_t count, double *out) {
  if (g.nne != 8 || !g.has_coords) {
    pf_set_error("mass vectors need an 8-node mesh with coordinates");
    return PF_INVALID;
  }
  if (count <= 0) return PF_OK;
  k_mass_vec_batch<<<pf_blocks(count * 8, 128), 128, 0, ctx->stream>>>(g.nn, g.ne, g.nodcoord.p, g.elemnod.p, rho,
                                                                      accel[0], accel[1], accel[2], first, count, out);
  PF_HIP(hipGetLastError());
  return PF_OK;
}

int pf_hex8_mass_mat(pf_ctx *ctx, const NodeGraph &g, const double *rho, int srho, int64_t first, int64_t count,
                     double *out) {
  if (g.nne != 8 || !g.has_coords) {
    pf_set_error("mass matrices need an 8-node mesh with coordinates");
    return PF_INVALID;
  }
  if (count <= 0) return PF_OK;
  k_mass_mat_batch<<<pf_blocks(count * 8, 128), 128, 0, ctx->stream>>>(g.nn, g.ne, g.nodcoord.p, g.elemnod.p, rho,
                                                                      srho, first, count, out);
  PF_HIP(hipGetLastError());
  return PF_OK;
}

int pf_assemble_mass(pf_ctx *ctx, const NodeGraph &g, const double *rho, int srho, const int64_t *row_ptr,
                     double *val) {
  if (g.nne != 8 || !g.cols_built || !g.has_coords) {
    pf_set_error("mass assembly needs an 8-node mesh with coordinates and its graph built");
    return PF_INVALID;
  }
  if (g.nn == 0) return PF_OK;
  k_mass_rows<<<pf_blocks(3 * g.nn, 128), 128, 0, ctx->stream>>>(g.nn, g.ne, g.nodcoord.p, g.elemnod.p, rho, srho,
                                                                g.eoff.p, g.elist.p, g.noff.p, g.ncol.p, row_ptr, val);
  PF_HIP(hipGetLastError());
  return PF_OK;
}

int pf_body_force(pf_ctx *ctx, const NodeGraph &g, const double *rho_dev, double rho_all, const double *accel,
                  double *rhs) {
  if (g.nne != 8 || !g.built || !g.has_coords) {
    pf_set_error("body force needs an 8-node mesh with coordinates and its node->element lists");
    return PF_INVALID;
  }
  if (g.nn == 0) return PF_OK;
  k_body_force<<<pf_blocks(g.nn, 128), 128, 0, ctx->stream>>>(g.nn, g.ne, g.nodcoord.p, g.elemnod.p, g.eoff.p,
                                                              g.elist.p, rho_dev, rho_all, accel[0], accel[1],
                                                              accel[2], rhs);
  PF_HIP(hipGetLastError());
  return PF_OK;
}

int pf_cell_stress(pf_ctx *ctx, const NodeGraph &g, const double *disp, const double *E, int sE, const double *nu,
                   int snu, int jmode, double *out) {
  if (g.nne != 8 || !g.has_coords) {
    pf_set_error("stress recovery needs an 8-node mesh with coordinates");
    return PF_INVALID;
  }
  if (g.ne == 0) return PF_OK;
  k_cell_stress<<<pf_blocks(g.ne, 128), 128, 0, ctx->stream>>>(g.nn, g.ne, g.nodcoord.p, g.elemnod.p, disp, E, sE,
                                                               nu, snu, jmode, out);
  PF_HIP(hipGetLastError());
  return PF_OK;
}

int pf_traction_boxes(pf_ctx *ctx, const NodeGraph &g, int32_t count, const int64_t *node_off_dev,
                      const double *boxes_dev, const double *pressures_dev, double *rhs) {
  if (g.nne != 8 || !g.built || !g.has_coords) {
    pf_set_error("traction needs an 8-node mesh with coordinates and its node->element lists");
    return PF_INVALID;
  }
  if (g.nn == 0 || count <= 0) return PF_OK;
  k_traction<true><<<pf_blocks(g.nn, 128), 128, 0, ctx->stream>>>(
      g.nn, g.ne, g.nodcoord.p, g.elemnod.p, g.eoff.p, g.elist.p, 0.0, 0.0, 0.0, 0.0, 0.0, 0.0, 0.0, count,
      boxes_dev, pressures_dev, node_off_dev, rhs);
  PF_HIP(hipGetLastError());
  return PF_OK;
}

// ---- overset penalty coupling (Gauss-point projection) ---------------------------------
// FEMDomain%overset(algorithm=GPP) (FEMDomainClass.f90:14606-14700) pairs every Gauss point
// of the own mesh with the element of the other mesh that contains it:
//   Mesh%nearestElementID (MeshClass.f90:8611-8665): inside the other mesh's node bounding
//   box (inclusive), take the nearest node (first strict minimum of the squared distance),
//   then the first of its elements (ascending id) that passes InsideOfElement's six plane
//   tests (:8746-8866).
// setEbOM (FEMSolverClass.f90:1891-1985) adds penalty * ConnectMatrix for each connect in
// order (FEMDomainClass.f90:13014-13100): Bc = [N_own(gp) I, -N_pair(xi) I] (3 x 48) and
// A = (Bc^T Bc) detJ, xi the local coordinate of the Gauss point in the paired element
// (getLocalCoordinate: inverse of the Gauss-point-averaged J applied to x - node centroid).
namespace {

__device__ inline double hex8_N_xi(int a, double x, double y, double z) {
  const double t0 = SX[a] > 0 ? 1.0 + x : 1.0 - x;
  const double t1 = SY[a] > 0 ? 1.0 + y : 1.0 - y;
  const double t2 = SZ[a] > 0 ? 1.0 + z : 1.0 - z;
  return ((0.125 * t0) * t1) * t2;
}

// InsideOfElementMesh, hex8: outside when (p1 x p2) . (x - X0) > 0 for any face triple
__device__ inline bool inside_hex8(const double (&X)[8][3], const double *p) {
  constexpr int T[6][3] = {{0, 3, 1}, {0, 1, 4}, {0, 4, 3}, {2, 6, 1}, {6, 7, 5}, {2, 3, 6}};
#pragma unroll
  for (int t = 0; t < 6; ++t) {
    const double *x0 = X[T[t][0]], *x1 = X[T[t][1]], *x2 = X[T[t][2]];
    const double a0 = x1[0] - x0[0], a1 = x1[1] - x0[1], a2 = x1[2] - x0[2];
    const double b0 = x2[0] - x0[0], b1 = x2[1] - x0[1], b2 = x2[2] - x0[2];
    const double o0 = p[0] - x0[0], o1 = p[1] - x0[1], o2 = p[2] - x0[2];
    const double n0 = a1 * b2 - a2 * b1, n1 = a2 * b0 - a0 * b2, n2 = a0 * b1 - a1 * b0;
    if ((n0 * o0 + n1 * o1) + n2 * o2 > 0.0) return false;
  }
  return true;
}

// node bounding box of one mesh: bb = xmin, ymin, zmin, xmax, ymax, zmax (one block)
__global__ void __launch_bounds__(256) k_node_bbox(int64_t nn, const double *__restrict__ xyz, double *bb) {
  __shared__ double smin[3][256], smax[3][256];
  double lo[3], hi[3];
  for (int d = 0; d < 3; ++d) lo[d] = hi[d] = xyz[(int64_t)d * nn];
  for (int64_t i = threadIdx.x; i < nn; i += 256)
    for (int d = 0; d < 3; ++d) {
      const double v = xyz[(int64_t)d * nn + i];
      lo[d] = fmin(lo[d], v);
      hi[d] = fmax(hi[d], v);
    }
  for (int d = 0; d < 3; ++d) {
    smin[d][threadIdx.x] = lo[d];
    smax[d][threadIdx.x] = hi[d];
  }
  __syncthreads();
  for (int s = 128; s > 0; s >>= 1) {
    if ((int)threadIdx.x < s)
      for (int d = 0; d < 3; ++d) {
        smin[d][threadIdx.x] = fmin(smin[d][threadIdx.x], smin[d][threadIdx.x + s]);
        smax[d][threadIdx.x] = fmax(smax[d][threadIdx.x], smax[d][threadIdx.x + s]);
      }
    __syncthreads();
  }
  if (threadIdx.x < 3) {
    bb[threadIdx.x] = smin[threadIdx.x][0];
    bb[3 + threadIdx.x] = smax[threadIdx.x][0];
  }
}

// GPP (P2P = false): one wavefront per (own element, Gauss point) w = 8e + gp; P2P: one per own
// node w (the point is the node itself, FEMDomainClass.f90:14701-14760, no element filter).
// pair[w] = containing element of the other mesh (0-based) or -1, pos[3w..] = the point.
template <bool P2P>
__global__ void __launch_bounds__(256) k_overset_gpp(int64_t nnA, int64_t neA, const double *__restrict__ xyzA,
                                                     const int32_t *__restrict__ enA, int64_t nnB, int64_t neB,
                                                     const double *__restrict__ xyzB,
                                                     const int32_t *__restrict__ enB,
                                                     const int64_t *__restrict__ eoffB,
                                                     const int32_t *__restrict__ elistB,
                                                     const double *__restrict__ bb, int32_t *__restrict__ pair,
                                                     double *__restrict__ pos) {
  const int lane = threadIdx.x & 63;
  const int64_t w = (blockIdx.x * (int64_t)blockDim.x + threadIdx.x) >> 6;
  if (w >= (P2P ? nnA : neA * 8)) return;  // uniform per wavefront
  bool hit = true;
  double p[3];
  if (P2P) {
    p[0] = xyzA[w];
    p[1] = xyzA[nnA + w];
    p[2] = xyzA[2 * nnA + w];
  } else {
    const int64_t e = w >> 3;
    const int gp = (int)(w & 7);
    int32_t nodes[8];
    double X[8][3];
    load_element(nnA, neA, xyzA, enA, e, nodes, X);
    // getElementList(xmin..zmax of the other mesh): element node range meets the box
#pragma unroll
    for (int d = 0; d < 3; ++d) {
      double lo = X[0][d], hi = X[0][d];
#pragma unroll
      for (int a = 1; a < 8; ++a) {
        lo = fmin(lo, X[a][d]);
        hi = fmax(hi, X[a][d]);
      }
      hit = hit && (hi >= bb[d]) && (lo <= bb[3 + d]);
    }
#pragma unroll
    for (int d = 0; d < 3; ++d) {
      double s = 0.0;
#pragma unroll
      for (int a = 0; a < 8; ++a) s = s + X[a][d] * hex8_N(a, gp);
      p[d] = s;
    }
  }
  hit = hit && (bb[0] <= p[0] && p[0] <= bb[3]) && (bb[1] <= p[1] && p[1] <= bb[4]) &&
        (bb[2] <= p[2] && p[2] <= bb[5]);
  if (lane == 0) {
    pos[3 * w] = p[0];
    pos[3 * w + 1] = p[1];
    pos[3 * w + 2] = p[2];
  }
  if (!hit) {
    if (lane == 0) pair[w] = -1;
    return;
  }
  // nearest node: each lane keeps its first strict minimum, ties go to the lower id
  double best = INFINITY;
  int32_t bi = INT32_MAX;
  for (int64_t i = lane; i < nnB; i += 64) {
    const double dx = p[0] - xyzB[i], dy = p[1] - xyzB[nnB + i], dz = p[2] - xyzB[2 * nnB + i];
    const double t = (dx * dx + dy * dy) + dz * dz;
    if (t < best) {
      best = t;
      bi = (int32_t)i;
    }
  }
#pragma unroll
  for (int off = 32; off > 0; off >>= 1) {
    const double ob = __shfl_xor(best, off, 64);
    const int32_t oi = __shfl_xor(bi, off, 64);
    if (ob < best || (ob == best && oi < bi)) {
      best = ob;
      bi = oi;
    }
  }
  int32_t found = -1;
  if (bi != INT32_MAX) {
    const int64_t s0 = eoffB[bi], s1 = eoffB[bi + 1];
    for (int64_t base = s0; base < s1 && found < 0; base += 64) {
      bool in = false;
      int32_t cand = -1;
      if (base + lane < s1) {
        cand = elistB[base + lane];
        int32_t nb[8];
        double Y[8][3];
        load_element(nnB, neB, xyzB, enB, cand, nb, Y);
        in = inside_hex8(Y, p);
      }
      const unsigned long long m = __ballot(in);
      if (m) found = __shfl(cand, __ffsll((long long)m) - 1, 64);
    }
  }
  if (lane == 0) pair[w] = found;
}

__global__ void k_overset_flags(int64_t m, const int32_t *__restrict__ pair, int32_t *__restrict__ flag) {
  const int64_t w = blockIdx.x * (int64_t)blockDim.x + threadIdx.x;
  if (w >= m) return;
  flag[w] = pair[w] >= 0 ? 1 : 0;
}

// connects in (own element, Gauss point) order, ids 1-based; P2P: (own node, 0)
__global__ void k_overset_scatter(int64_t m, bool p2p, const int32_t *__restrict__ pair,
                                  const double *__restrict__ pos, const int64_t *__restrict__ off,
                                  int32_t *__restrict__ elem, int32_t *__restrict__ gpo, int32_t *__restrict__ pe,
                                  double *__restrict__ po) {
  const int64_t w = blockIdx.x * (int64_t)blockDim.x + threadIdx.x;
  if (w >= m || pair[w] < 0) return;
  const int64_t k = off[w];
  elem[k] = p2p ? (int32_t)w + 1 : (int32_t)(w >> 3) + 1;
  gpo[k] = p2p ? 0 : (int32_t)(w & 7) + 1;
  pe[k] = pair[w] + 1;
  po[3 * k] = pos[3 * w];
  po[3 * k + 1] = pos[3 * w + 1];
  po[3 * k + 2] = pos[3 * w + 2];
}

// per connect: c16 = [0 + N_own(gp), 0 - N_pair(xi)], detJ of the own Gauss point, and the
// 16 node ids (0-based) of the own and the paired element
__global__ void __launch_bounds__(128) k_connect_coeffs(int64_t nn, int64_t ne, const double *__restrict__ xyz,
                                                        const int32_t *__restrict__ en, int64_t ncon,
                                                        const int32_t *__restrict__ my,
                                                        const int32_t *__restrict__ gpid,
                                                        const int32_t *__restrict__ pr,
                                                        const double *__restrict__ pos, double *__restrict__ c16,
                                                        double *__restrict__ detj, int32_t *__restrict__ nodes16) {
  const int64_t c = blockIdx.x * (int64_t)blockDim.x + threadIdx.x;
  if (c >= ncon) return;
  const bool p2p = gpid[c] == 0;  // P2P: my[c] is the own node (0-based)
  int32_t na[8], nb[8];
  double X[8][3], Y[8][3];
  if (!p2p) load_element(nn, ne, xyz, en, my[c], na, X);
  load_element(nn, ne, xyz, en, pr[c], nb, Y);
  const int gp = p2p ? 0 : gpid[c] - 1;
  double G[3][8];
  detj[c] = p2p ? 1.0 : hex8_gp_G(X, gp, 0, G, false);
  // local coordinate in the paired element
  double Jb[3][3] = {{0, 0, 0}, {0, 0, 0}, {0, 0, 0}};
#pragma unroll 1
  for (int g = 0; g < 8; ++g) {
    const double gx = (g == 1 || g == 2 || g == 5 || g == 6) ? GPV : -GPV;
    const double gy = (g == 2 || g == 3 || g == 6 || g == 7) ? GPV : -GPV;
    const double gz = (g >= 4) ? GPV : -GPV;
    double dN[3][8];
#pragma unroll
    for (int b = 0; b < 8; ++b) dN_node(b, gx, gy, gz, dN[0][b], dN[1][b], dN[2][b]);
#pragma unroll
    for (int i = 0; i < 3; ++i)
#pragma unroll
      for (int j = 0; j < 3; ++j) {
        double s = 0.0;
#pragma unroll
        for (int b = 0; b < 8; ++b) s = s + dN[i][b] * Y[b][j];
        Jb[i][j] = Jb[i][j] + s;
      }
  }
#pragma unroll
  for (int i = 0; i < 3; ++i)
#pragma unroll
    for (int j = 0; j < 3; ++j) Jb[i][j] = 0.125 * Jb[i][j];
  double x[3];
#pragma unroll
  for (int d = 0; d < 3; ++d) {
    double s = 0.0;
#pragma unroll
    for (int b = 0; b < 8; ++b) s = s + Y[b][d];
    x[d] = pos[3 * c + d] - 0.125 * s;
  }
  double Ji[3][3];
  inv3(Jb, Ji);
  double xi[3];
#pragma unroll
  for (int i = 0; i < 3; ++i) xi[i] = (Ji[i][0] * x[0] + Ji[i][1] * x[1]) + Ji[i][2] * x[2];
  if (p2p) {  // Bc = [I, -N_pair(xi) I] over the own node and the 8 paired nodes, no det J
    c16[c * 16] = 1.0;
    nodes16[c * 16] = my[c];
#pragma unroll
    for (int a = 0; a < 8; ++a) {
      c16[c * 16 + 1 + a] = 0.0 - hex8_N_xi(a, xi[0], xi[1], xi[2]);
      nodes16[c * 16 + 1 + a] = nb[a] - 1;
    }
#pragma unroll
    for (int a = 9; a < 16; ++a) {
      c16[c * 16 + a] = 0.0;
      nodes16[c * 16 + a] = -1;  // unused slot
    }
    return;
  }
#pragma unroll
  for (int a = 0; a < 8; ++a) {
    c16[c * 16 + a] = 0.0 + hex8_N(a, gp);
    c16[c * 16 + 8 + a] = 0.0 - hex8_N_xi(a, xi[0], xi[1], xi[2]);
    nodes16[c * 16 + a] = na[a] - 1;
    nodes16[c * 16 + 8 + a] = nb[a] - 1;
  }
}

// node -> (connect, slot) lists: count / fill / sort (t = 16c + slot ascends with c)
__global__ void k_con_count(int64_t total, const int32_t *__restrict__ nodes16, int32_t *__restrict__ cnt) {
  const int64_t t = blockIdx.x * (int64_t)blockDim.x + threadIdx.x;
  if (t >= total || nodes16[t] < 0) return;
  atomicAdd(&cnt[nodes16[t]], 1);
}

// (node << 32 | 16c + slot) keys: sorted, they give every node its (connect, slot) list in
// connect order
__global__ void k_con_keys(int64_t total, const int32_t *__restrict__ nodes16, uint64_t *__restrict__ keys) {
  const int64_t t = blockIdx.x * (int64_t)blockDim.x + threadIdx.x;
  if (t >= total) return;
  // unused slots sort behind every node's list (the node counts exclude them)
  keys[t] = nodes16[t] < 0 ? ~(uint64_t)0 : ((uint64_t)(uint32_t)nodes16[t] << 32) | (uint64_t)(uint32_t)t;
}

__global__ void k_con_list(int64_t total, const uint64_t *__restrict__ keys, int64_t *__restrict__ list) {
  const int64_t t = blockIdx.x * (int64_t)blockDim.x + threadIdx.x;
  if (t >= total) return;
  list[t] = (int64_t)(keys[t] & 0xffffffffu);
}

__device__ inline int64_t find_col(const int32_t *__restrict__ col_idx, int64_t rb, int64_t re, int32_t col,
                                   bool sorted) {
  if (sorted) {
    int64_t lo = rb, hi = re;
    while (lo < hi) {
      const int64_t mid = (lo + hi) >> 1;
      if (col_idx[mid] < col) lo = mid + 1;
      else hi = mid;
    }
    return (lo < re && col_idx[lo] == col) ? lo : -1;
  }
  for (int64_t k = rb; k < re; ++k)
    if (col_idx[k] == col) return k;
  return -1;
}

// rows longer than this are updated by one lane straight in global memory
constexpr int EB_MAXROW = 2048;

// one wavefront per row r = 3i + m: the connects touching node i in connect order; within a
// connect lane 3k + n adds entry (3j + m, 3k + n) of penalty * A (addMatrixValue,
// FEMSolverClass.f90:923-960) to the row held in LDS.  The 48 columns of one connect are
// distinct, and the barrier between connects keeps every entry's sum in connect order.
__global__ void __launch_bounds__(64) k_ebom_rows(int64_t nn, const int64_t *__restrict__ off,
                                                  const int64_t *__restrict__ list,
                                                  const int32_t *__restrict__ nodes16,
                                                  const double *__restrict__ c16, const double *__restrict__ detj,
                                                  double penalty, const int64_t *__restrict__ row_ptr,
                                                  const int32_t *__restrict__ col_idx, bool sorted,
                                                  double *__restrict__ val, int32_t *__restrict__ err) {
  __shared__ double sv[EB_MAXROW];
  const int lane = threadIdx.x;
  const int64_t r = blockIdx.x;
  if (r >= 3 * nn) return;
  const int64_t i = r / 3;
  const int m = (int)(r - 3 * i);
  const int64_t s0 = off[i], s1 = off[i + 1];
  if (s0 == s1) return;
  const int64_t rb = row_ptr[r] - 1, re = row_ptr[r + 1] - 1;
  const int64_t len = re - rb;
  const bool in_lds = len <= EB_MAXROW;
  if (!in_lds && lane != 0) return;
  if (in_lds) {
    for (int64_t q = lane; q < len; q += 64) sv[q] = val[rb + q];
    __syncthreads();
  }
  for (int64_t s = s0; s < s1; ++s) {
    const int64_t c = list[s] >> 4;
    const int j = (int)(list[s] & 15);
    const double cj = c16[c * 16 + j], dj = detj[c];
    for (int kn = in_lds ? lane : 0; kn < 48; kn += in_lds ? 64 : 1) {
      const int k = kn / 3, n = kn - 3 * (kn / 3);
      if (nodes16[c * 16 + k] < 0) continue;  // P2P connects use 9 of the 16 slots
      const double ck = c16[c * 16 + k];
      // (Bc^T Bc)(3j+m, 3k+n) = sum_l Bc(l,3j+m) Bc(l,3k+n), Bc(l, 3q+d) = c_q [l == d]
      double s3 = 0.0;
#pragma unroll
      for (int l = 0; l < 3; ++l) s3 = s3 + (l == m ? cj : 0.0) * (l == n ? ck : 0.0);
      const double v = penalty * (s3 * dj);
      const int64_t id = find_col(col_idx, rb, re, 3 * nodes16[c * 16 + k] + n + 1, sorted);
      if (id < 0) {
        atomicOr(err, 1);
        continue;
      }
      if (in_lds) sv[id - rb] = sv[id - rb] + v;
      else val[id] = val[id] + v;
    }
    if (in_lds) __syncthreads();
  }
  if (in_lds)
    for (int64_t q = lane; q < len; q += 64) val[rb + q] = sv[q];
}

// hex8 stiffness into any CRS holding the mesh's node blocks: one thread per row, incident
// elements in ascending id, every column located by search
__global__ void __launch_bounds__(128) k_assemble_rows_search(
    int64_t nn, int64_t ne, const double *__restrict__ xyz, const int32_t *__restrict__ en,
    const int64_t *__restrict__ eoff, const int32_t *__restrict__ elist, const double *__restrict__ E,
    const double *__restrict__ nu, int jmode, const int64_t *__restrict__ row_ptr,
    const int32_t *__restrict__ col_idx, bool sorted, double *__restrict__ val, int32_t *__restrict__ err) {
  const int64_t r = blockIdx.x * (int64_t)blockDim.x + threadIdx.x;
  if (r >= 3 * nn) return;
  const int64_t i = r / 3;
  const int d = (int)(r - 3 * i);
  const int64_t rb = row_ptr[r] - 1, re = row_ptr[r + 1] - 1;
  for (int64_t s = eoff[i]; s < eoff[i + 1]; ++s) {
    const int64_t e = elist[s];
    int32_t nodes[8];
    double X[8][3];
    load_element(nn, ne, xyz, en, e, nodes, X);
    int a = 0;
    while (a < 8 && nodes[a] - 1 != i) ++a;
    double acc[24];
    hex8_ke_row(X, E[e], nu[e], jmode, a, d, acc);
    for (int b = 0; b < 8; ++b)
      for (int dd = 0; dd < 3; ++dd) {
        const int64_t id = find_col(col_idx, rb, re, 3 * (nodes[b] - 1) + dd + 1, sorted);
        if (id < 0) {
          atomicOr(err, 1);
          continue;
        }
        val[id] = val[id] + acc[3 * b + dd];
      }
  }
}

}  // namespace

int pf_overset_search(pf_ctx *ctx, const NodeGraph &A, const NodeGraph &B, bool p2p, DevBuf<int32_t> &elem,
                      DevBuf<int32_t> &gp, DevBuf<int32_t> &pair_elem, DevBuf<double> &position, int64_t *count) {
  *count = 0;
  if (A.nne != 8 || B.nne != 8 || !A.has_coords || !B.has_coords || !B.built) {
    pf_set_error("overset: two hex8 meshes with coordinates (and the node->element lists of the other)");
    return PF_INVALID;
  }
  if ((p2p ? A.nn : A.ne) == 0 || B.nn == 0 || B.ne == 0) return PF_OK;
  const int64_t m = p2p ? A.nn : A.ne * 8;
  DevBuf<double> bb, pos;
  DevBuf<int32_t> pair, flag;
  DevBuf<int64_t> off, tmp;
  PF_CHECK(bb.ensure(6));
  PF_CHECK(pos.ensure((size_t)m * 3));
  PF_CHECK(pair.ensure((size_t)m));
  PF_CHECK(flag.ensure((size_t)m + 1));
  PF_CHECK(off.ensure((size_t)m + 1));
  k_node_bbox<<<1, 256, 0, ctx->stream>>>(B.nn, B.nodcoord.p, bb.p);
  if (p2p)
    k_overset_gpp<true><<<pf_blocks(m * 64, 256, INT32_MAX), 256, 0, ctx->stream>>>(
        A.nn, A.ne, A.nodcoord.p, A.elemnod.p, B.nn, B.ne, B.nodcoord.p, B.elemnod.p, B.eoff.p, B.elist.p, bb.p,
        pair.p, pos.p);
  else
    k_overset_gpp<false><<<pf_blocks(m * 64, 256, INT32_MAX), 256, 0, ctx->stream>>>(
        A.nn, A.ne, A.nodcoord.p, A.elemnod.p, B.nn, B.ne, B.nodcoord.p, B.elemnod.p, B.eoff.p, B.elist.p, bb.p,
        pair.p, pos.p);
  k_overset_flags<<<pf_blocks(m, 256, INT32_MAX), 256, 0, ctx->stream>>>(m, pair.p, flag.p);
  PF_HIP(hipGetLastError());
  PF_CHECK(pf_scan_i32_to_i64(ctx, flag.p, m, off.p, tmp));
  int64_t k = 0;
  PF_CHECK(pf_read_i64(ctx, off.p + m, &k));
  PF_CHECK(elem.ensure((size_t)k + 1));
  PF_CHECK(gp.ensure((size_t)k + 1));
  PF_CHECK(pair_elem.ensure((size_t)k + 1));
  PF_CHECK(position.ensure((size_t)k * 3 + 1));
  if (k > 0)
    k_overset_scatter<<<pf_blocks(m, 256, INT32_MAX), 256, 0, ctx->stream>>>(m, p2p, pair.p, pos.p, off.p, elem.p,
                                                                            gp.p, pair_elem.p, position.p);
  PF_HIP(hipGetLastError());
  PF_HIP(hipStreamSynchronize(ctx->stream));
  *count = k;
  return PF_OK;
}

int pf_ebom_assemble(pf_ctx *ctx, const NodeGraph &g, int64_t ncon, const int32_t *my, const int32_t *gp,
                     const int32_t *pr, const double *pos, double penalty, const int64_t *row_ptr,
                     const int32_t *col_idx, bool sorted, double *val, int32_t *err) {
  if (g.nne != 8 || !g.has_coords) {
    pf_set_error("setEbOM needs the hex8 mesh with coordinates");
    return PF_INVALID;
  }
  if (ncon <= 0) return PF_OK;
  if (ncon >= ((int64_t)1 << 28)) {
    pf_set_error("setEbOM: at most 2^28 connects per call");
    return PF_INVALID;
  }
  DevBuf<double> c16, detj;
  DevBuf<int32_t> nodes16, cnt;
  DevBuf<int64_t> off, list, tmp;
  const int64_t total = ncon * 16;
  PF_CHECK(c16.ensure((size_t)total));
  PF_CHECK(detj.ensure((size_t)ncon));
  PF_CHECK(nodes16.ensure((size_t)total));
  PF_CHECK(cnt.ensure((size_t)g.nn + 1));
  PF_CHECK(off.ensure((size_t)g.nn + 1));
  PF_CHECK(list.ensure((size_t)total));
  k_connect_coeffs<<<pf_blocks(ncon, 128, INT32_MAX), 128, 0, ctx->stream>>>(
      g.nn, g.ne, g.nodcoord.p, g.elemnod.p, ncon, my, gp, pr, pos, c16.p, detj.p, nodes16.p);
  PF_HIP(hipMemsetAsync(cnt.p, 0, sizeof(int32_t) * (g.nn + 1), ctx->stream));
  k_con_count<<<pf_blocks(total, 256, INT32_MAX), 256, 0, ctx->stream>>>(total, nodes16.p, cnt.p);
  PF_CHECK(pf_scan_i32_to_i64(ctx, cnt.p, g.nn, off.p, tmp));
  DevBuf<uint64_t> ka, kb;
  PF_CHECK(ka.ensure((size_t)total));
  PF_CHECK(kb.ensure((size_t)total));
  k_con_keys<<<pf_blocks(total, 256, INT32_MAX), 256, 0, ctx->stream>>>(total, nodes16.p, ka.p);
  PF_HIP(hipGetLastError());
  const int end_bit = 64;  // the unused-slot keys set every bit
  bool in_b = false;
  PF_CHECK(pf_sort_u64(ctx, ka.p, kb.p, total, end_bit, &in_b));
  k_con_list<<<pf_blocks(total, 256, INT32_MAX), 256, 0, ctx->stream>>>(total, in_b ? kb.p : ka.p, list.p);
  k_ebom_rows<<<(unsigned)(3 * g.nn), 64, 0, ctx->stream>>>(g.nn, off.p, list.p, nodes16.p, c16.p, detj.p, penalty,
                                                          row_ptr, col_idx, sorted, val, err);
  PF_HIP(hipGetLastError());
  PF_HIP(hipStreamSynchronize(ctx->stream));
  return PF_OK;
}

int pf_assemble_hex8_search(pf_ctx *ctx, const NodeGraph &g, const double *E, const double *nu, int jmode,
                            const int64_t *row_ptr, const int32_t *col_idx, bool sorted, double *val,
                            int32_t *err) {
  if (g.nn == 0) return PF_OK;
  k_assemble_rows_search<<<pf_blocks(3 * g.nn, 128, INT32_MAX), 128, 0, ctx->stream>>>(
      g.nn, g.ne, g.nodcoord.p, g.elemnod.p, g.eoff.p, g.elist.p, E, nu, jmode, row_ptr, col_idx, sorted, val,
      err);
  PF_HIP(hipGetLastError());
  return PF_OK;
}
