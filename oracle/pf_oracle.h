/*
 * pf_oracle.h -- CPU restatement of plantFEM's static linear-elastic solve path.
 *
 * TEST INFRASTRUCTURE ONLY.  This library is the checker for the HIP path
 * (plantfem_amd/csrc) and the timed CPU baseline in bench.py.  Only tests/,
 * __graft_entry__.smoke() and bench.py's cpu_baseline leg may load it.  The
 * product library never links, loads or calls it.
 *
 * Conventions follow the reference's Fortran storage:
 *   - node coordinates NodCoord(nn,3) column-major  -> nodcoord[d*nn + i]
 *   - connectivity   ElemNod(ne,nne) column-major    -> elemnod[a*ne + e], 1-based node ids
 *   - CRS: row_ptr int64 (n+1, 1-based), col_idx int32 (nnz, 1-based), val double
 *   - element matrices Ke(24,24) column-major        -> ke[c*24 + r]
 *
 * Each function cites the reference file:line it restates (paths relative to
 * the plantFEM source tree).
 */
#ifndef PF_ORACLE_H
#define PF_ORACLE_H
#include <stdint.h>

#ifdef __cplusplus
extern "C" {
#endif

/* status codes shared with the product library */
#define ORC_OK 0
#define ORC_BREAKDOWN_VV 1      /* "Bicgstab devide by zero" at v.v == 0  */
#define ORC_BREAKDOWN_C2C3 2    /* "Bicgstab devide by zero" at c2*c3 == 0 */
#define ORC_NOT_CONVERGED 3     /* loop ran to itrmax (silent in the reference) */
#define ORC_INVALID 4

/* A1: Cube3D structured box (MeshClass.f90:6348-6361, :6745-6808, :6992-7090) */
void orc_cube3d(int xn, int yn, int zn, double lx, double ly, double lz,
                double *nodcoord, int32_t *elemnod);

/* A3: ZeroMatrix_as_CRS_FEMDomain, non-legacy branch (FEMDomainClass.f90:10665-10838).
 * Returns nnz; row_ptr/col_idx may be NULL (count only).  Returns -1 on invalid input. */
int64_t orc_zero_matrix(int64_t nn, int64_t ne, int nne, const int32_t *elemnod, int dof,
                        int64_t *row_ptr, int32_t *col_idx);

/* A3 legacy branch: COO append + to_CRS + DOF extension (FEMDomainClass.f90:10597-10655,
 * SparseClass.f90:358-409, :450-500).  Independent second route to the same arrays. */
int64_t orc_zero_matrix_legacy(int64_t nn, int64_t ne, int nne, const int32_t *elemnod, int dof,
                               int64_t *row_ptr, int32_t *col_idx);

/* A4-A8: StiffnessMatrixFEMDomain for one hex8 element (FEMDomainClass.f90:11457-11515).
 * X is the element coordinate table ElemCoord(8,3) row-major x[a*3+d].
 * jmode 0 = reference B (J^-T dN/dxi, FEMDomainClass.f90:12289), 1 = textbook J^-1 dN/dxi.
 * ke is 24x24 column-major (Fortran layout). */
void orc_hex8_stiffness(const double *X, double E, double nu, int jmode, double *ke);

/* batched form over a mesh: ke[(e*24*24) + c*24 + r] */
void orc_hex8_stiffness_all(int64_t nn, int64_t ne, const double *nodcoord, const int32_t *elemnod,
                            const double *E, const double *nu, int jmode, double *ke);

/* A9: StiffnessMatrix_as_CRS_FEMDomain serial order (FEMDomainClass.f90:11193-11229):
 * val += Ke in ascending element order with a linear search of the row.  val must hold
 * the ZeroMatrix pattern's nnz entries (it is zeroed here).  Returns 0 or ORC_INVALID. */
int orc_assemble_stiffness(int64_t nn, int64_t ne, const double *nodcoord, const int32_t *elemnod,
                           const double *E, const double *nu, int jmode,
                           const int64_t *row_ptr, const int32_t *col_idx, double *val);

/* A10: setValueFEMSolver matrix branch for one element (FEMSolverClass.f90:1020-1057). */
int orc_set_element_matrix(int64_t ne, int nne, const int32_t *elemnod, int64_t elem /*1-based*/,
                           int dof, const double *ke /* (nne*dof)^2 col-major */,
                           const int64_t *row_ptr, const int32_t *col_idx, double *val);

/* A11: TractionVector_by_elemFEMDomain summed over all elements (FEMDomainClass.f90:17210-17292,
 * facets :8192-8204).  Facets whose 4 nodes lie in the inclusive box get p*area*n/4 per node. */
void orc_traction_box(int64_t nn, int64_t ne, const double *nodcoord, const int32_t *elemnod,
                      const double *box /* xmin,xmax,ymin,ymax,zmin,zmax */,
                      double pressure, double *rhs /* 3*nn, accumulated */);

/* A12: selectFEMDomain box form (FEMDomainClass.f90:12854-12926); returns count, fills ids (1-based)
 * when ids != NULL. */
int64_t orc_select_box(int64_t nn, const double *nodcoord, const double *box, int32_t *ids);

/* MassVectorFEMDomain(ElementID, Density, DOF=3, Accel) (FEMDomainClass.f90:10993-11073) of one
 * hex8: mv[3a+k] = rho * sum_gp N_a(gp) accel_k detJ(gp) (unit Gauss weights). */
void orc_hex8_mass_vector(const double *X, double rho, const double *accel, double *mv);
/* sum of the element mass vectors in ascending element order added to rhs (the setVector loop
 * of Tutorial/fem/benchmark_bar_in_gravity_field.f90:20-31); rho per element. */
void orc_body_force(int64_t nn, int64_t ne, const double *nodcoord, const int32_t *elemnod,
                    const double *rho, const double *accel, double *rhs);
/* StressMatrixFEMDomain without GaussPoint (FEMDomainClass.f90:11912-12032): cell average over
 * the 8 Gauss points of D B u_e; s[3*j + i] = sigma(i+1, j+1).  disp is the solver's x
 * (node-major, ArrayClass reshape). */
void orc_hex8_cell_stress(const double *X, const double *ue, double E, double nu, int jmode, double *s);
void orc_cell_stress_all(int64_t nn, int64_t ne, const double *nodcoord, const int32_t *elemnod,
                         const double *disp, const double *E, const double *nu, int jmode, double *sigma);

/* A14 steps 1-4: Dirichlet elimination of solveFEMSolver (FEMSolverClass.f90:1707-1756). */
void orc_apply_dirichlet(int64_t n, const int64_t *row_ptr, const int32_t *col_idx, double *val,
                         double *rhs, const uint8_t *fixed, const double *ubar);

/* A16: sub_crs_matvec_generic (LinearSolverClass.f90:3728-3776): y = A x with an OpenMP
 * row loop and an atomic add per nonzero, as in the reference.  nthreads<=0: OpenMP default. */
void orc_spmv(int64_t n, const int64_t *row_ptr, const int32_t *col_idx, const double *val,
              const double *x, double *y, int nthreads);

/* A15: bicgstab_CRS_2 (FEMSolverClass.f90:2199-2298).  x holds x0 on entry.  Returns a status
 * code; *iters = iterations run, *rr_out = final r.r, *rr1_out = r.r after iteration 1.
 * hist (may be NULL, length >= itrmax) receives sqrt(rr) per iteration. */
int orc_bicgstab_crs2(int64_t n, const int64_t *row_ptr, const int32_t *col_idx, const double *val,
                      const double *b, double *x, int itrmax, double er0, double relative_er,
                      int nthreads, int *iters, double *rr_out, double *rr1_out, double *hist);

/* A17: bicgstab_CRS (LinearSolverClass.f90:1493-1569), the LinearSolver_ form: the same
 * recurrence as bicgstab_CRS_2, but er0 = er, init_rr = r.r of the initial residual, a return
 * before the loop when init_rr < er (:1527) and the exit rr/init_rr < er with no square root
 * (:1561).  *iters = 0 for the early return.  hist as above. */
int orc_bicgstab_crs(int64_t n, const int64_t *row_ptr, const int32_t *col_idx, const double *val,
                     const double *b, double *x, int itrmax, double er, int nthreads, int *iters,
                     double *rr_out, double *init_rr_out, double *hist);

/* A19: crs_spmv_real64 / crs_spmv_real32 (SparseClass-c.c:38-66): new_vector(row) =
 * new_vector(row) + val*old_vector(col), entries of each row in stored order, serial. */
void orc_crs_spmv_real64(const double *val, const int64_t *row_ptr, const int32_t *col_idx,
                         const double *old_vector, double *new_vector, int64_t n);
void orc_crs_spmv_real32(const float *val, const int64_t *row_ptr, const int32_t *col_idx,
                         const float *old_vector, float *new_vector, int64_t n);

/* orc_assemble_stiffness restricted to the DOF rows of the nodes with row_node[node] != 0:
 * the same ascending element loop and row searches, so each computed entry equals the full
 * assembly's bitwise; val (nnz) is zeroed here and entries of other rows stay zero.  For
 * spot checks of large meshes (only elements touching a marked node are evaluated). */
int orc_assemble_stiffness_rows(int64_t nn, int64_t ne, const double *nodcoord, const int32_t *elemnod,
                                const double *E, const double *nu, int jmode, const uint8_t *row_node,
                                const int64_t *row_ptr, const int32_t *col_idx, double *val);

/* Whole solveFEMSolver (FEMSolverClass.f90:1697-1825) on a copy: elimination, BiCGSTAB, x_fixed = ubar. */
int orc_solve(int64_t n, const int64_t *row_ptr, const int32_t *col_idx, double *val, double *rhs,
              const uint8_t *fixed, const double *ubar, double *x, int itrmax, double er0,
              double relative_er, int nthreads, int *iters, double *rr_out, double *rr1_out);

/* JacobiPreconditionerCRS (FEMSolverClass.f90:2508-2529): rows scaled by 1/A_ii in place. */
void orc_jacobi_scale(int64_t n, const int64_t *row_ptr, const int32_t *col_idx, double *val, double *rhs);

/* ILUCRS(0) (SparseClass.f90:1629-1700) in place, columns k < i taken in stored order. */
void orc_ilu0(int64_t n, const int64_t *row_ptr, const int32_t *col_idx, double *val);

/* ILU_matvecCRS (SparseClass.f90:1795-1845): y = D_U^-1 L^-1 v (its backward loop is a no-op). */
void orc_ilu_apply(int64_t n, const int64_t *row_ptr, const int32_t *col_idx, const double *fval,
                   const double *v, double *y);

/* bicgstab_CRS_ILU (LinearSolverClass.f90:4104-4229) with the factor fval from orc_ilu0. */
int orc_bicgstab_ilu(int64_t n, const int64_t *row_ptr, const int32_t *col_idx, const double *val,
                     const double *fval, const double *b, double *x, int itrmax, double er0,
                     double relative_er, int nthreads, int *iters, double *rr_out, double *rr1_out);

/* MassMatrixFEMDomain (FEMDomainClass.f90:10874-10945), DOF = 3: 24x24 column-major. */
void orc_hex8_mass_matrix(const double *X, double rho, double *me);

/* MassMatrix_as_CRS_FEMDomain (FEMDomainClass.f90:10364-10467) on the ZeroMatrix pattern,
 * ascending element order; rho per element. */
int orc_mass_matrix_crs(int64_t nn, int64_t ne, const double *nodcoord, const int32_t *elemnod,
                        const double *rho, const int64_t *row_ptr, const int32_t *col_idx, double *val);

/* multi-domain assembly with overset penalties (FEMDomainClass.f90:14606-14700, :13014-13100;
 * MeshClass.f90:8611-8866; FEMSolverClass.f90:466-676, :1891-1985) */
void orc_gauss_point_position(const double *X, int gp, double *pos);
int orc_inside_hex8(const double *X, const double *p);
int64_t orc_nearest_element(int64_t nn, int64_t ne, const double *xyz, const int32_t *en, const double *p);
int64_t orc_overset_gpp(int64_t nnA, int64_t neA, const double *xyzA, const int32_t *enA, int64_t nnB, int64_t neB,
                        const double *xyzB, const int32_t *enB, int32_t *my_elem, int32_t *gp_out,
                        int32_t *pair_elem, double *pos_out, int64_t cap);
int64_t orc_overset_p2p(int64_t nnA, const double *xyzA, int64_t nnB, int64_t neB, const double *xyzB,
                        const int32_t *enB, int32_t *node_out, int32_t *pair_elem, double *pos_out, int64_t cap);
void orc_connect_coeffs_p2p(const double *Xpair, const double *pos, double *c9);
void orc_local_coordinate(const double *X, const double *p, double *xi);
void orc_connect_coeffs(const double *Xmy, int gp, const double *Xpair, const double *pos, double *c16, double *detJ);
void orc_connect_coeffs_all(int64_t nn, int64_t ne, const double *xyz, const int32_t *en, int64_t ncon,
                            const int32_t *elem, const int32_t *gp, const int32_t *pair, const double *pos,
                            double *c16, double *detJ);
int64_t orc_multi_crs(int64_t nn, int64_t ne, const int32_t *en, int any_overset, int64_t ncon, const int32_t *nodes16,
                      int64_t *row_ptr, int32_t *col_idx);
int orc_ebom(int64_t n, const int64_t *row_ptr, const int32_t *col_idx, double *val, int64_t ncon,
             const int32_t *nodes16, const double *c16, const double *detJ, double penalty);

/* Direct solve behind solve(algorithm="GaussJordan") (FEMSolverClass.f90:1795-1798; the
 * reference's gauss_jordan_crs, FEMSolverClass.f90:2413-2460, stops "Under construction", so
 * the oracle restates the textbook method the GPU path uses: Cholesky of a symmetric positive
 * definite band matrix).  Band layout of LAPACK's lower form: ab[j*ldab + (i - j)] = A[i][j]
 * for j <= i <= j + kd, ldab >= kd + 1.
 * orc_pbtrf_lower: LAPACK dpbtf2 ('L'), L L^T in place; 0, or j+1 when the pivot of column j
 * (0-based) is not positive.  orc_pbtrs_lower: LAPACK dpbtrs ('L'), b := (L L^T)^-1 b.
 * orc_band_from_crs: lower band of P A P^T from a CRS (1-based), perm[q] = original row of
 * band row q (0-based); 1 when an entry falls outside the band. */
int orc_pbtrf_lower(int64_t n, int32_t kd, double *ab, int32_t ldab);
void orc_pbtrs_lower(int64_t n, int32_t kd, const double *ab, int32_t ldab, double *b);
int orc_band_from_crs(int64_t n, const int64_t *row_ptr, const int32_t *col_idx, const double *val,
                      const int32_t *perm, int32_t kd, double *ab, int32_t ldab);

#ifdef __cplusplus
}
#endif
#endif
