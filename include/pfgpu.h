/*
 * pfgpu.h -- C-ABI of the MI355X (gfx950) implementation of plantFEM's static
 * linear-elastic solve path: hex8 CRS graph build, element stiffness assembly,
 * Dirichlet elimination and BiCGSTAB over CRS.
 *
 * Plain pointers and sizes only.  Array conventions are the reference's Fortran ones:
 *   NodCoord(nn,3) real64 column-major      nodcoord[d*nn + i]
 *   ElemNod(ne,nne) int32 column-major      elemnod[a*ne + e], node ids 1-based
 *   CRS_: row_ptr int64 (n+1, 1-based), col_idx int32 (nnz, 1-based), val real64
 *         (SparseClass.f90:128-131, FEMSolverClass.f90:24-27)
 *   element matrices Ke(24,24) real64 column-major   ke[c*24 + r]
 * Functions named pf_dev_* take DEVICE pointers (e.g. torch.cuda tensors); all
 * others take HOST pointers unless stated.
 *
 * Every function returns a status (PF_OK = 0).  The message of the last failure on
 * the calling thread is available from pf_last_error().
 */
#ifndef PFGPU_H
#define PFGPU_H
#include <stddef.h>
#include <stdint.h>

#ifdef __cplusplus
extern "C" {
#endif

/* ---- status codes --------------------------------------------------------------- */
#define PF_OK 0
#define PF_BREAKDOWN_VV 1    /* v.v == 0: reference stops "Bicgstab devide by zero" (FEMSolverClass.f90:2266) */
#define PF_BREAKDOWN_C2C3 2  /* c2*c3 == 0: same stop (FEMSolverClass.f90:2295) */
#define PF_NOT_CONVERGED 3   /* itrmax reached; the reference ends silently (FEMSolverClass.f90:2241) */
#define PF_INVALID 4         /* bad ids/sizes, missing CRS slot ("fill-in request" :1040-1047), call order */
#define PF_NOT_SPD 5         /* direct solve: a non-positive pivot in the Cholesky factorization; the
                                reference's gauss_jordan_crs stops "A is singular" on a zero pivot
                                (FEMSolverClass.f90:2455) */
#define PF_ERR_HIP -1        /* HIP runtime error */
#define PF_ERR_NOMEM -2      /* device allocation failed */

typedef struct pf_ctx pf_ctx;       /* one device + one HIP stream */
typedef struct pf_system pf_system; /* device-resident FEMSolver_ state (mesh, CRS, RHS, BCs, x) */

/* Solver options; defaults from FEMSolverClass.f90:69-71.  er0 and relative_er default
 * to dble(1.0e-10): a REAL(4) literal widened to REAL(8), i.e. 1.00000001335e-10. */
typedef struct {
  int32_t itrmax;        /* 100000 */
  int32_t precision;     /* 64: plain FP64 BiCGSTAB (bicgstab_CRS_2); 32: FP32 BiCGSTAB with FP64 reliable
                            updates (FP64 solution accumulation and true-residual replacement) */
  double er0;            /* absolute stop sqrt(r.r) < er0 */
  double relative_er;    /* stop sqrt(r.r / r1.r1) < relative_er (r1 = residual after iteration 1) */
  double inner_tol;      /* precision 32: replace the FP32 recursive residual by the FP64 true residual
                            once its norm fell below inner_tol times its largest value since the last
                            replacement (default 0.05) */
  double target_relres;  /* precision 32: stop when ||b-Ax||/||b|| (FP64, at a replacement) < target
                            (default 1e-8) */
  int32_t preconditioning; /* solve(preconditioning=) (FEMSolverClass.f90:1768-1790), FP64 only:
                              0 none: bicgstab_CRS_2 (default);
                              1 "PointJacobi": JacobiPreconditionerCRS on the eliminated system (in
                                place, as the reference), then bicgstab_CRS_2 on the scaled system;
                              2 "ILU": bicgstab_CRS_ILU with the ILU(0) factor of the eliminated
                                matrix, computed at the first such solve and kept (ILU_MATRIX);
                              3 "PointJacobi" as the reference's branch runs it: the scaling and no
                                solver call, so x = x0 (zeros) with the fixed values */
  int32_t spmv_kind;     /* 0 auto (blocked when the CRS has DOF=3 node blocks and pads little, else CRS
                            rows, with x staged in LDS for the unpreconditioned FP64 solve from 2^18
                            rows on), 1 CRS row kernel, 2 blocked, 4 CRS rows with x staged in LDS per
                            row block (FP64, no ILU) */
  int32_t check_every;   /* iterations enqueued between host polls of the stop flag (default 32) */
  int32_t use_graph;     /* capture the polled chunk in a hipGraph (default 1) */
  int32_t debug;         /* print sqrt(r.r) per poll */
  int32_t stop_mode;     /* 0: bicgstab_CRS_2's tests (default, FEMSolverClass.f90:2275-2289);
                            1: sqrt(r.r) < relative_er * ||b||;
                            2: bicgstab_CRS's tests (LinearSolverClass.f90:1493-1569) with er = er0:
                               return before the loop when r0.r0 < er (:1527), stop when
                               r.r / r0.r0 < er, no square root (:1561); FP64, no ILU */
  int32_t kernel_timing; /* 1: bracket every kernel of the iteration loop with hipEvents (no hipGraph)
                            and report the average device time of each in pf_stats;
                            2: the same for y = A p only (two events per iteration) */
  int32_t jacobi;        /* precision 32, scaling of the FP32 inner system (D = diag(A)):
                            2 (default): rows, D^-1 A, and every replacement restarts the inner
                              BiCGSTAB from the FP64 residual (iterative refinement);
                            1: symmetric, D^-1/2 A D^-1/2, reliable updates without restarts;
                            0: none */
  int32_t inner_max;     /* precision 32: also replace the residual after this many iterations without
                            a replacement (default 300; 0 = never) */
  int32_t use_x0;        /* multi-RHS solve: X holds the initial guess on entry (default 0: x0 = 0) */
} pf_solve_opts;

typedef struct {
  int32_t status;           /* PF_* code of the solve */
  int32_t iterations;       /* BiCGSTAB iterations (FP64 or, for precision 32, FP32) */
  int32_t outer_iterations; /* FP64 residual replacements (mixed), 0 for FP64 */
  int32_t spmv_kind_used;   /* 1 CRS row kernel, 2 blocked, 4 CRS rows with x staged in LDS */
  double rr;                /* final r.r of the recurrence (FP64 form) */
  double rr1;               /* r.r after iteration 1 */
  double true_relres;       /* ||b - A x|| / ||b|| in FP64 after the solve (eliminated system) */
  double solve_ms;          /* device time of the solve, hipEvents on the context stream */
  double setup_ms;          /* device time of per-solve setup (blocked layout, FP32 copy) */
  /* kernel_timing = 1: average device ms of y = A p, v = A e, the x/r update and the p update
     (FP32 inner kernels for precision 32); kernel_timing = 2: y = A p only, the others 0;
     0 otherwise */
  double spmv_p_ms, spmv_e_ms, update_ms, pupdate_ms;
  int64_t timed_iterations;
} pf_stats;

void pf_solve_opts_default(pf_solve_opts *opts);
const char *pf_last_error(void);
int pf_version(void);

/* ---- context --------------------------------------------------------------------- */
int pf_ctx_create(int device, pf_ctx **out);
int pf_ctx_destroy(pf_ctx *ctx);
int pf_ctx_synchronize(pf_ctx *ctx);
/* The context's HIP stream (hipStream_t as void*), for callers that order their own work. */
void *pf_ctx_stream(pf_ctx *ctx);

/* ---- stateless host-pointer entry points ---------------------------------------- */

/* ZeroMatrix_as_CRS_FEMDomain(DOF) (FEMDomainClass.f90:10581-10840): CRS graph of the
 * mesh in the reference's first-appearance column order.  Two-phase: call with
 * row_ptr == col_idx == NULL to get *nnz, then with arrays of n*dof+1 and *nnz. */
int pf_crs_graph(pf_ctx *ctx, int64_t nn, int64_t ne, int32_t nne, const int32_t *elemnod,
                 int32_t dof, int64_t *nnz, int64_t *row_ptr, int32_t *col_idx);

/* StiffnessMatrixFEMDomain(ElementID,E,v) for elements first..first+count-1 (1-based)
 * (FEMDomainClass.f90:11457-11515).  E, nu indexed by element (length ne).
 * jmode 0 = reference B = J^-T dN/dxi (FEMDomainClass.f90:12289); 1 = textbook J^-1 dN/dxi.
 * ke receives count matrices of 24x24, column-major. */
int pf_hex8_stiffness(pf_ctx *ctx, int64_t nn, int64_t ne, const double *nodcoord,
                      const int32_t *elemnod, const double *E, const double *nu, int32_t jmode,
                      int64_t first, int64_t count, double *ke);

/* MassMatrixFEMDomain(ElementID, Density, DOF=3) (FEMDomainClass.f90:10874-10945) for elements
 * first..first+count-1 (1-based): me receives count 24x24 column-major matrices; density has
 * 1 or ne entries (per element). */
int pf_hex8_mass_matrix(pf_ctx *ctx, int64_t nn, int64_t ne, const double *nodcoord, const int32_t *elemnod,
                        const double *density, int64_t ndensity, int64_t first, int64_t count, double *me);

/* StiffnessMatrix_as_CRS_FEMDomain(YoungModulus(:),PoissonRatio(:))
 * (FEMDomainClass.f90:11123-11232): val (nnz) of the ZeroMatrix pattern given by
 * row_ptr/col_idx, summed in ascending element order. */
int pf_hex8_stiffness_crs(pf_ctx *ctx, int64_t nn, int64_t ne, const double *nodcoord,
                          const int32_t *elemnod, const double *E, const double *nu,
                          int32_t jmode, const int64_t *row_ptr, const int32_t *col_idx,
                          double *val);

/* sub_crs_matvec (LinearSolverClass.f90:3728-3776): y = A x, 1-based CRS, host arrays. */
int pf_spmv_crs(pf_ctx *ctx, int64_t n, const int64_t *row_ptr, const int32_t *col_idx,
                const double *val, const double *x, double *y);

/* bicgstab_CRS_2 (FEMSolverClass.f90:2199-2298) on host arrays; x holds x0 on entry and
 * the solution on exit.  This is the plug-in LinearSolver(row_ptr,col_idx,val,rhs,x) of
 * solveFEMSolver_UserDefinedLinearSolver (FEMSolverClass.f90:1638-1648); like that hook
 * it applies no Dirichlet elimination.  opts NULL = defaults; stats may be NULL. */
int pf_bicgstab_crs(pf_ctx *ctx, int64_t n, const int64_t *row_ptr, const int32_t *col_idx,
                    const double *val, const double *rhs, double *x, const pf_solve_opts *opts,
                    pf_stats *stats);

/* ---- preconditioners (solveFEMSolver preconditioning=, FEMSolverClass.f90:1768-1790) ---- */
/* JacobiPreconditionerCRS (FEMSolverClass.f90:2508-2529) in place on host arrays: every row
 * whose diagonal (last stored entry of the column) is nonzero is multiplied by its
 * reciprocal, rhs(i) too (rhs may be NULL). */
int pf_jacobi_crs(pf_ctx *ctx, int64_t n, const int64_t *row_ptr, const int32_t *col_idx, double *val,
                  double *rhs);

/* ILUCRS(fill_in_order=0) (SparseClass.f90:1629-1700): fval (nnz) receives the factor of val
 * on the same pattern, columns taken in their stored order as the reference does.  Rows may
 * hold at most 512 entries and no column twice (PF_INVALID otherwise). */
int pf_ilu0_crs(pf_ctx *ctx, int64_t n, const int64_t *row_ptr, const int32_t *col_idx, const double *val,
                double *fval);

/* ILU_matvecCRS (SparseClass.f90:1795-1845) of a factor fval: y = D^-1 L^-1 v, L the unit
 * lower triangle of fval and D its diagonal (1 where 0); the reference's backward loop
 * leaves its vector unchanged, so the strict upper triangle is not applied. */
int pf_ilu_apply_crs(pf_ctx *ctx, int64_t n, const int64_t *row_ptr, const int32_t *col_idx, const double *fval,
                     const double *v, double *y);

/* ---- device-pointer entry points (torch interop) ---------------------------------- */
/* y = A x on device arrays; 1-based CRS as above. */
int pf_dev_spmv_crs(pf_ctx *ctx, int64_t n, const int64_t *row_ptr, const int32_t *col_idx,
                    const double *val, const double *x, double *y);
int pf_dev_spmv_crs_f32(pf_ctx *ctx, int64_t n, const int64_t *row_ptr, const int32_t *col_idx,
                        const float *val, const float *x, float *y);

/* crs_spmv_real64 / crs_spmv_real32 (SparseClass-c.c:38-66, bound in SparseClass.f90:24-50):
 * new_vector += A old_vector on host arrays (the product is ADDED, as the reference's C
 * kernels do; matmul_c_CRS zeroes new_vector first, SparseClass.f90:727-742).  1-based
 * row_ptr (n+1) and col_idx; col_size = size(col_idx) >= row_ptr(n+1)-1.  FP32 values and
 * vectors accumulate in FP32 like crs_spmv_real32. */
int pf_crs_spmv_real64(pf_ctx *ctx, const double *val, const int64_t *row_ptr, const int32_t *col_idx,
                       const double *old_vector, double *new_vector, int64_t n, int64_t col_size);
int pf_crs_spmv_real32(pf_ctx *ctx, const float *val, const int64_t *row_ptr, const int32_t *col_idx,
                       const float *old_vector, float *new_vector, int64_t n, int64_t col_size);
/* The same under the reference's own C names and signatures, so a program linked against this
 * library in place of SparseClass-c.o runs matmul_c_CRS / matmul_c_real32_CRS on the GPU
 * (device 0, a context created on first use).  These return nothing, like the reference's:
 * an invalid CRS or a HIP error prints a message and aborts the process. */
void crs_spmv_real64(double val[], long row_ptr[], int col_idx[], double old_vector[], double new_vector[],
                     size_t n, size_t col_size);
void crs_spmv_real32(float val[], long row_ptr[], int col_idx[], float old_vector[], float new_vector[],
                     size_t n, size_t col_size);

/* ---- allocation-free solve ---------------------------------------------------------
 * No counterpart in the reference: pf_bicgstab_crs (bicgstab_CRS_2, FEMSolverClass.f90:
 * 2199-2298) on DEVICE arrays for callers that cannot allocate device memory (code under
 * hipStream capture, tight serving loops, harnesses that forbid hipMalloc).  The work vectors,
 * partial sums and scalars live in a static device array of each GPU (about 30 MB, held until
 * the process ends), so at most PF_NOALLOC_MAX_ROWS rows; larger systems are rejected with
 * PF_INVALID.  Makes no device or host allocation and captures no hipGraph.  FP64 on the CRS
 * layout; opts as pf_bicgstab_crs (precision 64, preconditioning 0, stop_mode 0/1/2); x holds x0
 * on entry.  The CRS is checked on the device first (nondecreasing row_ptr from 1, columns in
 * 1..n); col_idx and val must hold row_ptr(n+1)-1 entries.  Calls on one GPU share its scratch
 * array: calls that overlap in time on one GPU race. */
#define PF_NOALLOC_MAX_ROWS 524288
int pf_dev_bicgstab_crs_noalloc(pf_ctx *ctx, int64_t n, const int64_t *row_ptr, const int32_t *col_idx,
                                const double *val, const double *rhs, double *x, const pf_solve_opts *opts,
                                pf_stats *stats);

/* ---- FEMSolver_ state on the device (FEMSolverClass.f90:13-164) ------------------ */
int pf_system_create(pf_ctx *ctx, pf_system **out);
int pf_system_destroy(pf_system *sys);

/* FEMDomain mesh upload (setDomain, FEMSolverClass.f90:329-419).  nne = 8 for hex8. */
int pf_system_set_mesh(pf_system *sys, int64_t nn, int64_t ne, int32_t nne,
                       const double *nodcoord, const int32_t *elemnod);

/* setCRS(DOF) single-domain branch (FEMSolverClass.f90:679-694): ZeroMatrix on the device,
 * zero RHS and x.  Returns nnz. */
int pf_system_zero_matrix(pf_system *sys, int32_t dof, int64_t *nnz);

/* setCRS(CRS_) (FEMSolverClass.f90:421-429): upload an existing CRS (host arrays). */
int pf_system_set_crs(pf_system *sys, int64_t n, int64_t nnz, const int64_t *row_ptr,
                      const int32_t *col_idx, const double *val);

/* Sizes of the current CRS. */
int pf_system_sizes(pf_system *sys, int64_t *n, int64_t *nnz, int64_t *nn, int64_t *ne);

/* getCRS: download (any pointer may be NULL). */
int pf_system_get_crs(pf_system *sys, int64_t *row_ptr, int32_t *col_idx, double *val);

/* StiffnessMatrix(YoungModulus(:),PoissonRatio(:)) assembled into the CRS values on the
 * device (overwrites val).  E, nu: host arrays of length ne, kept resident on the device;
 * E == nu == NULL reuses the resident arrays of the previous call. */
int pf_system_stiffness(pf_system *sys, const double *E, const double *nu, int32_t jmode);

/* Element-matrix kernel of later pf_system_stiffness / _dev calls on a single-domain CRS:
 * 0 (default) the reference's operation order of StiffnessMatrixFEMDomain
 * (FEMDomainClass.f90:11457-11515, CRS values bitwise); 1 the isotropic split
 * Ke(3a+d,3b+e) = lam M_ab(d,e) + mu M_ab(e,d) + mu [d==e] tr M_ab with
 * M_ab(p,q) = sum_gp detJ G(p,a) G(q,b), contraction on the vector units; 2 the same with the
 * contraction on the FP64 matrix cores.  1 and 2 agree with 0 to rounding (not bitwise).  No
 * counterpart in the reference: it selects how the element contraction is evaluated. */
int pf_system_ke_kernel(pf_system *sys, int32_t kind);

/* setMatrix(DomainID=1,ElementID,DOF,Matrix) for a batch (FEMSolverClass.f90:901-909,
 * :1020-1057): elem_ids 1-based, ke count matrices of (nne*dof)^2 column-major.  Adds into
 * val; a missing slot returns PF_INVALID like the reference's fill-in stop. */
int pf_system_add_element_matrices(pf_system *sys, int64_t count, const int32_t *elem_ids,
                                   const double *ke);

/* setRHS (FEMSolverClass.f90:431-437) / setVector batch (:913-920, :1059-1069) / getRHS */
int pf_system_set_rhs(pf_system *sys, const double *rhs);
int pf_system_add_element_vectors(pf_system *sys, int64_t count, const int32_t *elem_ids,
                                  const double *vec);
int pf_system_get_rhs(pf_system *sys, double *rhs);

/* TractionVector_by_elemFEMDomain summed over elements, added to the RHS
 * (FEMDomainClass.f90:17210-17292).  box = xmin,xmax,ymin,ymax,zmin,zmax (open box, as
 * Range_%inside).  pressure > 0 pushes along the outward facet normal. */
int pf_system_add_traction(pf_system *sys, const double *box, double pressure);

/* fix(DomainID=1,IDs,FixValue|FixValues) (FEMSolverClass.f90:1073-1136): ids are 1-based
 * DOF ids; ids < 1 or > n are skipped; values NULL -> fixvalue for all. */
int pf_system_fix(pf_system *sys, int64_t nids, const int32_t *ids, const double *values,
                  double fixvalue);

/* solve() (FEMSolverClass.f90:1697-1825): Dirichlet elimination on the device CRS and RHS
 * (destructive, as the reference), BiCGSTAB from x0 (NULL = zeros), x_fixed = ubar.
 * x_out (host, length n) may be NULL to keep x on the device. */
int pf_system_solve(pf_system *sys, const pf_solve_opts *opts, const double *x0, double *x_out,
                    pf_stats *stats);
int pf_system_get_x(pf_system *sys, double *x);

/* MassMatrix_as_CRS_FEMDomain(Density(:), DOF=3) (FEMDomainClass.f90:10364-10467) assembled
 * into the CRS values on the device (overwrites val, element contributions in ascending element
 * order); density: one value per element (host array), or NULL to use density_all. */
int pf_system_mass_matrix(pf_system *sys, const double *density, double density_all);

/* zerosFEMSolver (FEMSolverClass.f90:1310-1322): CRS values and RHS set to zero. */
int pf_system_zeros(pf_system *sys);

/* ---- several domains coupled by overset penalties (Soybean%getDisplacement) -------- */
/* FEMDomain%overset(FEMDomain, DomainID, algorithm=FEMDomain_Overset_GPP, MyDomainID)
 * (FEMDomainClass.f90:14606-14700): the Gauss points of mesh A (the own domain) that lie in
 * elements of mesh B (Mesh%nearestElementID, MeshClass.f90:8611-8866).  Meshes as
 * pf_system_set_mesh (hex8, nodcoord nn x 3 and elemnod ne x 8 column-major, 1-based).
 * Connects come in the reference's order (own elements ascending among those whose node
 * range meets B's node box, Gauss points 1..8): own element, Gauss point, element of B (all
 * 1-based) and the Gauss point's position (3 doubles).  At most cap are written; *count is
 * the number found (call again with a larger cap when it exceeds cap). */
int pf_overset_gpp(pf_ctx *ctx, int64_t nnA, int64_t neA, const double *nodcoordA, const int32_t *elemnodA,
                   int64_t nnB, int64_t neB, const double *nodcoordB, const int32_t *elemnodB, int64_t cap,
                   int64_t *count, int32_t *elem, int32_t *gauss_point, int32_t *pair_elem, double *position);

/* FEMDomain%overset(..., algorithm=FEMDomain_Overset_P2P) (FEMDomainClass.f90:14701-14760):
 * every node of mesh A that lies in an element of mesh B, in node order: own node id, element
 * of B (1-based) and the node's position.  As pf_overset_gpp otherwise. */
int pf_overset_p2p(pf_ctx *ctx, int64_t nnA, int64_t neA, const double *nodcoordA, const int32_t *elemnodA,
                   int64_t nnB, int64_t neB, const double *nodcoordB, const int32_t *elemnodB, int64_t cap,
                   int64_t *count, int32_t *node, int32_t *pair_elem, double *position);

/* setCRS(DOF) with several domains (FEMSolverClass.f90:466-676) on the mesh given to
 * pf_system_set_mesh, which is the domains' meshes concatenated (node and element ids of
 * domain d offset by the sizes of domains 1..d-1; CRS_ID_Starts_From(d) = 1 + dof * nodes
 * before d).  The pattern is the union of the ZeroMatrix blocks and of every node pair of
 * the ncon connects (elem, pair_elem: element ids of the concatenated mesh, 1-based; a
 * connect with gauss_point 0 is a P2P one whose elem is the own NODE id and which couples 9
 * nodes; gauss_point may be NULL when every connect is GPP), each row's columns ascending;
 * values zero.  any_overset = 0 (no domain was overset) gives the reference's empty CRS (every
 * row empty).  pf_system_stiffness then assembles into this pattern.  Returns nnz. */
int pf_system_multi_crs(pf_system *sys, int32_t dof, int32_t any_overset, int64_t ncon, const int32_t *elem,
                        const int32_t *gauss_point, const int32_t *pair_elem, int64_t *nnz);

/* setEbOM(penalty, DOF) (FEMSolverClass.f90:1891-1985): for each connect in order, adds
 * penalty * ConnectMatrix (FEMDomainClass.f90:13014-13100) over the 16 nodes of the own and
 * the paired element: A = (Bc^T Bc) detJ, Bc = [N_own(gp) I, -N_pair(xi) I], detJ of the own
 * Gauss point, xi the local coordinate of the position in the paired element.  Element ids of
 * the concatenated mesh (1-based), Gauss points 1..8, positions 3 per connect.  Gauss point 0
 * marks a P2P connect (elem = own node id): A = Bc^T Bc with Bc = [I, -N_pair(xi) I] over the
 * node and the 8 paired nodes, no det J (FEMDomainClass.f90:13100-13127).  DOF = 3. */
int pf_system_ebom(pf_system *sys, int64_t ncon, const int32_t *elem, const int32_t *gauss_point,
                   const int32_t *pair_elem, const double *position, double penalty, int32_t dof);

/* keepThisMatrixAs(As) (FEMSolverClass.f90:1544-1604): a named copy of the current CRS values
 * (the eigen solver reads "A" = stiffness and "B" = mass). */
int pf_system_keep_matrix(pf_system *sys, const char *name);

/* fix_eig(IDs) (FEMSolverClass.f90:1607-1628): 1-based DOF ids whose amplitudes are zero in
 * the eigen problem (merged with earlier calls, unique). */
int pf_system_fix_eig(pf_system *sys, int64_t nids, const int32_t *ids);

/* eigFEMSolver (FEMSolverClass.f90:1384-1541): A u = lambda B u with the kept "A" and "B", the
 * fix_eig DOFs removed.  nev <= 0: every eigenpair of the reduced problem, as the reference's
 * default LAPACK DSPGVD path (dense on the device, at most 16384 free DOFs); nev > 0: the nev
 * smallest by LOBPCG (its use_LOBPCG option; nev <= 32; tol on max_i |A x_i - l_i B x_i| / (max_j |l_j| |B x_i|),
 * default 1e-8 for tol <= 0; maxit default 10000).  eigen_values (ascending) and eigen_vectors
 * (n x count column-major, B-orthonormal, zero at the removed DOFs) are host arrays and may be
 * NULL; *count receives the number of pairs (pf_system_eig_count gives it for nev <= 0).
 * stats: status, iterations (LOBPCG), true_relres = the worst relative residual, solve_ms. */
int pf_system_eig_count(pf_system *sys, int32_t *count);
int pf_system_eig(pf_system *sys, int32_t nev, double tol, int32_t maxit, double *eigen_values,
                  double *eigen_vectors, int32_t *count, pf_stats *stats);

/* The ILU(0) factor kept by the system (FEMSolver_%ILU_MATRIX, FEMSolverClass.f90:67,
 * :1777-1784): pf_system_ilu_factor computes it from the eliminated matrix when none is kept
 * (elimination applied first) and downloads its values (fval may be NULL) and the number of
 * levels of its schedule (levels may be NULL); pf_system_ilu_apply computes y = M^-1 v with
 * it (host arrays of length n); pf_system_ilu_reset drops it so the next ILU solve factors
 * the current matrix again. */
int pf_system_ilu_factor(pf_system *sys, double *fval, int32_t *levels);
int pf_system_ilu_apply(pf_system *sys, const double *v, double *y);
int pf_system_ilu_reset(pf_system *sys);

/* y = A x with the system's current CRS (elimination included once solve() has run).
 * kind 1: CRS row kernel, FP64; 2: blocked kernel, FP64; 3: blocked kernel, FP32 values and
 * vectors (x rounded to FP32, y widened back); 4: CRS row kernel with x staged in LDS per
 * block of rows (the distinct columns of the block gathered once, 16-bit positions per entry;
 * the same sums as kind 1, bitwise).  x, y are host arrays of length n (x NULL:
 * reuse the last x; y NULL: keep on device).  The kernel is launched `repeat` times and *ms
 * (may be NULL) receives the average device time of one launch (hipEvents). */
int pf_system_spmv(pf_system *sys, int32_t kind, const double *x, double *y, int32_t repeat, double *ms);

/* Device pointers of the system's arrays, for torch interop (valid until the next call
 * that reallocates: set_mesh, zero_matrix, set_crs, destroy). */
typedef struct {
  int64_t n, nnz, nn, ne;
  int64_t *row_ptr;
  int32_t *col_idx;
  double *val;
  double *rhs;
  double *x;
  double *nodcoord;
  int32_t *elemnod;
} pf_system_view;
int pf_system_view_get(pf_system *sys, pf_system_view *view);

/* ---- batched form: many independent meshes in one system --------------------------
 * The reference solves organ meshes one at a time, one FEMSolver_ per organ
 * (Soybean%getDisplacement, SoybeanClass.f90:8417-8533; solve() per organ,
 * FEMSolverClass.f90:1697-1825).  Here the meshes are concatenated into one pf_system
 * (node ids of mesh m shifted by its first node, element lists concatenated) so mesh setup,
 * ZeroMatrix, assembly, tractions and fix run once for all of them; the CRS is then
 * block-diagonal and every mesh's block equals the reference's CRS of that mesh alone.
 * pf_system_batch_solve runs bicgstab_CRS_2 for every mesh in one persistent kernel. */

/* Traction boxes per mesh: box m (6 doubles, as pf_system_add_traction) and pressure m act
 * on the facets of mesh m only (nodes node_off[m] .. node_off[m+1]-1, 0-based). */
int pf_system_add_tractions(pf_system *sys, int32_t nmesh, const int64_t *node_off, const double *boxes,
                            const double *pressures);

/* Declare the mesh partition: node_off[0] = 0 < node_off[1] < ... < node_off[nmesh] = nn.
 * The CRS must be block-diagonal over it (checked).  cost (may be NULL) sets the order in
 * which meshes are taken from the queue, largest first (default: the mesh's nonzeros). */
int pf_system_batch_set(pf_system *sys, int32_t nmesh, const int64_t *node_off, const double *cost);

/* Launch shape of the persistent kernel. */
typedef struct {
  int32_t layout;        /* 1: blocked Sell3 (64-node slices, DOF=3 node blocks); 0: CRS rows */
  int32_t sub;           /* layout 0: lanes per CRS row (8, 16 or 32) */
  int32_t block;         /* threads per workgroup: 256, 512 or 1024 */
  int32_t blocks_per_cu; /* resident workgroups per CU in the grid (0: the occupancy limit) */
} pf_batch_tune;

/* Result of one mesh. */
typedef struct {
  int32_t status;        /* PF_OK, PF_BREAKDOWN_*, PF_NOT_CONVERGED */
  int32_t iterations;    /* BiCGSTAB iterations of this mesh */
  int32_t replacements;  /* precision 32: FP64 residual replacements */
  int32_t workgroup;     /* workgroup that solved it */
  double rr, rr1;        /* last r.r and r.r after iteration 1 (precision 64) */
  double true_relres;    /* FP64 ||b - A x|| / ||b|| of this mesh */
  double start_ms, end_ms; /* device wall clock, from the start of the first mesh */
} pf_batch_mesh_stats;

/* Solve every mesh (x0 = 0) with the options' precision (64 or 32), stop rule, itrmax and
 * tolerances, applying each mesh's Dirichlet elimination first as pf_system_solve does.
 * tune NULL: the shape chosen by pf_system_batch_autotune, else a default.  The queue
 * reset and the kernel are captured in one hipGraph (opts->use_graph).  opts->preconditioning
 * 1 applies JacobiPreconditionerCRS to the eliminated system first (in place, row-local, so
 * each mesh is scaled alone); 0 none; others are rejected.  mesh_stats (nmesh entries) and
 * x_out (n) may be NULL.  Returns PF_OK when every mesh converged, else the status of the
 * first mesh (in mesh order) that did not. */
int pf_system_batch_solve(pf_system *sys, const pf_solve_opts *opts, const pf_batch_tune *tune, double *x_out,
                          pf_batch_mesh_stats *mesh_stats, pf_stats *total);

/* Autotuner: times every launch shape (blocked / CRS with 8, 16, 32 lanes per row; 256,
 * 512, 1024 threads per workgroup) on this batch with `iters` iterations per mesh, keeps
 * the fastest for later pf_system_batch_solve calls of the same precision and returns it
 * in *best; up to max_results (shape, ms) pairs go to tried / tried_ms, *ntried = count.
 * Overwrites the system's x. */
int pf_system_batch_autotune(pf_system *sys, const pf_solve_opts *opts, int32_t iters, pf_batch_tune *best,
                             int32_t max_results, pf_batch_tune *tried, double *tried_ms, int32_t *ntried);

/* ---- direct solve: solve(algorithm="GaussJordan") ------------------------------------
 * The reference's solve() routes algorithm="GaussJordan" (FEMSolverClass.f90:1795-1798) to
 * gauss_jordan_crs (FEMSolverClass.f90:2413-2460), which stops "Under construction" after its
 * first pivot test.  Here that branch solves the eliminated system directly: each mesh's
 * nodes (DOF=3 node blocks) or DOFs are renumbered by the candidate of smallest bandwidth b
 * (natural, reverse Cuthill-McKee, coordinate sorts), the symmetric positive definite matrix
 * is factored L L^T in band storage (n x (b+1) doubles per mesh, PF_ERR_NOMEM when that does
 * not fit the device), then forward/backward substitution and `refine` steps of iterative
 * refinement on the FP64 residual b - A x of the CRS values.  The factor is kept while the
 * matrix is unchanged: a later solve with a new right-hand side only substitutes.  A
 * non-positive pivot gives PF_NOT_SPD and x = 0 for that mesh. */

/* Whole system as one mesh.  stats: status, iterations = 1 + refine substitutions,
 * outer_iterations = refine, true_relres (FP64, eliminated system), setup_ms (band assembly,
 * device), solve_ms (factorization, substitutions and residuals, device). */
int pf_system_solve_direct(pf_system *sys, int32_t refine, double *x_out, pf_stats *stats);

/* Every mesh of a batch (pf_system_batch_set) in one persistent kernel: one workgroup per mesh
 * from a largest-first queue (cost = n b^2).  mesh_stats (may be NULL): status, iterations
 * (1 + refine), true_relres, start_ms/end_ms of the mesh's factorization and first
 * substitution.  Returns PF_OK, or the status of the first mesh (in mesh order) that failed. */
int pf_system_batch_direct(pf_system *sys, int32_t refine, double *x_out, pf_batch_mesh_stats *mesh_stats,
                           pf_stats *total);

/* Ordering and factor of mesh `mesh` (0 for pf_system_solve_direct) of the last direct solve:
 * *n DOFs, half bandwidth *bandwidth, *ordering 0 natural, 1 reverse Cuthill-McKee, 2/3/4
 * coordinate sort with x/y/z the slowest key; perm (n entries, may be NULL): mesh-local DOF
 * (0-based) of band row q; factor (n*(b+1), may be NULL): L of the permuted matrix in LAPACK
 * lower band layout, factor[c*(b+1) + r - c] = L[r][c]. */
int pf_system_direct_info(pf_system *sys, int32_t mesh, int64_t *n, int32_t *bandwidth, int32_t *ordering,
                          int32_t *perm, double *factor);

/* Tile products of the factorization: 1 (default) FP64 MFMA 16x16x4, 0 the same products on the
 * vector units.  Drops a held factor. */
int pf_system_direct_kernel(pf_system *sys, int32_t mfma);

/* ---- body force and stress recovery ----------------------------------------------------- */

/* MassVectorFEMDomain(ElementID, Density, DOF=3, Accel) (FEMDomainClass.f90:10993-11073) for
 * elements first..first+count-1 (1-based): mv[q*24 + 3a + k] = rho sum_gp N_a accel_k detJ. */
int pf_hex8_mass_vector(pf_ctx *ctx, int64_t nn, int64_t ne, const double *nodcoord, const int32_t *elemnod,
                        double density, const double *accel, int64_t first, int64_t count, double *mv);

/* setVector(MassVector(...)) for every element in ascending order, added to the RHS (the body
 * force loop of Tutorial/fem/benchmark_bar_in_gravity_field.f90:20-31).  density: per element
 * (ne values) or NULL for density_all; accel: 3 values. */
int pf_system_add_body_force(pf_system *sys, const double *density, double density_all, const double *accel);

/* getElementCauchyStress / StressMatrix without GaussPoint (FEMDomainClass.f90:15249-15380,
 * :11912-12032): the cell average over the 8 Gauss points of D B u_e, per element:
 * sigma[e*9 + 3j + i] = sigma(i+1, j+1).  disp: node-major displacement (the solver's x);
 * E, nu: 1 or ne values, as the reference accepts. */
int pf_element_stress(pf_ctx *ctx, int64_t nn, int64_t ne, const double *nodcoord, const int32_t *elemnod,
                      const double *disp, const double *E, int64_t nE, const double *nu, int64_t nnu,
                      int32_t jmode, double *sigma);
/* the same for the system's mesh and its current solution x */
int pf_system_element_stress(pf_system *sys, const double *E, int64_t nE, const double *nu, int64_t nnu,
                             int32_t jmode, double *sigma);

/* ---- differentiable form: device arrays, many load cases ----------------------------
 * The reference has no adjoint; these entry points serve a torch.autograd.Function whose
 * forward is the static solve (StiffnessMatrix(YoungModulus,PoissonRatio) + fix + solve,
 * FEMDomainClass.f90:11123-11232, FEMSolverClass.f90:1697-1825) for R load cases, and whose
 * backward is one more solve with K^T = K plus a per-element contraction. */

/* StiffnessMatrix(YoungModulus(:),PoissonRatio(:)) from DEVICE arrays of length ne. */
int pf_system_stiffness_dev(pf_system *sys, const double *E_dev, const double *nu_dev, int32_t jmode);

/* bicgstab_CRS_2 for nrhs (1..64) right-hand sides of the system's matrix at once, after the
 * Dirichlet elimination (the fixed rows of every right-hand side are set to 0; nonzero
 * prescribed values are rejected).  B_dev, X_dev: DEVICE arrays, n x nrhs interleaved
 * (dof i of case l at i*nrhs + l); x0 = 0.  Every case stops on its own by the options' stop
 * rule; per_rhs (nrhs entries) and total may be NULL.  Returns PF_OK when every case
 * converged, else the status of the first case that did not. */
int pf_system_solve_multi_dev(pf_system *sys, int32_t nrhs, const double *B_dev, double *X_dev,
                              const pf_solve_opts *opts, pf_stats *per_rhs, pf_stats *total);

/* out_dev[e] = sum_l lam_e(l)^T Ke(E_e, nu_e) u_e(l) for every element (DEVICE arrays, U and
 * Lam interleaved as above, fixed DOFs excluded).  Ke is linear in E, so this is
 * d/d(log E_e) of sum_l lam(l)^T K u(l); the adjoint gradient of a loss L(u) is -out with
 * lam solving K lam = dL/du. */
int pf_system_elem_sensitivity_dev(pf_system *sys, int32_t nrhs, const double *U_dev, const double *Lam_dev,
                                   double *out_dev);

#ifdef __cplusplus
}
#endif
#endif /* PFGPU_H */
