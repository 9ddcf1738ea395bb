! plantfem_gpu -- Fortran FEMDomain_/FEMSolver_ surface of plantFEM's static linear-elastic
! path, calling the gfx950 library libpfgpu.so through iso_c_binding (include/pfgpu.h).
!
! The type-bound names and argument meanings follow the reference:
!   FEMDomain_%create("Cube3D"|"Leaf3D"|"Cylinder",x_num,...,z_len)  FEMDomainClass.f90:7215-7310,
!                                                                   MeshClass.f90:6017-6276, :6506-6743
!   FEMDomain_%ZeroMatrix(DOF)                                      FEMDomainClass.f90:10581-10840
!   FEMDomain_%StiffnessMatrix(ElementID,E,v)                       FEMDomainClass.f90:11457-11515
!   FEMDomain_%StiffnessMatrix(YoungModulus,PoissonRatio)           FEMDomainClass.f90:11123-11232
!   FEMDomain_%select(x_min,...,z_max)                              FEMDomainClass.f90:12854-12926
!   FEMSolver_%init/setDomain/setCRS/setRHS/setMatrix/setVector/fix/solve
!                                                                   FEMSolverClass.f90:84-163
!   FEMDomain_%overset(FEMDomain,DomainID,algorithm,MyDomainID)     FEMDomainClass.f90:14606-14700
!   FEMSolver_%init(NumDomain)/setDomain(DomainID)/setCRS/setEbOM   FEMSolverClass.f90:466-676,
!                                                                   :1891-1985 (several domains)
! plus pf_gpu_bicgstab, a subroutine with the exact LinearSolver interface of
! solveFEMSolver_UserDefinedLinearSolver (FEMSolverClass.f90:1638-1648), so the reference's own
! FEMSolver_ can hand its CRS system to the GPU: x = solver%solve(LinearSolver=pf_gpu_bicgstab).
module plantfem_gpu
   use iso_c_binding
   use iso_fortran_env, only: int32, int64, real64, error_unit
   implicit none
   private

   public :: Mesh_, CRS_, FEMDomain_, FEMSolver_, pf_solve_opts, pf_stats
   public :: pf_gpu_bicgstab, pf_gpu_bicgstab_func, pf_gpu_bicgstab_crs, pf_gpu_spmv, pf_gpu_version, FEMDomainp_
   integer(int32), parameter, public :: FEMDomain_Overset_GPP = 1, FEMDomain_Overset_P2P = 2

   integer(c_int), parameter, public :: PF_OK = 0, PF_BREAKDOWN_VV = 1, PF_BREAKDOWN_C2C3 = 2, &
                                        PF_NOT_CONVERGED = 3, PF_INVALID = 4, PF_NOT_SPD = 5

   type, bind(C) :: pf_solve_opts
      integer(c_int32_t) :: itrmax, precision
      real(c_double) :: er0, relative_er, inner_tol, target_relres
      integer(c_int32_t) :: preconditioning, spmv_kind, check_every, use_graph, debug, stop_mode, &
                            kernel_timing, jacobi, inner_max, use_x0
   end type

   type, bind(C) :: pf_stats
      integer(c_int32_t) :: status, iterations, outer_iterations, spmv_kind_used
      real(c_double) :: rr, rr1, true_relres, solve_ms, setup_ms, spmv_p_ms, spmv_e_ms, update_ms, pupdate_ms
      integer(c_int64_t) :: timed_iterations
   end type

   interface
      subroutine c_pf_solve_opts_default(o) bind(C, name="pf_solve_opts_default")
         import :: pf_solve_opts
         type(pf_solve_opts), intent(inout) :: o
      end subroutine
      function c_pf_last_error() bind(C, name="pf_last_error")
         import :: c_ptr
         type(c_ptr) :: c_pf_last_error
      end function
      integer(c_int) function c_pf_version() bind(C, name="pf_version")
         import :: c_int
      end function
      integer(c_int) function c_pf_ctx_create(device, ctx) bind(C, name="pf_ctx_create")
         import :: c_int, c_ptr
         integer(c_int), value :: device
         type(c_ptr), intent(out) :: ctx
      end function
      integer(c_int) function c_pf_crs_graph(ctx, nn, ne, nne, elemnod, dof, nnz, row_ptr, col_idx) &
         bind(C, name="pf_crs_graph")
         import :: c_int, c_ptr, c_int64_t, c_int32_t
         type(c_ptr), value :: ctx
         integer(c_int64_t), value :: nn, ne
         integer(c_int32_t), value :: nne, dof
         integer(c_int32_t), intent(in) :: elemnod(*)
         integer(c_int64_t), intent(out) :: nnz
         type(c_ptr), value :: row_ptr, col_idx
      end function
      integer(c_int) function c_pf_hex8_stiffness(ctx, nn, ne, nodcoord, elemnod, E, nu, jmode, first, count, ke) &
         bind(C, name="pf_hex8_stiffness")
         import :: c_int, c_ptr, c_int64_t, c_int32_t, c_double
         type(c_ptr), value :: ctx
         integer(c_int64_t), value :: nn, ne, first, count
         real(c_double), intent(in) :: nodcoord(*), E(*), nu(*)
         integer(c_int32_t), intent(in) :: elemnod(*)
         integer(c_int32_t), value :: jmode
         real(c_double), intent(out) :: ke(*)
      end function
      integer(c_int) function c_pf_hex8_stiffness_crs(ctx, nn, ne, nodcoord, elemnod, E, nu, jmode, row_ptr, &
                                                      col_idx, val) bind(C, name="pf_hex8_stiffness_crs")
         import :: c_int, c_ptr, c_int64_t, c_int32_t, c_double
         type(c_ptr), value :: ctx
         integer(c_int64_t), value :: nn, ne
         real(c_double), intent(in) :: nodcoord(*), E(*), nu(*)
         integer(c_int32_t), intent(in) :: elemnod(*), col_idx(*)
         integer(c_int64_t), intent(in) :: row_ptr(*)
         integer(c_int32_t), value :: jmode
         real(c_double), intent(out) :: val(*)
      end function
      integer(c_int) function c_pf_spmv_crs(ctx, n, row_ptr, col_idx, val, x, y) bind(C, name="pf_spmv_crs")
         import :: c_int, c_ptr, c_int64_t, c_int32_t, c_double
         type(c_ptr), value :: ctx
         integer(c_int64_t), value :: n
         integer(c_int64_t), intent(in) :: row_ptr(*)
         integer(c_int32_t), intent(in) :: col_idx(*)
         real(c_double), intent(in) :: val(*), x(*)
         real(c_double), intent(out) :: y(*)
      end function
      integer(c_int) function c_pf_bicgstab_crs(ctx, n, row_ptr, col_idx, val, rhs, x, opts, stats) &
         bind(C, name="pf_bicgstab_crs")
         import :: c_int, c_ptr, c_int64_t, c_int32_t, c_double, pf_solve_opts, pf_stats
         type(c_ptr), value :: ctx
         integer(c_int64_t), value :: n
         integer(c_int64_t), intent(in) :: row_ptr(*)
         integer(c_int32_t), intent(in) :: col_idx(*)
         real(c_double), intent(in) :: val(*), rhs(*)
         real(c_double), intent(inout) :: x(*)
         type(pf_solve_opts), intent(in) :: opts
         type(pf_stats), intent(out) :: stats
      end function
      integer(c_int) function c_pf_system_create(ctx, sys) bind(C, name="pf_system_create")
         import :: c_int, c_ptr
         type(c_ptr), value :: ctx
         type(c_ptr), intent(out) :: sys
      end function
      integer(c_int) function c_pf_system_destroy(sys) bind(C, name="pf_system_destroy")
         import :: c_int, c_ptr
         type(c_ptr), value :: sys
      end function
      integer(c_int) function c_pf_system_set_mesh(sys, nn, ne, nne, nodcoord, elemnod) &
         bind(C, name="pf_system_set_mesh")
         import :: c_int, c_ptr, c_int64_t, c_int32_t, c_double
         type(c_ptr), value :: sys
         integer(c_int64_t), value :: nn, ne
         integer(c_int32_t), value :: nne
         real(c_double), intent(in) :: nodcoord(*)
         integer(c_int32_t), intent(in) :: elemnod(*)
      end function
      integer(c_int) function c_pf_system_zero_matrix(sys, dof, nnz) bind(C, name="pf_system_zero_matrix")
         import :: c_int, c_ptr, c_int64_t, c_int32_t
         type(c_ptr), value :: sys
         integer(c_int32_t), value :: dof
         integer(c_int64_t), intent(out) :: nnz
      end function
      integer(c_int) function c_pf_system_set_crs(sys, n, nnz, row_ptr, col_idx, val) bind(C, name="pf_system_set_crs")
         import :: c_int, c_ptr, c_int64_t, c_int32_t, c_double
         type(c_ptr), value :: sys
         integer(c_int64_t), value :: n, nnz
         integer(c_int64_t), intent(in) :: row_ptr(*)
         integer(c_int32_t), intent(in) :: col_idx(*)
         real(c_double), intent(in) :: val(*)
      end function
      integer(c_int) function c_pf_system_get_crs(sys, row_ptr, col_idx, val) bind(C, name="pf_system_get_crs")
         import :: c_int, c_ptr
         type(c_ptr), value :: sys, row_ptr, col_idx, val
      end function
      integer(c_int) function c_pf_system_stiffness(sys, E, nu, jmode) bind(C, name="pf_system_stiffness")
         import :: c_int, c_ptr, c_int32_t, c_double
         type(c_ptr), value :: sys
         real(c_double), intent(in) :: E(*), nu(*)
         integer(c_int32_t), value :: jmode
      end function
      integer(c_int) function c_pf_system_add_element_matrices(sys, count, ids, ke) &
         bind(C, name="pf_system_add_element_matrices")
         import :: c_int, c_ptr, c_int64_t, c_int32_t, c_double
         type(c_ptr), value :: sys
         integer(c_int64_t), value :: count
         integer(c_int32_t), intent(in) :: ids(*)
         real(c_double), intent(in) :: ke(*)
      end function
      integer(c_int) function c_pf_system_add_element_vectors(sys, count, ids, vec) &
         bind(C, name="pf_system_add_element_vectors")
         import :: c_int, c_ptr, c_int64_t, c_int32_t, c_double
         type(c_ptr), value :: sys
         integer(c_int64_t), value :: count
         integer(c_int32_t), intent(in) :: ids(*)
         real(c_double), intent(in) :: vec(*)
      end function
      integer(c_int) function c_pf_system_set_rhs(sys, rhs) bind(C, name="pf_system_set_rhs")
         import :: c_int, c_ptr, c_double
         type(c_ptr), value :: sys
         real(c_double), intent(in) :: rhs(*)
      end function
      integer(c_int) function c_pf_system_get_rhs(sys, rhs) bind(C, name="pf_system_get_rhs")
         import :: c_int, c_ptr, c_double
         type(c_ptr), value :: sys
         real(c_double), intent(out) :: rhs(*)
      end function
      integer(c_int) function c_pf_system_add_traction(sys, box, pressure) bind(C, name="pf_system_add_traction")
         import :: c_int, c_ptr, c_double
         type(c_ptr), value :: sys
         real(c_double), intent(in) :: box(6)
         real(c_double), value :: pressure
      end function
      integer(c_int) function c_pf_system_fix(sys, nids, ids, values, fixvalue) bind(C, name="pf_system_fix")
         import :: c_int, c_ptr, c_int64_t, c_int32_t, c_double
         type(c_ptr), value :: sys
         integer(c_int64_t), value :: nids
         integer(c_int32_t), intent(in) :: ids(*)
         type(c_ptr), value :: values
         real(c_double), value :: fixvalue
      end function
      integer(c_int) function c_pf_system_solve(sys, opts, x0, x_out, stats) bind(C, name="pf_system_solve")
         import :: c_int, c_ptr, pf_solve_opts, pf_stats
         type(c_ptr), value :: sys
         type(pf_solve_opts), intent(in) :: opts
         type(c_ptr), value :: x0, x_out
         type(pf_stats), intent(out) :: stats
      end function
      integer(c_int) function c_pf_system_solve_direct(sys, refine, x_out, stats) bind(C, name="pf_system_solve_direct")
         import :: c_int, c_ptr, c_int32_t, pf_stats
         type(c_ptr), value :: sys
         integer(c_int32_t), value :: refine
         type(c_ptr), value :: x_out
         type(pf_stats), intent(out) :: stats
      end function
      integer(c_int) function c_pf_hex8_mass_matrix(ctx, nn, ne, nodcoord, elemnod, density, ndensity, first, &
                                                    count, me) bind(C, name="pf_hex8_mass_matrix")
         import :: c_int, c_ptr, c_int64_t, c_int32_t, c_double
         type(c_ptr), value :: ctx
         integer(c_int64_t), value :: nn, ne, ndensity, first, count
         real(c_double), intent(in) :: nodcoord(*), density(*)
         integer(c_int32_t), intent(in) :: elemnod(*)
         real(c_double), intent(out) :: me(*)
      end function
      integer(c_int) function c_pf_system_zeros(sys) bind(C, name="pf_system_zeros")
         import :: c_int, c_ptr
         type(c_ptr), value :: sys
      end function
      integer(c_int) function c_pf_system_keep_matrix(sys, name) bind(C, name="pf_system_keep_matrix")
         import :: c_int, c_ptr, c_char
         type(c_ptr), value :: sys
         character(kind=c_char), intent(in) :: name(*)
      end function
      integer(c_int) function c_pf_system_fix_eig(sys, nids, ids) bind(C, name="pf_system_fix_eig")
         import :: c_int, c_ptr, c_int64_t, c_int32_t
         type(c_ptr), value :: sys
         integer(c_int64_t), value :: nids
         integer(c_int32_t), intent(in) :: ids(*)
      end function
      integer(c_int) function c_pf_system_eig_count(sys, count) bind(C, name="pf_system_eig_count")
         import :: c_int, c_ptr, c_int32_t
         type(c_ptr), value :: sys
         integer(c_int32_t), intent(out) :: count
      end function
      integer(c_int) function c_pf_system_eig(sys, nev, tol, maxit, values, vectors, count, stats) &
         bind(C, name="pf_system_eig")
         import :: c_int, c_ptr, c_int32_t, c_double, pf_stats
         type(c_ptr), value :: sys
         integer(c_int32_t), value :: nev, maxit
         real(c_double), value :: tol
         real(c_double), intent(out) :: values(*), vectors(*)
         integer(c_int32_t), intent(out) :: count
         type(pf_stats), intent(out) :: stats
      end function
      integer(c_int) function c_pf_overset_gpp(ctx, nnA, neA, xyzA, enA, nnB, neB, xyzB, enB, cap, count, &
                                                elem, gp, pair, pos) bind(C, name="pf_overset_gpp")
         import :: c_int, c_int64_t, c_int32_t, c_double, c_ptr
         type(c_ptr), value :: ctx
         integer(c_int64_t), value :: nnA, neA, nnB, neB, cap
         real(c_double), intent(in) :: xyzA(*), xyzB(*)
         integer(c_int32_t), intent(in) :: enA(*), enB(*)
         integer(c_int64_t), intent(out) :: count
         integer(c_int32_t), intent(out) :: elem(*), gp(*), pair(*)
         real(c_double), intent(out) :: pos(*)
      end function
      integer(c_int) function c_pf_overset_p2p(ctx, nnA, neA, xyzA, enA, nnB, neB, xyzB, enB, cap, count, &
                                                node, pair, pos) bind(C, name="pf_overset_p2p")
         import :: c_int, c_int64_t, c_int32_t, c_double, c_ptr
         type(c_ptr), value :: ctx
         integer(c_int64_t), value :: nnA, neA, nnB, neB, cap
         real(c_double), intent(in) :: xyzA(*), xyzB(*)
         integer(c_int32_t), intent(in) :: enA(*), enB(*)
         integer(c_int64_t), intent(out) :: count
         integer(c_int32_t), intent(out) :: node(*), pair(*)
         real(c_double), intent(out) :: pos(*)
      end function
      integer(c_int) function c_pf_system_multi_crs(sys, dof, any_overset, ncon, elem, gp, pair, nnz) &
         bind(C, name="pf_system_multi_crs")
         import :: c_int, c_int64_t, c_int32_t, c_ptr
         type(c_ptr), value :: sys
         integer(c_int32_t), value :: dof, any_overset
         integer(c_int64_t), value :: ncon
         integer(c_int32_t), intent(in) :: elem(*), gp(*), pair(*)
         integer(c_int64_t), intent(out) :: nnz
      end function
      integer(c_int) function c_pf_system_ebom(sys, ncon, elem, gp, pair, pos, penalty, dof) &
         bind(C, name="pf_system_ebom")
         import :: c_int, c_int64_t, c_int32_t, c_double, c_ptr
         type(c_ptr), value :: sys
         integer(c_int64_t), value :: ncon
         integer(c_int32_t), intent(in) :: elem(*), gp(*), pair(*)
         real(c_double), intent(in) :: pos(*)
         real(c_double), value :: penalty
         integer(c_int32_t), value :: dof
      end function
      function c_strlen(s) bind(C, name="strlen")
         import :: c_ptr, c_size_t
         type(c_ptr), value :: s
         integer(c_size_t) :: c_strlen
      end function
   end interface

   type :: Mesh_
      real(real64), allocatable :: NodCoord(:, :)
      integer(int32), allocatable :: ElemNod(:, :)
   end type

   type :: CRS_
      integer(int32), allocatable :: col_idx(:)
      integer(int64), allocatable :: row_ptr(:)
      real(real64), allocatable :: val(:)
   end type

   type :: FEMDomain_
      type(Mesh_) :: mesh
      ! OversetConnect_ list as arrays: own element (GPP) or own node (P2P), Gauss point (0 for
      ! P2P), paired element (ids in the paired domain), position (3, k), DomainIDs12 (2, k);
      ! allocated by the first overset
      integer(int32) :: num_oversetconnect = 0, overset_kind = 0  ! 0 none, 1 GPP, 2 P2P
      integer(int32), allocatable :: ov_elem(:), ov_gp(:), ov_pair(:), ov_domains(:, :)
      real(real64), allocatable :: ov_position(:, :)
   contains
      procedure :: create => createFEMDomain
      procedure :: move => moveFEMDomain
      procedure :: overset => oversetFEMDomain
      procedure :: numOversetElements => numOversetElementsFEMDomain
      procedure :: nn => nnFEMDomain
      procedure :: ne => neFEMDomain
      procedure :: nne => nneFEMDomain
      procedure :: nd => ndFEMDomain
      procedure :: z_min => z_minFEMDomain
      procedure :: z_max => z_maxFEMDomain
      procedure :: x_min => x_minFEMDomain
      procedure :: x_max => x_maxFEMDomain
      procedure :: select => selectFEMDomain
      procedure :: ZeroMatrix => ZeroMatrixFEMDomain
      procedure, pass :: StiffnessMatrixElementFEMDomain
      procedure, pass :: StiffnessMatrixCRSFEMDomain
      generic :: StiffnessMatrix => StiffnessMatrixElementFEMDomain, StiffnessMatrixCRSFEMDomain
      procedure :: MassMatrix => MassMatrixFEMDomain
   end type

   type :: FEMDomainp_
      type(FEMDomain_), pointer :: femdomainp => null()
   end type

   type :: FEMSolver_
      type(c_ptr) :: sys = c_null_ptr
      type(FEMDomain_), pointer :: domain => null()
      ! several domains: slots, the concatenated mesh, element offsets, CRS_ID_Starts_From
      type(FEMDomainp_), allocatable :: FEMDomains(:)
      type(FEMDomain_), pointer :: joint => null()
      integer(int32), allocatable :: elem_off(:), node_off(:)
      integer(int64), allocatable :: CRS_ID_Starts_From(:)
      logical :: initialized = .false.
      logical :: debug = .false.
      integer(int32) :: itrmax = 100000
      real(real64) :: er0 = dble(1.0e-10)
      real(real64) :: relative_er = dble(1.0e-10)
      integer(int32) :: DOF = 0
      integer(int64) :: n = 0, nnz = 0
      ! host copies, filled by getCRS() like the reference's public CRS_* components
      real(real64), allocatable :: CRS_val(:), CRS_RHS(:), CRS_x(:)
      integer(int32), allocatable :: CRS_Index_Col(:)
      integer(int64), allocatable :: CRS_Index_Row(:)
      type(pf_stats) :: stats
      ! setMatrix calls staged here and scattered on the device in call order
      integer(int32), allocatable, private :: pend_ids(:)
      real(real64), allocatable, private :: pend_mat(:)
      integer(int64), private :: npend = 0, mat_len = 0
   contains
      procedure :: init => initFEMSolver
      procedure :: setDomain => setDomainFEMSolver
      procedure, pass :: setCRSFEMSolver
      procedure, pass :: setCRS_CRSobjFEMSolver
      generic :: setCRS => setCRSFEMSolver, setCRS_CRSobjFEMSolver
      procedure :: setRHS => setRHSFEMSolver
      procedure :: getCRS => getCRSFEMSolver
      procedure :: saveMatrix => saveMatrixFEMSolver
      procedure :: setMatrix => setMatrixFEMSolver
      procedure :: setVector => setVectorFEMSolver
      procedure :: setStiffness => setStiffnessFEMSolver
      procedure :: addTraction => addTractionFEMSolver
      procedure :: fix => fixFEMSolver
      procedure :: solve => solveFEMSolver
      procedure :: flush => flushFEMSolver
      procedure :: remove => removeFEMSolver
      procedure :: zeros => zerosFEMSolver
      procedure :: keepThisMatrixAs => keepThisMatrixAsFEMSolver
      procedure :: fix_eig => fix_eigFEMSolver
      procedure :: eig => eigFEMSolver
      procedure :: setEbOM => setEbOMFEMSolver
   end type

   type(c_ptr), save :: g_ctx = c_null_ptr

contains

   ! ------------------------------------------------------------------------ plumbing
   function last_error() result(msg)
      character(len=:), allocatable :: msg
      type(c_ptr) :: p
      character(kind=c_char), pointer :: chars(:)
      integer(c_size_t) :: k, m
      p = c_pf_last_error()
      m = c_strlen(p)
      call c_f_pointer(p, chars, [m])
      allocate (character(len=m) :: msg)
      do k = 1, m
         msg(k:k) = chars(k)
      end do
   end function

   subroutine check(code, where)
      integer(c_int), intent(in) :: code
      character(*), intent(in) :: where
      if (code == PF_OK) return
      write (error_unit, '(a,a,a,i0,a,a)') "[plantfem_gpu] ", where, " failed (", code, "): ", last_error()
      error stop 1
   end subroutine

   function ctx() result(c)
      type(c_ptr) :: c
      integer(c_int) :: dev, code
      character(len=16) :: env
      integer :: stat, ln
      if (.not. c_associated(g_ctx)) then
         dev = 0
         call get_environment_variable("PF_DEVICE", env, ln, stat)
         if (stat == 0 .and. ln > 0) read (env, *) dev
         code = c_pf_ctx_create(dev, g_ctx)
         call check(code, "pf_ctx_create")
      end if
      c = g_ctx
   end function

   integer function pf_gpu_version()
      pf_gpu_version = int(c_pf_version())
   end function

   function default_opts() result(o)
      type(pf_solve_opts) :: o
      call c_pf_solve_opts_default(o)
   end function

   ! ------------------------------------------------------------------------ FEMDomain_
   ! Cube3D numbering (MeshClass.f90:6745-6808, :6992-7090): x fastest, lx/dble(xn)*dble(i-1)
   subroutine createFEMDomain(this, meshtype, x_num, y_num, z_num, x_len, y_len, z_len, shaperatio, species, &
                              SoyWidthRatio)
      class(FEMDomain_), intent(inout) :: this
      character(*), intent(in) :: meshtype
      integer(int32), optional, intent(in) :: x_num, y_num, z_num, species
      real(real64), optional, intent(in) :: x_len, y_len, z_len, shaperatio, SoyWidthRatio
      integer(int32) :: xn, yn, zn
      real(real64) :: lx, ly, lz
      xn = 10; yn = 10; zn = 10; lx = 1.0d0; ly = 1.0d0; lz = 1.0d0
      if (present(x_num)) xn = x_num
      if (present(y_num)) yn = y_num
      if (present(z_num)) zn = z_num
      if (present(x_len)) lx = x_len
      if (present(y_len)) ly = y_len
      if (present(z_len)) lz = z_len
      select case (meshtype)
      case ("Cube3D", "Cube", "cube3D", "rectangular3D")
         call cube3d_mesh(this%mesh, xn, yn, zn, lx, ly, lz)
      case ("Leaf3D")
         call cube3d_mesh(this%mesh, xn, yn, zn, lx, ly, lz)
         call leaf3d_shape(this%mesh, shaperatio, species, SoyWidthRatio)
      case ("Cylinder")
         call cylinder3d_mesh(this%mesh, xn, yn, zn, lx, ly, lz)
      case ("Cylinder3D")
         ! FEMDomain_%create sends "Cylinder3D" to to_cylinder (FEMDomainClass.f90:7279-7282,
         ! :16136-16157), a different mesh from Mesh%create("Cylinder"); not built here
         stop "plantfem_gpu: meshtype Cylinder3D (to_cylinder) is not on this path; use Cylinder"
      case default
         stop "plantfem_gpu: meshtype is Cube3D, Cube, Cylinder or Leaf3D on this path"
      end select
   end subroutine

   subroutine cube3d_mesh(mesh, xn, yn, zn, lx, ly, lz)
      type(Mesh_), intent(inout) :: mesh
      integer(int32), intent(in) :: xn, yn, zn
      real(real64), intent(in) :: lx, ly, lz
      integer(int32) :: i, j, k, n, e, a
      integer(int64) :: nn2, ne2
      integer(int32) :: q(4)
      nn2 = int(xn + 1, int64)*(yn + 1)
      ne2 = int(xn, int64)*yn
      if (allocated(mesh%NodCoord)) deallocate (mesh%NodCoord)
      if (allocated(mesh%ElemNod)) deallocate (mesh%ElemNod)
      allocate (mesh%NodCoord(nn2*(zn + 1), 3), mesh%ElemNod(ne2*zn, 8))
      n = 0
      do k = 1, zn + 1
         do j = 1, yn + 1
            do i = 1, xn + 1
               n = n + 1
               mesh%NodCoord(n, 1) = lx/dble(xn)*dble(i - 1)
               mesh%NodCoord(n, 2) = ly/dble(yn)*dble(j - 1)
               mesh%NodCoord(n, 3) = lz/dble(zn)*dble(k - 1)
            end do
         end do
      end do
      e = 0
      do k = 1, zn
         do j = 1, yn
            do i = 1, xn
               e = e + 1
               q = [i + (j - 1)*(xn + 1), i + 1 + (j - 1)*(xn + 1), xn + 2 + i + (j - 1)*(xn + 1), &
                    xn + 1 + i + (j - 1)*(xn + 1)]
               do a = 1, 4
                  mesh%ElemNod(e, a) = q(a) + int(nn2, int32)*(k - 1)
                  mesh%ElemNod(e, a + 4) = q(a) + int(nn2, int32)*k
               end do
            end do
         end do
      end do
   end subroutine

   ! Leaf3D (MeshClass.f90:6017-6276): the rectangular3D mesh (x width, y thickness, z length)
   ! centred in x and y, then the default outline (shaperatio) or the soybean one (species 1)
   subroutine leaf3d_shape(mesh, shaperatio, species, SoyWidthRatio)
      type(Mesh_), intent(inout) :: mesh
      real(real64), optional, intent(in) :: shaperatio, SoyWidthRatio
      integer(int32), optional, intent(in) :: species
      real(real64) :: s, thickness_, width, thickness_ratio, zl, z20, zc, ratio, length, sw, alpha, r, xx, yy, yy_, &
                      xmax, xmin, theta
      real(real64), parameter :: lin_curve_ratio = 0.50d0
      integer(int32) :: i, sp
      mesh%NodCoord(:, 1) = mesh%NodCoord(:, 1) - (maxval(mesh%NodCoord(:, 1)) - minval(mesh%NodCoord(:, 1)))*0.50d0
      mesh%NodCoord(:, 2) = mesh%NodCoord(:, 2) - (maxval(mesh%NodCoord(:, 2)) - minval(mesh%NodCoord(:, 2)))*0.50d0
      sp = 0
      if (present(species)) sp = species
      if (sp == 0) then
         s = 0.30d0
         if (present(shaperatio)) s = shaperatio
         thickness_ = maxval(mesh%NodCoord(:, 2)) - minval(mesh%NodCoord(:, 2))
         width = maxval(mesh%NodCoord(:, 1)) - minval(mesh%NodCoord(:, 1))
         thickness_ratio = (width*1.0d0/10.0d0)/thickness_
         zl = maxval(mesh%NodCoord(:, 3)) - minval(mesh%NodCoord(:, 3))
         z20 = 1.0d0/20.0d0*zl
         do i = 1, size(mesh%NodCoord, 1)
            zc = mesh%NodCoord(i, 3)
            if (zc <= z20) then
               ratio = 1.0d0/10.0d0
               mesh%NodCoord(i, 2) = mesh%NodCoord(i, 2)*thickness_ratio
            elseif (zc <= zl*s) then
               ratio = 1.0d0/10.0d0 + 0.90d0/(zl*s - z20)*(zc - z20)
            else
               ratio = 1.0d0 - 0.90d0/(zl - s*zl)*(zc - s*zl)
            end if
            mesh%NodCoord(i, 1) = mesh%NodCoord(i, 1)*ratio
         end do
      elseif (sp == 1) then  ! PF_GLYCINE_MAX
         sw = 0.50d0
         if (present(SoyWidthRatio)) sw = SoyWidthRatio
         if (sw >= 1.0d0 .or. sw <= 0.0d0) stop "ERROR  >> mesh%create(leaf3d, PF_SOYBEAN) >> invalid SoyWidthRatio"
         length = maxval(mesh%NodCoord(:, 3)) - minval(mesh%NodCoord(:, 3))
         width = maxval(mesh%NodCoord(:, 1)) - minval(mesh%NodCoord(:, 1))
         do i = 1, size(mesh%NodCoord, 1)
            xmax = maxval(mesh%NodCoord(:, 1))
            xmin = minval(mesh%NodCoord(:, 1))
            xx = mesh%NodCoord(i, 3)
            if (mesh%NodCoord(i, 1) <= (xmax + xmin)*0.50d0) then
               alpha = sw*width
            else
               alpha = (1.0d0 - sw)*width
            end if
            r = (alpha*alpha + (length - alpha)*(length - alpha))/(2*alpha)*1.20d0
            if (xx <= 1.0d0/25.0d0*length) then
               mesh%NodCoord(i, 1) = mesh%NodCoord(i, 1)*1.0d0/10.0d0
               mesh%NodCoord(i, 2) = mesh%NodCoord(i, 2)*5.0d0
               cycle
            elseif (xx < alpha) then
               yy = sqrt(alpha*alpha - (xx - alpha)*(xx - alpha))
               yy_ = xx
            else
               yy_ = alpha + (-alpha)/(length - alpha)*(xx - alpha)
               yy = alpha - r + sqrt(r*r - (xx - alpha)*(xx - alpha))
            end if
            yy = abs(lin_curve_ratio*yy + (1.0d0 - lin_curve_ratio)*yy_)
            mesh%NodCoord(i, 1) = mesh%NodCoord(i, 1)*(yy/alpha)
         end do
         do i = 1, size(mesh%NodCoord, 1)
            theta = abs(mesh%NodCoord(i, 1)/(width/2.0d0))
            mesh%NodCoord(i, 2) = mesh%NodCoord(i, 2) - 0.10d0*theta*(theta - 1.0d0)
         end do
      else
         stop "[ERROR] Mesh%create => No such species"
      end if
   end subroutine

   ! Cylinder (MeshClass.f90:6506-6522): Circle2D (:6524-6743) on a unit disc, cleanMesh,
   ! Convert2Dto3D(thickness=lz, division=zn), resized to lx, ly
   subroutine cylinder3d_mesh(mesh, x_num, y_num, zn, lx, ly, lz)
      type(Mesh_), intent(inout) :: mesh
      integer(int32), intent(in) :: x_num, y_num, zn
      real(real64), intent(in) :: lx, ly, lz
      real(real64), allocatable :: c2(:, :), c3(:, :)
      integer(int32), allocatable :: q2(:, :), used(:), newid(:)
      real(real64) :: xx, yy, RR, theta, alpha, rr_, ang, cf
      real(real64), parameter :: pi = 3.14159265358979323846d0
      integer(int32) :: xn, m, sq, nsq, nring, i, j, k, n, ini, nn2, ne2, e, a
      xn = max(x_num/2 + 1, y_num/2 + 1)
      ! the ring layers index nodes up to nsq + 8*xn*(xn+2), which exist only from xn = 2 on
      if (xn < 2) stop "plantfem_gpu: Cylinder needs x_num or y_num >= 2"
      m = 2*xn
      nsq = (m + 1)*(m + 1)
      nring = m*4*m
      sq = nsq
      allocate (c2(nsq + nring, 2), q2(m*m + 8*xn*(xn + 2), 4))
      ! rectangular2D (2m x 2m, side 2) shifted to [-1, 1]^2, then rounded
      n = 0
      do j = 1, m + 1
         do i = 1, m + 1
            n = n + 1
            c2(n, 1) = 2.0d0/dble(m)*dble(i - 1) - 1.0d0
            c2(n, 2) = 2.0d0/dble(m)*dble(j - 1) - 1.0d0
         end do
      end do
      e = 0
      do j = 1, m
         do i = 1, m
            e = e + 1
            q2(e, :) = [i + (j - 1)*(m + 1), i + 1 + (j - 1)*(m + 1), m + 2 + i + (j - 1)*(m + 1), &
                        m + 1 + i + (j - 1)*(m + 1)]
         end do
      end do
      theta = (1.0d0 - 0.80d0)/(1.0d0 - 1.0d0/sqrt(2.0d0))
      do i = 1, nsq
         xx = c2(i, 1)
         yy = c2(i, 2)
         RR = 1.00d0
         if (xx /= 0.0d0) then
            if (abs(yy/xx) > 1.0d0) then
               if (yy /= 0.0d0) RR = sqrt((xx/yy)*(xx/yy) + 1)
            else
               RR = sqrt((yy/xx)*(yy/xx) + 1)
            end if
         end if
         alpha = theta*((1/RR) - 1.0d0/sqrt(2.0d0)) + 0.790d0
         c2(i, 1) = (xx*alpha)*(sqrt(2.0d0))*0.95d0
         c2(i, 2) = (yy*alpha)*(sqrt(2.0d0))*0.95d0
      end do
      cf = dble(2*xn - 1)/dble(2*xn)
      c2(1:nsq, 1) = cf*c2(1:nsq, 1)/sqrt(2.0d0)
      c2(1:nsq, 2) = cf*c2(1:nsq, 2)/sqrt(2.0d0)
      do i = 1, m
         rr_ = 1.0d0 + dble(i)*(1.0d0/dble(m))
         do j = 1, 4*m
            ang = 2.0d0*pi/4.0d0/dble(m)*dble(j - 1)
            c2(nsq + (i - 1)*4*m + j, 1) = rr_*cos(ang)
            c2(nsq + (i - 1)*4*m + j, 2) = rr_*sin(ang)
         end do
      end do
      ! the square's boundary tied to ring 1, then ring layers 1..xn+1
      j = 0
      do i = 1, xn
         j = j + 1
         q2(m*m + j, :) = [(m + 1)*(xn + i), sq + j, sq + j + 1, (m + 1)*(xn + i + 1)]
      end do
      do i = 1, m
         j = j + 1
         q2(m*m + j, :) = [sq - i + 1, sq + j, sq + j + 1, sq - i]
      end do
      do i = 1, m
         j = j + 1
         q2(m*m + j, :) = [sq - (m + 1) + 1 - (i - 1)*(m + 1), sq + j, sq + j + 1, sq - (m + 1) + 1 - i*(m + 1)]
      end do
      do i = 1, m
         j = j + 1
         q2(m*m + j, :) = [i, sq + j, sq + j + 1, i + 1]
      end do
      do i = 1, xn
         j = j + 1
         q2(m*m + j, :) = [(m + 1)*i, sq + j, sq + j + 1, (m + 1)*(i + 1)]
      end do
      q2(m*m + j, 3) = sq + 1
      do i = 1, xn + 1
         ini = j + 1
         do k = 1, 8*xn - 1
            j = j + 1
            q2(m*m + j, :) = [sq + j - 8*xn, sq + j, sq + j + 1, sq + j + 1 - 8*xn]
         end do
         j = j + 1
         q2(m*m + j, :) = [sq + j - 8*xn, sq + j, sq + ini, sq + ini - 8*xn]
      end do
      ! Circle2D's resize to a unit disc (over every node, the unused outer rings included)
      c2(:, 1) = 1.0d0/(maxval(c2(:, 1)) - minval(c2(:, 1)))*c2(:, 1)
      c2(:, 2) = 1.0d0/(maxval(c2(:, 2)) - minval(c2(:, 2)))*c2(:, 2)
      ! cleanMesh: drop the unused nodes
      allocate (used(size(c2, 1)), newid(size(c2, 1)))
      used = 0
      do e = 1, size(q2, 1)
         do a = 1, 4
            used(q2(e, a)) = 1
         end do
      end do
      nn2 = 0
      do i = 1, size(c2, 1)
         if (used(i) == 1) nn2 = nn2 + 1
         newid(i) = nn2
      end do
      ne2 = size(q2, 1)
      ! Convert2Dto3D and the final resize
      if (allocated(mesh%NodCoord)) deallocate (mesh%NodCoord)
      if (allocated(mesh%ElemNod)) deallocate (mesh%ElemNod)
      allocate (c3(nn2, 2), mesh%NodCoord(nn2*(zn + 1), 3), mesh%ElemNod(ne2*zn, 8))
      do i = 1, size(c2, 1)
         if (used(i) == 1) c3(newid(i), :) = c2(i, :)
      end do
      do k = 1, zn + 1
         mesh%NodCoord(nn2*(k - 1) + 1:nn2*k, 1:2) = c3
         mesh%NodCoord(nn2*(k - 1) + 1:nn2*k, 3) = lz/dble(zn)*dble(k - 1)
      end do
      do k = 1, zn
         do e = 1, ne2
            do a = 1, 4
               mesh%ElemNod(ne2*(k - 1) + e, a) = newid(q2(e, a)) + nn2*(k - 1)
               mesh%ElemNod(ne2*(k - 1) + e, a + 4) = newid(q2(e, a)) + nn2*k
            end do
         end do
      end do
      mesh%NodCoord(:, 1) = lx/(maxval(mesh%NodCoord(:, 1)) - minval(mesh%NodCoord(:, 1)))*mesh%NodCoord(:, 1)
      mesh%NodCoord(:, 2) = ly/(maxval(mesh%NodCoord(:, 2)) - minval(mesh%NodCoord(:, 2)))*mesh%NodCoord(:, 2)
   end subroutine

   integer(int32) function nnFEMDomain(this)
      class(FEMDomain_), intent(in) :: this
      nnFEMDomain = size(this%mesh%NodCoord, 1)
   end function

   integer(int32) function neFEMDomain(this)
      class(FEMDomain_), intent(in) :: this
      neFEMDomain = size(this%mesh%ElemNod, 1)
   end function

   integer(int32) function nneFEMDomain(this)
      class(FEMDomain_), intent(in) :: this
      nneFEMDomain = size(this%mesh%ElemNod, 2)
   end function

   integer(int32) function ndFEMDomain(this)
      class(FEMDomain_), intent(in) :: this
      ndFEMDomain = size(this%mesh%NodCoord, 2)
   end function

   real(real64) function x_minFEMDomain(this)
      class(FEMDomain_), intent(in) :: this
      x_minFEMDomain = minval(this%mesh%NodCoord(:, 1))
   end function

   real(real64) function x_maxFEMDomain(this)
      class(FEMDomain_), intent(in) :: this
      x_maxFEMDomain = maxval(this%mesh%NodCoord(:, 1))
   end function

   real(real64) function z_minFEMDomain(this)
      class(FEMDomain_), intent(in) :: this
      z_minFEMDomain = minval(this%mesh%NodCoord(:, 3))
   end function

   real(real64) function z_maxFEMDomain(this)
      class(FEMDomain_), intent(in) :: this
      z_maxFEMDomain = maxval(this%mesh%NodCoord(:, 3))
   end function

   ! selectFEMDomain box form with InOrOut's closed box (FEMDomainClass.f90:12854-12926)
   function selectFEMDomain(this, x_min, x_max, y_min, y_max, z_min, z_max) result(NodeList)
      class(FEMDomain_), intent(in) :: this
      real(real64), optional, intent(in) :: x_min, x_max, y_min, y_max, z_min, z_max
      integer(int32), allocatable :: NodeList(:)
      real(real64) :: lo(3), hi(3)
      logical, allocatable :: inside(:)
      integer(int32) :: i, d, m
      lo = -1.0d14; hi = 1.0d14
      if (present(x_min)) lo(1) = x_min
      if (present(y_min)) lo(2) = y_min
      if (present(z_min)) lo(3) = z_min
      if (present(x_max)) hi(1) = x_max
      if (present(y_max)) hi(2) = y_max
      if (present(z_max)) hi(3) = z_max
      allocate (inside(this%nn()))
      inside = .true.
      do d = 1, 3
         inside = inside .and. (lo(d) <= this%mesh%NodCoord(:, d)) .and. (this%mesh%NodCoord(:, d) <= hi(d))
      end do
      allocate (NodeList(count(inside)))
      m = 0
      do i = 1, this%nn()
         if (inside(i)) then
            m = m + 1
            NodeList(m) = i
         end if
      end do
   end function

   function ZeroMatrixFEMDomain(this, DOF) result(ZeroMatrix)
      class(FEMDomain_), intent(in) :: this
      integer(int32), intent(in) :: DOF
      type(CRS_) :: ZeroMatrix
      integer(c_int64_t) :: nnz
      integer(c_int) :: code
      integer(int64), allocatable, target :: rp(:)
      integer(int32), allocatable, target :: ci(:)
      code = c_pf_crs_graph(ctx(), int(this%nn(), c_int64_t), int(this%ne(), c_int64_t), &
                            int(this%nne(), c_int32_t), this%mesh%ElemNod, int(DOF, c_int32_t), nnz, &
                            c_null_ptr, c_null_ptr)
      call check(code, "ZeroMatrix")
      allocate (rp(int(DOF, int64)*this%nn() + 1), ci(nnz))
      code = c_pf_crs_graph(ctx(), int(this%nn(), c_int64_t), int(this%ne(), c_int64_t), &
                            int(this%nne(), c_int32_t), this%mesh%ElemNod, int(DOF, c_int32_t), nnz, &
                            c_loc(rp), c_loc(ci))
      call check(code, "ZeroMatrix")
      call move_alloc(rp, ZeroMatrix%row_ptr)
      call move_alloc(ci, ZeroMatrix%col_idx)
      allocate (ZeroMatrix%val(nnz))
      ZeroMatrix%val(:) = 0.0d0
   end function

   function StiffnessMatrixElementFEMDomain(this, ElementID, E, v) result(StiffnessMatrix)
      class(FEMDomain_), intent(in) :: this
      integer(int32), intent(in) :: ElementID
      real(real64), intent(in) :: E, v
      real(real64), allocatable :: StiffnessMatrix(:, :)
      real(real64), allocatable :: Ea(:), va(:)
      integer(c_int) :: code
      allocate (StiffnessMatrix(24, 24), Ea(this%ne()), va(this%ne()))
      Ea = 0.0d0; va = 0.0d0
      Ea(ElementID) = E
      va(ElementID) = v
      code = c_pf_hex8_stiffness(ctx(), int(this%nn(), c_int64_t), int(this%ne(), c_int64_t), this%mesh%NodCoord, &
                                 this%mesh%ElemNod, Ea, va, 0_c_int32_t, int(ElementID, c_int64_t), 1_c_int64_t, &
                                 StiffnessMatrix)
      call check(code, "StiffnessMatrix(ElementID,E,v)")
   end function

   ! MassMatrixFEMDomain(ElementID, Density, DOF) (FEMDomainClass.f90:10874-10945), DOF = 3
   function MassMatrixFEMDomain(this, ElementID, Density, DOF) result(MassMatrix)
      class(FEMDomain_), intent(in) :: this
      integer(int32), intent(in) :: ElementID
      real(real64), optional, intent(in) :: Density
      integer(int32), optional, intent(in) :: DOF
      real(real64), allocatable :: MassMatrix(:, :)
      real(real64) :: rho(1)
      integer(c_int) :: code
      if (present(DOF)) then
         if (DOF /= 3) then
            print *, "plantfem_gpu :: MassMatrix :: DOF=3 only"
            error stop
         end if
      end if
      rho(1) = 1.0d0
      if (present(Density)) rho(1) = Density
      allocate (MassMatrix(24, 24))
      code = c_pf_hex8_mass_matrix(ctx(), int(this%nn(), c_int64_t), int(this%ne(), c_int64_t), this%mesh%NodCoord, &
                                   this%mesh%ElemNod, rho, 1_c_int64_t, int(ElementID, c_int64_t), 1_c_int64_t, &
                                   MassMatrix)
      call check(code, "MassMatrix(ElementID,Density,DOF)")
   end function

   function StiffnessMatrixCRSFEMDomain(this, YoungModulus, PoissonRatio) result(StiffnessMatrix)
      class(FEMDomain_), intent(in) :: this
      real(real64), intent(in) :: YoungModulus(:), PoissonRatio(:)
      type(CRS_) :: StiffnessMatrix
      integer(c_int) :: code
      StiffnessMatrix = this%ZeroMatrix(DOF=3)
      code = c_pf_hex8_stiffness_crs(ctx(), int(this%nn(), c_int64_t), int(this%ne(), c_int64_t), &
                                     this%mesh%NodCoord, this%mesh%ElemNod, YoungModulus, PoissonRatio, 0_c_int32_t, &
                                     StiffnessMatrix%row_ptr, StiffnessMatrix%col_idx, StiffnessMatrix%val)
      call check(code, "StiffnessMatrix(YoungModulus,PoissonRatio)")
   end function

   subroutine moveFEMDomain(this, x, y, z)
      class(FEMDomain_), intent(inout) :: this
      real(real64), optional, intent(in) :: x, y, z
      if (present(x)) this%mesh%NodCoord(:, 1) = this%mesh%NodCoord(:, 1) + x
      if (present(y)) this%mesh%NodCoord(:, 2) = this%mesh%NodCoord(:, 2) + y
      if (present(z)) this%mesh%NodCoord(:, 3) = this%mesh%NodCoord(:, 3) + z
   end subroutine

   ! overset(FEMDomain, DomainID, algorithm, MyDomainID) (FEMDomainClass.f90:14606-14700): the
   ! Gauss points of this mesh inside elements of FEMDomain become connects, searched on the GPU
   subroutine oversetFEMDomain(this, FEMDomain, DomainID, algorithm, MyDomainID)
      class(FEMDomain_), intent(inout) :: this
      type(FEMDomain_), intent(in) :: FEMDomain
      integer(int32), intent(in) :: DomainID, algorithm
      integer(int32), optional, intent(in) :: MyDomainID
      integer(c_int64_t) :: cap, k
      integer(int32), allocatable :: e(:), g(:), p(:), ebuf(:), gbuf(:), pbuf(:), dbuf(:, :)
      real(real64), allocatable :: x(:), xbuf(:, :)
      integer(int32) :: my, n0
      integer(c_int) :: code
      if (algorithm /= FEMDomain_Overset_GPP .and. algorithm /= FEMDomain_Overset_P2P) &
         stop "plantfem_gpu: overset algorithm is FEMDomain_Overset_GPP or FEMDomain_Overset_P2P"
      if (this%overset_kind /= 0 .and. this%overset_kind /= algorithm) &
         stop "[ERROR] FEMDomainClass > oversetFEMDomain >> P2P & GPP cannot be mixed."
      this%overset_kind = algorithm
      my = 1
      if (present(MyDomainID)) my = MyDomainID
      if (algorithm == FEMDomain_Overset_P2P) then
         cap = int(this%nn(), c_int64_t)
      else
         cap = 8_c_int64_t*this%ne()
      end if
      allocate (e(max(cap, 1_c_int64_t)), g(max(cap, 1_c_int64_t)), p(max(cap, 1_c_int64_t)), &
                x(3*max(cap, 1_c_int64_t)))
      if (algorithm == FEMDomain_Overset_P2P) then  ! own node ids, Gauss point 0
         code = c_pf_overset_p2p(ctx(), int(this%nn(), c_int64_t), int(this%ne(), c_int64_t), this%mesh%NodCoord, &
                                 this%mesh%ElemNod, int(FEMDomain%nn(), c_int64_t), int(FEMDomain%ne(), c_int64_t), &
                                 FEMDomain%mesh%NodCoord, FEMDomain%mesh%ElemNod, cap, k, e, p, x)
         g = 0
      else
         code = c_pf_overset_gpp(ctx(), int(this%nn(), c_int64_t), int(this%ne(), c_int64_t), this%mesh%NodCoord, &
                                 this%mesh%ElemNod, int(FEMDomain%nn(), c_int64_t), int(FEMDomain%ne(), c_int64_t), &
                                 FEMDomain%mesh%NodCoord, FEMDomain%mesh%ElemNod, cap, k, e, g, p, x)
      end if
      call check(code, "overset")
      n0 = this%num_oversetconnect
      allocate (ebuf(n0 + k), gbuf(n0 + k), pbuf(n0 + k), dbuf(2, n0 + k), xbuf(3, n0 + k))
      if (n0 > 0) then
         ebuf(1:n0) = this%ov_elem
         gbuf(1:n0) = this%ov_gp
         pbuf(1:n0) = this%ov_pair
         dbuf(:, 1:n0) = this%ov_domains
         xbuf(:, 1:n0) = this%ov_position
      end if
      ebuf(n0 + 1:) = e(1:k)
      gbuf(n0 + 1:) = g(1:k)
      pbuf(n0 + 1:) = p(1:k)
      dbuf(1, n0 + 1:) = my
      dbuf(2, n0 + 1:) = DomainID
      xbuf(:, n0 + 1:) = reshape(x(1:3*k), [3, int(k)])
      call move_alloc(ebuf, this%ov_elem)
      call move_alloc(gbuf, this%ov_gp)
      call move_alloc(pbuf, this%ov_pair)
      call move_alloc(dbuf, this%ov_domains)
      call move_alloc(xbuf, this%ov_position)
      this%num_oversetconnect = n0 + int(k)
   end subroutine

   integer(int32) function numOversetElementsFEMDomain(this)
      class(FEMDomain_), intent(in) :: this
      numOversetElementsFEMDomain = this%num_oversetconnect
   end function

   ! ------------------------------------------------------------------------ FEMSolver_
   ! init(NumDomain) (FEMSolverClass.f90:167-210)
   subroutine initFEMSolver(this, NumDomain)
      class(FEMSolver_), intent(inout) :: this
      integer(int32), optional, intent(in) :: NumDomain
      integer(c_int) :: code
      integer(int32) :: nd
      nd = 1
      if (present(NumDomain)) nd = NumDomain
      if (nd < 1) stop "plantfem_gpu: init(NumDomain >= 1)"
      if (allocated(this%FEMDomains)) deallocate (this%FEMDomains)
      allocate (this%FEMDomains(nd))
      if (.not. c_associated(this%sys)) then
         code = c_pf_system_create(ctx(), this%sys)
         call check(code, "init")
      end if
      this%initialized = .true.
   end subroutine

   ! setDomain(FEMDomain, DomainID) (FEMSolverClass.f90:329-419)
   subroutine setDomainFEMSolver(this, FEMDomain, DomainID)
      class(FEMSolver_), intent(inout) :: this
      type(FEMDomain_), target, intent(in) :: FEMDomain
      integer(int32), optional, intent(in) :: DomainID
      integer(c_int) :: code
      integer(int32) :: id
      if (.not. this%initialized) call this%init(NumDomain=1)
      id = 1
      if (present(DomainID)) id = DomainID
      if (id < 1 .or. id > size(this%FEMDomains)) stop "ERROR :: setFEMDomains >> invalid domain IDs"
      this%FEMDomains(id)%femdomainp => FEMDomain
      if (size(this%FEMDomains) >= 2) return  ! joined at setCRS
      this%domain => FEMDomain
      code = c_pf_system_set_mesh(this%sys, int(FEMDomain%nn(), c_int64_t), int(FEMDomain%ne(), c_int64_t), &
                                  int(FEMDomain%nne(), c_int32_t), FEMDomain%mesh%NodCoord, FEMDomain%mesh%ElemNod)
      call check(code, "setDomain")
   end subroutine

   ! multi-domain setCRS (FEMSolverClass.f90:466-676): the domains' meshes concatenated, the
   ! ZeroMatrix blocks merged with the node pairs of every overset connect, columns ascending
   subroutine setCRSMultiFEMSolver(this, DOF)
      class(FEMSolver_), intent(inout) :: this
      integer(int32), intent(in) :: DOF
      integer(int32) :: nd, d, nn, ne, k
      integer(int32), allocatable :: own(:), pair(:), gp(:)
      integer(c_int64_t) :: nnz
      integer(c_int32_t) :: any_overset
      integer(c_int) :: code
      type(FEMDomain_), pointer :: f
      nd = size(this%FEMDomains)
      if (allocated(this%node_off)) deallocate (this%node_off)
      allocate (this%node_off(nd + 1))
      if (allocated(this%elem_off)) deallocate (this%elem_off)
      allocate (this%elem_off(nd + 1))
      this%node_off(1) = 0
      this%elem_off(1) = 0
      do d = 1, nd
         if (.not. associated(this%FEMDomains(d)%femdomainp)) stop "plantfem_gpu: setCRS with an empty domain slot"
         this%node_off(d + 1) = this%node_off(d) + this%FEMDomains(d)%femdomainp%nn()
         this%elem_off(d + 1) = this%elem_off(d) + this%FEMDomains(d)%femdomainp%ne()
      end do
      if (associated(this%joint)) deallocate (this%joint)
      allocate (this%joint)
      nn = this%node_off(nd + 1)
      ne = this%elem_off(nd + 1)
      allocate (this%joint%mesh%NodCoord(nn, 3), this%joint%mesh%ElemNod(ne, 8))
      any_overset = 0
      k = 0
      do d = 1, nd
         f => this%FEMDomains(d)%femdomainp
         this%joint%mesh%NodCoord(this%node_off(d) + 1:this%node_off(d + 1), :) = f%mesh%NodCoord(:, 1:3)
         this%joint%mesh%ElemNod(this%elem_off(d) + 1:this%elem_off(d + 1), :) = f%mesh%ElemNod(:, 1:8) + &
            this%node_off(d)
         if (allocated(f%ov_elem)) any_overset = 1
         k = k + f%num_oversetconnect
      end do
      allocate (own(max(k, 1)), pair(max(k, 1)), gp(max(k, 1)))
      call connects_of(this, own, pair, gp)
      this%domain => this%joint
      code = c_pf_system_set_mesh(this%sys, int(nn, c_int64_t), int(ne, c_int64_t), 8_c_int32_t, &
                                  this%joint%mesh%NodCoord, this%joint%mesh%ElemNod)
      call check(code, "setCRS (joint mesh)")
      code = c_pf_system_multi_crs(this%sys, int(DOF, c_int32_t), any_overset, int(k, c_int64_t), own, gp, pair, nnz)
      call check(code, "setCRS")
      if (allocated(this%CRS_ID_Starts_From)) deallocate (this%CRS_ID_Starts_From)
      allocate (this%CRS_ID_Starts_From(nd))
      do d = 1, nd
         this%CRS_ID_Starts_From(d) = 1_int64 + int(DOF, int64)*this%node_off(d)
      end do
      this%DOF = DOF
      this%n = int(DOF, int64)*nn
      this%nnz = nnz
      this%npend = 0
   end subroutine

   ! every connect of every domain in domain order, as element ids of the joint mesh
   subroutine connects_of(this, own, pair, gp, pos)
      class(FEMSolver_), intent(in) :: this
      integer(int32), intent(out) :: own(:), pair(:)
      integer(int32), optional, intent(out) :: gp(:)
      real(real64), optional, intent(out) :: pos(:, :)
      integer(int32) :: d, i, k
      type(FEMDomain_), pointer :: f
      k = 0
      do d = 1, size(this%FEMDomains)
         f => this%FEMDomains(d)%femdomainp
         do i = 1, f%num_oversetconnect
            k = k + 1
            if (f%ov_gp(i) == 0) then  ! P2P: own node of the joint mesh
               own(k) = f%ov_elem(i) + this%node_off(f%ov_domains(1, i))
            else
               own(k) = f%ov_elem(i) + this%elem_off(f%ov_domains(1, i))
            end if
            pair(k) = f%ov_pair(i) + this%elem_off(f%ov_domains(2, i))
            if (present(gp)) gp(k) = f%ov_gp(i)
            if (present(pos)) pos(:, k) = f%ov_position(:, i)
         end do
      end do
   end subroutine

   ! setEbOM(penalty, DOF) (FEMSolverClass.f90:1891-1985)
   subroutine setEbOMFEMSolver(this, penalty, DOF)
      class(FEMSolver_), intent(inout) :: this
      real(real64), intent(in) :: penalty
      integer(int32), intent(in) :: DOF
      integer(int32) :: d, k
      integer(int32), allocatable :: own(:), pair(:), gp(:)
      real(real64), allocatable :: pos(:, :)
      integer(c_int) :: code
      call this%flush()
      if (size(this%FEMDomains) < 2) return
      k = 0
      do d = 1, size(this%FEMDomains)
         k = k + this%FEMDomains(d)%femdomainp%num_oversetconnect
      end do
      allocate (own(max(k, 1)), pair(max(k, 1)), gp(max(k, 1)), pos(3, max(k, 1)))
      call connects_of(this, own, pair, gp, pos)
      code = c_pf_system_ebom(this%sys, int(k, c_int64_t), own, gp, pair, pos, penalty, int(DOF, c_int32_t))
      call check(code, "setEbOM")
   end subroutine

   ! setCRS(DOF) single-domain branch (FEMSolverClass.f90:679-694)
   subroutine setCRSFEMSolver(this, DOF)
      class(FEMSolver_), intent(inout) :: this
      integer(int32), intent(in) :: DOF
      integer(c_int64_t) :: nnz
      integer(c_int) :: code
      if (allocated(this%FEMDomains)) then
         if (size(this%FEMDomains) >= 2) then
            call setCRSMultiFEMSolver(this, DOF)
            return
         end if
      end if
      code = c_pf_system_zero_matrix(this%sys, int(DOF, c_int32_t), nnz)
      call check(code, "setCRS")
      this%DOF = DOF
      this%n = int(DOF, int64)*this%domain%nn()
      this%nnz = nnz
      this%npend = 0
   end subroutine

   ! setCRS(CRS_) (FEMSolverClass.f90:421-429)
   subroutine setCRS_CRSobjFEMSolver(this, CRS)
      class(FEMSolver_), intent(inout) :: this
      type(CRS_), intent(in) :: CRS
      integer(c_int) :: code
      if (.not. this%initialized) call this%init(NumDomain=1)
      code = c_pf_system_set_crs(this%sys, int(size(CRS%row_ptr) - 1, c_int64_t), int(size(CRS%col_idx), c_int64_t), &
                                 CRS%row_ptr, CRS%col_idx, CRS%val)
      call check(code, "setCRS(CRS)")
      this%n = size(CRS%row_ptr) - 1
      this%nnz = size(CRS%col_idx)
   end subroutine

   subroutine setRHSFEMSolver(this, RHS)
      class(FEMSolver_), intent(inout) :: this
      real(real64), intent(in) :: RHS(:)
      integer(c_int) :: code
      code = c_pf_system_set_rhs(this%sys, RHS)
      call check(code, "setRHS")
   end subroutine

   subroutine flushFEMSolver(this)
      class(FEMSolver_), intent(inout) :: this
      integer(c_int) :: code
      if (this%npend == 0) return
      code = c_pf_system_add_element_matrices(this%sys, int(this%npend, c_int64_t), this%pend_ids, this%pend_mat)
      call check(code, "setMatrix (fill-in request)")
      this%npend = 0
   end subroutine

   ! setMatrix(DomainID,ElementID,DOF,Matrix) (FEMSolverClass.f90:901-909, :1020-1057)
   subroutine setMatrixFEMSolver(this, DomainID, ElementID, DOF, Matrix)
      class(FEMSolver_), intent(inout) :: this
      integer(int32), intent(in) :: DomainID, ElementID, DOF
      real(real64), intent(in) :: Matrix(:, :)
      integer(int64), parameter :: batch = 8192
      integer(int64) :: m
      if (this%n == 0) call this%setCRS(DOF=DOF)
      if (DomainID < 1 .or. DomainID > size(this%FEMDomains)) stop "plantfem_gpu: setMatrix with an invalid DomainID"
      m = size(Matrix, kind=int64)
      if (.not. allocated(this%pend_ids) .or. this%mat_len /= m) then
         call this%flush()
         if (allocated(this%pend_ids)) deallocate (this%pend_ids, this%pend_mat)
         allocate (this%pend_ids(batch), this%pend_mat(batch*m))
         this%mat_len = m
      end if
      this%npend = this%npend + 1
      this%pend_ids(this%npend) = ElementID
      if (allocated(this%elem_off)) this%pend_ids(this%npend) = ElementID + this%elem_off(DomainID)
      this%pend_mat((this%npend - 1)*m + 1:this%npend*m) = reshape(Matrix, [m])
      if (this%npend == batch) call this%flush()
   end subroutine

   subroutine setVectorFEMSolver(this, DomainID, ElementID, DOF, Vector)
      class(FEMSolver_), intent(inout) :: this
      integer(int32), intent(in) :: DomainID, ElementID, DOF
      real(real64), intent(in) :: Vector(:)
      integer(c_int) :: code
      integer(int32) :: ids(1)
      if (DomainID < 1 .or. DomainID > size(this%FEMDomains)) stop "plantfem_gpu: setVector with an invalid DomainID"
      call this%flush()
      ids(1) = ElementID
      if (allocated(this%elem_off)) ids(1) = ElementID + this%elem_off(DomainID)
      code = c_pf_system_add_element_vectors(this%sys, 1_c_int64_t, ids, Vector)
      call check(code, "setVector")
   end subroutine

   ! device assembly of StiffnessMatrix(YoungModulus(:),PoissonRatio(:)) into the solver's CRS
   subroutine setStiffnessFEMSolver(this, YoungModulus, PoissonRatio)
      class(FEMSolver_), intent(inout) :: this
      real(real64), intent(in) :: YoungModulus(:), PoissonRatio(:)
      integer(c_int) :: code
      call this%flush()
      code = c_pf_system_stiffness(this%sys, YoungModulus, PoissonRatio, 0_c_int32_t)
      call check(code, "setStiffness")
      this%DOF = 3
      this%n = 3_int64*this%domain%nn()
   end subroutine

   ! TractionVector_by_elem summed over elements (open box as Range_), added to the RHS
   subroutine addTractionFEMSolver(this, box, SurfacePressure)
      class(FEMSolver_), intent(inout) :: this
      real(real64), intent(in) :: box(6), SurfacePressure
      integer(c_int) :: code
      code = c_pf_system_add_traction(this%sys, box, SurfacePressure)
      call check(code, "addTraction")
   end subroutine

   ! fix(DomainID,IDs,FixValue|FixValues) (FEMSolverClass.f90:1073-1136)
   subroutine fixFEMSolver(this, DomainID, IDs, FixValue, FixValues)
      class(FEMSolver_), intent(inout) :: this
      integer(int32), optional, intent(in) :: DomainID
      integer(int32), intent(in) :: IDs(:)
      real(real64), optional, intent(in) :: FixValue
      real(real64), optional, target, intent(in) :: FixValues(:)
      integer(c_int) :: code
      real(real64) :: fv
      type(c_ptr) :: pv
      if (present(DomainID)) then
         ! the reference's DomainID >= 2 form marks IDs but stores values at IDs + earlier nodes
         if (DomainID /= 1) stop "plantfem_gpu: fix(DomainID>=2) is not mirrored; pass DOF ids of the whole system"
      end if
      fv = 0.0d0
      if (present(FixValue)) fv = FixValue
      pv = c_null_ptr
      if (present(FixValues)) pv = c_loc(FixValues)
      code = c_pf_system_fix(this%sys, int(size(IDs), c_int64_t), IDs, pv, fv)
      call check(code, "fix")
   end subroutine

   ! solve(algorithm, preconditioning, x0) (FEMSolverClass.f90:1697-1825): elimination, then
   ! bicgstab_CRS_2, or with preconditioning="PointJacobi" the Jacobi-scaled system (the
   ! reference calls no solver after the scaling; "PointJacobi-noSolve" does the same), or with
   ! "ILU"/"incompleteLU" bicgstab_CRS_ILU; x_fixed = ubar.  Without preconditioning,
   ! algorithm="GaussJordan" (FEMSolverClass.f90:1795-1798, whose gauss_jordan_crs stops "Under
   ! construction") is a direct solve: banded Cholesky after a bandwidth-reducing renumbering,
   ! one step of FP64 refinement; a non-positive pivot stops " A is singular " as the
   ! reference's pivot test does (:2455).  Other algorithm values run bicgstab_CRS_2.
   function solveFEMSolver(this, algorithm, preconditioning, x0, precision) result(x)
      class(FEMSolver_), intent(inout) :: this
      character(*), optional, intent(in) :: algorithm, preconditioning
      real(real64), optional, target, intent(in) :: x0(:)
      integer(int32), optional, intent(in) :: precision
      real(real64), allocatable :: x(:)
      real(real64), allocatable, target :: xl(:)
      type(pf_solve_opts) :: o
      type(c_ptr) :: px0
      integer(c_int) :: code
      call this%flush()
      o = default_opts()
      o%itrmax = this%itrmax
      o%er0 = this%er0
      o%relative_er = this%relative_er
      o%debug = merge(1, 0, this%debug)
      if (present(precision)) o%precision = precision
      if (.not. present(preconditioning) .and. present(algorithm)) then
         if (algorithm == "GaussJordan") then
            allocate (xl(this%n))
            code = c_pf_system_solve_direct(this%sys, 1_c_int32_t, c_loc(xl), this%stats)
            if (code == PF_NOT_SPD) stop " A is singular "
            call check(code, "solve")
            this%CRS_x = xl
            call move_alloc(xl, x)
            return
         end if
      end if
      if (present(preconditioning)) then
         select case (preconditioning)
         case ("PointJacobi")
            o%preconditioning = 1
         case ("ILU", "incompleteLU")
            o%preconditioning = 2
         case ("PointJacobi-noSolve")
            o%preconditioning = 3
         case default
            print *, "[Warning!] :: FEMSolver :: invalid preconditioning"
            stop
         end select
      end if
      allocate (xl(this%n))
      px0 = c_null_ptr
      if (present(x0)) px0 = c_loc(x0)
      code = c_pf_system_solve(this%sys, o, px0, c_loc(xl), this%stats)
      if (code == PF_BREAKDOWN_VV .or. code == PF_BREAKDOWN_C2C3) stop "Bicgstab devide by zero"
      if (code /= PF_NOT_CONVERGED) call check(code, "solve")
      this%CRS_x = xl
      call move_alloc(xl, x)
   end function

   ! zerosFEMSolver (FEMSolverClass.f90:1310-1322)
   subroutine zerosFEMSolver(this)
      class(FEMSolver_), intent(inout) :: this
      call this%flush()
      call check(c_pf_system_zeros(this%sys), "zeros")
   end subroutine

   ! keepThisMatrixAsFEMSolver (FEMSolverClass.f90:1544-1604)
   subroutine keepThisMatrixAsFEMSolver(this, As)
      class(FEMSolver_), intent(inout) :: this
      character(*), intent(in) :: As
      call this%flush()
      call check(c_pf_system_keep_matrix(this%sys, trim(As)//c_null_char), "keepThisMatrixAs")
   end subroutine

   ! fix_eigFEMSolver (FEMSolverClass.f90:1607-1628)
   subroutine fix_eigFEMSolver(this, IDs)
      class(FEMSolver_), intent(inout) :: this
      integer(int32), intent(in) :: IDs(:)
      call check(c_pf_system_fix_eig(this%sys, int(size(IDs), c_int64_t), IDs), "fix_eig")
   end subroutine

   ! eigFEMSolver (FEMSolverClass.f90:1384-1541): every eigenpair of the kept "A", "B" with the
   ! fix_eig DOFs removed, or the num_eigen smallest (LOBPCG); eigen_vectors(n, count)
   subroutine eigFEMSolver(this, num_eigen, eigen_value, eigen_vectors)
      class(FEMSolver_), intent(inout) :: this
      integer(int32), optional, intent(in) :: num_eigen
      real(real64), allocatable, intent(inout) :: eigen_value(:)
      real(real64), allocatable, intent(inout) :: eigen_vectors(:, :)
      integer(c_int32_t) :: cnt, nev, got
      integer(c_int) :: code
      nev = 0
      if (present(num_eigen)) nev = num_eigen
      if (nev > 0) then
         cnt = nev
      else
         call check(c_pf_system_eig_count(this%sys, cnt), "eig")
      end if
      if (allocated(eigen_value)) deallocate (eigen_value)
      if (allocated(eigen_vectors)) deallocate (eigen_vectors)
      allocate (eigen_value(max(cnt, 1)), eigen_vectors(this%n, max(cnt, 1)))
      code = c_pf_system_eig(this%sys, nev, 1.0d-8, 10000_c_int32_t, eigen_value, eigen_vectors, got, this%stats)
      if (code /= PF_NOT_CONVERGED) call check(code, "eig")
      eigen_value = eigen_value(1:got)
      eigen_vectors = eigen_vectors(:, 1:got)
   end subroutine

   ! getCRS: host copies of the device CRS and RHS into the public CRS_* components
   subroutine getCRSFEMSolver(this)
      class(FEMSolver_), target, intent(inout) :: this
      integer(c_int) :: code
      call this%flush()
      if (allocated(this%CRS_Index_Row)) deallocate (this%CRS_Index_Row, this%CRS_Index_Col, this%CRS_val)
      allocate (this%CRS_Index_Row(this%n + 1), this%CRS_Index_Col(this%nnz), this%CRS_val(this%nnz))
      code = c_pf_system_get_crs(this%sys, c_loc(this%CRS_Index_Row), c_loc(this%CRS_Index_Col), c_loc(this%CRS_val))
      call check(code, "getCRS")
      if (allocated(this%CRS_RHS)) deallocate (this%CRS_RHS)
      allocate (this%CRS_RHS(this%n))
      code = c_pf_system_get_rhs(this%sys, this%CRS_RHS)
      call check(code, "getRHS")
   end subroutine

   ! saveMatrixFEMSolver (FEMSolverClass.f90:1225-1307), CRS form: name_data.txt (one value per
   ! line, list-directed), name_indices.txt and name_indptr.txt (one integer and a blank per line)
   subroutine saveMatrixFEMSolver(this, name)
      class(FEMSolver_), target, intent(inout) :: this
      character(*), intent(in) :: name
      integer :: u, i
      call this%getCRS()
      open (newunit=u, file=name//"_data.txt", status="replace", action="write")
      do i = 1, size(this%CRS_val)
         write (u, *) this%CRS_val(i)
      end do
      close (u)
      open (newunit=u, file=name//"_indices.txt", status="replace", action="write")
      do i = 1, size(this%CRS_Index_Col)
         write (u, '(I0,A)') this%CRS_Index_Col(i), " "
      end do
      close (u)
      open (newunit=u, file=name//"_indptr.txt", status="replace", action="write")
      do i = 1, size(this%CRS_Index_Row)
         write (u, '(I0,A)') int(this%CRS_Index_Row(i)), " "
      end do
      close (u)
   end subroutine

   subroutine removeFEMSolver(this)
      class(FEMSolver_), intent(inout) :: this
      integer(c_int) :: code
      if (c_associated(this%sys)) then
         code = c_pf_system_destroy(this%sys)
         this%sys = c_null_ptr
      end if
      this%initialized = .false.
      this%n = 0
      this%npend = 0
      ! several domains: the joined mesh and the offsets go with the system
      if (associated(this%joint)) deallocate (this%joint)
      if (allocated(this%FEMDomains)) deallocate (this%FEMDomains)
      if (allocated(this%elem_off)) deallocate (this%elem_off)
      if (allocated(this%node_off)) deallocate (this%node_off)
      if (allocated(this%CRS_ID_Starts_From)) deallocate (this%CRS_ID_Starts_From)
      this%domain => null()
   end subroutine

   ! ------------------------------------------------------------------- plug-in solvers
   ! LinearSolver(row_ptr,col_idx,val,rhs,x) of solveFEMSolver_UserDefinedLinearSolver
   ! (FEMSolverClass.f90:1638-1648): bicgstab_CRS_2 with the reference's defaults on the GPU.
   subroutine pf_gpu_bicgstab(row_ptr, col_idx, val, rhs, x)
      real(real64), intent(in) :: val(:), rhs(:)
      real(real64), intent(inout) :: x(:)
      integer(int32), intent(in) :: col_idx(:)
      integer(int64), intent(in) :: row_ptr(:)
      type(pf_solve_opts) :: o
      type(pf_stats) :: st
      integer(c_int) :: code
      o = default_opts()
      code = c_pf_bicgstab_crs(ctx(), int(size(row_ptr) - 1, c_int64_t), row_ptr, col_idx, val, rhs, x, o, st)
      if (code == PF_BREAKDOWN_VV .or. code == PF_BREAKDOWN_C2C3) stop "Bicgstab devide by zero"
      if (code /= PF_NOT_CONVERGED) call check(code, "pf_gpu_bicgstab")
   end subroutine

   ! function form (FEMSolverClass.f90:1674-1683)
   function pf_gpu_bicgstab_func(row_ptr, col_idx, val, rhs, x0) result(x)
      real(real64), intent(in) :: val(:), rhs(:)
      real(real64), intent(in) :: x0(:)
      real(real64), allocatable :: x(:)
      integer(int32), intent(in) :: col_idx(:)
      integer(int64), intent(in) :: row_ptr(:)
      x = x0
      call pf_gpu_bicgstab(row_ptr, col_idx, val, rhs, x)
   end function

   ! bicgstab_CRS(a, ptr_i, index_j, x, b, itrmax, er, debug) of LinearSolver_
   ! (LinearSolverClass.f90:1493-1569) on the GPU: the same arguments, x holds x0 on entry; a
   ! return before the loop when r0.r0 < er, the exit on r.r / r0.r0 < er (stop_mode 2)
   subroutine pf_gpu_bicgstab_crs(a, ptr_i, index_j, x, b, itrmax, er, debug)
      real(real64), intent(in) :: a(:), b(:)
      real(real64), intent(inout) :: x(:)
      integer(int32), intent(in) :: index_j(:), itrmax
      integer(int64), intent(in) :: ptr_i(:)
      real(real64), intent(in) :: er
      logical, optional, intent(in) :: debug
      type(pf_solve_opts) :: o
      type(pf_stats) :: st
      integer(c_int) :: code
      o = default_opts()
      o%stop_mode = 2
      o%er0 = er
      o%itrmax = itrmax
      if (present(debug)) then
         if (debug) o%debug = 1
      end if
      code = c_pf_bicgstab_crs(ctx(), int(size(ptr_i) - 1, c_int64_t), ptr_i, index_j, a, b, x, o, st)
      if (code == PF_BREAKDOWN_VV .or. code == PF_BREAKDOWN_C2C3) stop "Bicgstab devide by zero"
      if (code /= PF_NOT_CONVERGED) call check(code, "pf_gpu_bicgstab_crs")
   end subroutine

   ! sub_crs_matvec (LinearSolverClass.f90:3728-3776): new_vector = A old_vector
   subroutine pf_gpu_spmv(row_ptr, col_idx, val, old_vector, new_vector)
      real(real64), intent(in) :: val(:), old_vector(:)
      integer(int32), intent(in) :: col_idx(:)
      integer(int64), intent(in) :: row_ptr(:)
      real(real64), allocatable, intent(inout) :: new_vector(:)
      integer(c_int) :: code
      if (.not. allocated(new_vector)) allocate (new_vector(size(row_ptr) - 1))
      code = c_pf_spmv_crs(ctx(), int(size(row_ptr) - 1, c_int64_t), row_ptr, col_idx, val, old_vector, new_vector)
      call check(code, "pf_gpu_spmv")
   end subroutine

end module plantfem_gpu
