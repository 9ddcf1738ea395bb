! plugin_solver -- the reference's external-solver hook with the GPU BiCGSTAB plugged in.
! solveFEMSolver_UserDefinedLinearSolver (FEMSolverClass.f90:1632-1662) calls
! LinearSolver(row_ptr, col_idx, val, rhs, x); pf_gpu_bicgstab has exactly that interface.
! The system is the 4x4 CRS of Tutorial/std/ex0004_linearsolvers.f90:19-36.  The same system
! also goes through pf_gpu_bicgstab_crs, LinearSolver_'s bicgstab_CRS argument list.
program plugin_solver
   use plantfem_gpu
   use iso_fortran_env, only: int32, int64, real64
   implicit none
   integer(int64) :: row_ptr(5)
   integer(int32) :: col_idx(10)
   real(real64) :: val(10), rhs(4), x(4), y(4), x2(4)
   real(real64), allocatable :: ax(:)
   row_ptr = [1_int64, 3_int64, 6_int64, 9_int64, 11_int64]
   col_idx = [1, 2, 1, 2, 3, 2, 3, 4, 3, 4]
   val = [2.0d0, -1.0d0, -1.0d0, 2.0d0, -1.0d0, 2.0d0, 3.0d0, 1.0d0, 1.0d0, 2.0d0]
   rhs = [1.0d0, 2.0d0, 3.0d0, 4.0d0]
   x = 0.0d0
   call pf_gpu_bicgstab(row_ptr, col_idx, val, rhs, x)
   call pf_gpu_spmv(row_ptr, col_idx, val, x, ax)
   y = ax
   print '(a,4es24.16)', "x=", x
   print '(a,es12.5)', "residual=", maxval(abs(y - rhs))
   x2 = 0.0d0
   call pf_gpu_bicgstab_crs(val, row_ptr, col_idx, x2, rhs, 1000, 1.0d-22)
   print '(a,4es24.16)', "x_crs=", x2
end program
