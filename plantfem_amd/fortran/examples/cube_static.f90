! cube_static -- static solve of a clamped Cube3D box under a top traction through the
! plantfem_gpu FEMSolver_ surface, written the way plantFEM's tutorials drive FEMSolver_
! (Tutorial/fem/simpleDeformationByFEMSolver.f90).
!
!   cube_static --version            library version (no device needed)
!   cube_static N [out.bin] [mode] [algorithm]
!                                    N^3 hex box; mode "elem" assembles with per-element
!                                    StiffnessMatrix + setMatrix, "crs" (default) with the
!                                    device assembly; algorithm "GaussJordan" solves directly
!                                    (default bicgstab_CRS_2); writes x (real64 stream) to out.bin
program cube_static
   use plantfem_gpu
   use iso_fortran_env, only: int32, int64, real64
   implicit none
   type(FEMDomain_), target :: cube
   type(FEMSolver_) :: solver
   integer(int32), allocatable :: FixBoundary(:)
   real(real64), allocatable :: x(:), E(:), v(:)
   real(real64) :: box(6), eps
   integer(int32) :: n, i, u
   character(len=256) :: arg, outfile, mode, algorithm

   call get_command_argument(1, arg)
   if (trim(arg) == "--version") then
      print '(a,i0)', "plantfem_gpu library version ", pf_gpu_version()
      stop
   end if
   n = 8
   if (len_trim(arg) > 0) read (arg, *) n
   outfile = ""
   call get_command_argument(2, outfile)
   mode = "crs"
   call get_command_argument(3, arg)
   if (len_trim(arg) > 0) mode = arg
   algorithm = "bicgstab"
   call get_command_argument(4, arg)
   if (len_trim(arg) > 0) algorithm = arg

   call cube%create("Cube3D", x_num=n, y_num=n, z_num=n)
   allocate (E(cube%ne()), v(cube%ne()))
   do i = 1, cube%ne()
      E(i) = 1.0d7*(1.0d0 + 0.5d0*sin(dble(i)))
      v(i) = 0.3d0
   end do

   call solver%init(NumDomain=1)
   call solver%setDomain(FEMDomain=cube, DomainID=1)
   call solver%setCRS(DOF=3)
   if (trim(mode) == "elem") then
      do i = 1, cube%ne()
         call solver%setMatrix(DomainID=1, ElementID=i, DOF=3, &
                               Matrix=cube%StiffnessMatrix(ElementID=i, E=E(i), v=v(i)))
      end do
   else
      call solver%setStiffness(E, v)
   end if

   ! uniform downward traction 1e4 Pa on the top face
   eps = 1.0d-9
   box = [-1.0d300, 1.0d300, -1.0d300, 1.0d300, cube%z_max() - eps, cube%z_max() + eps]
   call solver%addTraction(box, -1.0d4)

   ! clamp the bottom face
   FixBoundary = cube%select(z_max=cube%z_min())*3 - 2
   call solver%fix(DomainID=1, IDs=FixBoundary, FixValue=0.0d0)
   FixBoundary = cube%select(z_max=cube%z_min())*3 - 1
   call solver%fix(DomainID=1, IDs=FixBoundary, FixValue=0.0d0)
   FixBoundary = cube%select(z_max=cube%z_min())*3 - 0
   call solver%fix(DomainID=1, IDs=FixBoundary, FixValue=0.0d0)

   solver%relative_er = 1.0d-14
   solver%er0 = 1.0d-30
   x = solver%solve(algorithm=trim(algorithm))
   print '(a,i0,a,i0,a,es12.5,a,es12.5)', "n_dof=", size(x), " iterations=", solver%stats%iterations, &
      " relres=", solver%stats%true_relres, " min_uz=", minval(x(3::3))
   if (len_trim(outfile) > 0) then
      open (newunit=u, file=trim(outfile), access="stream", form="unformatted", status="replace")
      write (u) x
      close (u)
   end if
   call solver%remove()
end program
