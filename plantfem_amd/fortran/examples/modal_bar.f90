! modal_bar -- modal analysis of a Cube3D bar through the plantfem_gpu FEMSolver_ surface,
! written the way Tutorial/fem/ModalAnalysis_FixBoundary.f90 drives FEMSolver_: element mass
! matrices set and kept as "B", zeros(), element stiffness matrices set and kept as "A",
! top and bottom faces removed with fix_eig, then eig.
!
!   modal_bar NX NY NZ [nev]   nev > 0: the nev lowest pairs (LOBPCG); absent: every pair.
!                              Prints the count and the eigenvalues, one per line.
program modal_bar
   use plantfem_gpu
   use iso_fortran_env, only: int32, real64
   implicit none
   type(FEMDomain_), target :: bar
   type(FEMSolver_) :: solver
   real(real64), allocatable :: eigen_value(:), eigen_vectors(:, :)
   integer(int32), allocatable :: nodes(:), ids(:)
   integer(int32) :: nx, ny, nz, nev, i
   character(len=64) :: arg

   call get_command_argument(1, arg); read (arg, *) nx
   call get_command_argument(2, arg); read (arg, *) ny
   call get_command_argument(3, arg); read (arg, *) nz
   nev = 0
   call get_command_argument(4, arg)
   if (len_trim(arg) > 0) read (arg, *) nev

   call bar%create("Cube3D", x_num=nx, y_num=ny, z_num=nz)
   call solver%init(NumDomain=1)
   call solver%setDomain(FEMDomain=bar, DomainID=1)
   call solver%setCRS(DOF=3)
   do i = 1, bar%ne()
      call solver%setMatrix(DomainID=1, ElementID=i, DOF=3, &
                            Matrix=bar%MassMatrix(ElementID=i, Density=2.70d0, DOF=3))
   end do
   call solver%keepThisMatrixAs("B")
   call solver%zeros()
   do i = 1, bar%ne()
      call solver%setMatrix(DomainID=1, ElementID=i, DOF=3, &
                            Matrix=bar%StiffnessMatrix(ElementID=i, E=70000.0d0, v=0.300d0))
   end do
   call solver%keepThisMatrixAs("A")

   nodes = bar%select(z_min=bar%z_max())
   ids = [3*nodes - 2, 3*nodes - 1, 3*nodes]
   call solver%fix_eig(IDs=ids)
   nodes = bar%select(z_max=bar%z_min())
   ids = [3*nodes - 2, 3*nodes - 1, 3*nodes]
   call solver%fix_eig(IDs=ids)

   if (nev > 0) then
      call solver%eig(num_eigen=nev, eigen_value=eigen_value, eigen_vectors=eigen_vectors)
   else
      call solver%eig(eigen_value=eigen_value, eigen_vectors=eigen_vectors)
   end if
   print '(a,i0)', "count ", size(eigen_value)
   do i = 1, size(eigen_value)
      print '(es25.17)', eigen_value(i)
   end do
end program modal_bar
