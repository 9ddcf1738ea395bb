! overset_organs -- two overlapping Cube3D organs solved as one system, in the order
! Soybean%getDisplacement drives FEMSolver_ (SoybeanClass.f90:8340-8533): overset, init with
! several domains, setDomain, setCRS, per-domain setMatrix, setEbOM, fix, solve.
!
!   overset_organs out.bin [P2P]   writes the connect count, nnz and x (real64 stream);
!                                  P2P couples the lower organ's nodes instead of its Gauss points
program overset_organs
   use plantfem_gpu
   use iso_fortran_env, only: int32, int64, real64
   implicit none
   type(FEMDomain_), target :: organ(2)
   type(FEMSolver_) :: solver
   integer(int32), allocatable :: FixBoundary(:)
   real(real64), allocatable :: x(:), rhs(:)
   integer(int32) :: d, e, i, u, nn1
   character(len=256) :: outfile, alg
   integer(int32) :: algorithm

   call get_command_argument(1, outfile)
   call get_command_argument(2, alg)
   algorithm = FEMDomain_Overset_GPP
   if (trim(alg) == "P2P") algorithm = FEMDomain_Overset_P2P
   call organ(1)%create("Cube3D", x_num=3, y_num=3, z_num=4, x_len=1.0d0, y_len=1.0d0, z_len=2.0d0)
   call organ(2)%create("Cube3D", x_num=3, y_num=3, z_num=3, x_len=1.0d0, y_len=1.0d0, z_len=1.2d0)
   call organ(2)%move(x=0.05d0, y=0.07d0, z=1.43d0)

   call organ(1)%overset(FEMDomain=organ(2), DomainID=2, MyDomainID=1, algorithm=algorithm)

   call solver%init(NumDomain=2)
   call solver%setDomain(FEMDomain=organ(1), DomainID=1)
   call solver%setDomain(FEMDomain=organ(2), DomainID=2)
   call solver%setCRS(DOF=3)
   do d = 1, 2
      do e = 1, organ(d)%ne()
         call solver%setMatrix(DomainID=d, ElementID=e, DOF=3, &
                               Matrix=organ(d)%StiffnessMatrix(ElementID=e, E=1.0d4, v=0.3d0))
      end do
   end do
   call solver%setEbOM(penalty=1.0d6, DOF=3)

   ! load along x on the top face of the upper organ
   nn1 = organ(1)%nn()
   allocate (rhs(3*(nn1 + organ(2)%nn())))
   rhs = 0.0d0
   do i = 1, organ(2)%nn()
      if (organ(2)%mesh%NodCoord(i, 3) > organ(2)%z_max() - 1.0d-9) rhs(3*(nn1 + i) - 2) = 0.1d0
   end do
   call solver%setRHS(rhs)
   FixBoundary = organ(1)%select(z_max=organ(1)%z_min())*3 - 2
   call solver%fix(IDs=FixBoundary, FixValue=0.0d0)
   FixBoundary = organ(1)%select(z_max=organ(1)%z_min())*3 - 1
   call solver%fix(IDs=FixBoundary, FixValue=0.0d0)
   FixBoundary = organ(1)%select(z_max=organ(1)%z_min())*3
   call solver%fix(IDs=FixBoundary, FixValue=0.0d0)
   solver%relative_er = 1.0d-12
   solver%er0 = 1.0d-30
   x = solver%solve()
   print '(a,i0,a,i0,a,i0)', "connects ", organ(1)%numOversetElements(), " nnz ", solver%nnz, &
      " iterations ", solver%stats%iterations
   if (len_trim(outfile) > 0) then
      open (newunit=u, file=trim(outfile), access="stream", form="unformatted", status="replace")
      write (u) int(organ(1)%numOversetElements(), int64), solver%nnz, x
      close (u)
   end if
   call solver%remove()
end program
