! organ_meshes -- the organ meshes of FEMDomain_%create: a default Leaf3D blade, a soybean
! Leaf3D blade and a Cylinder stem, written for inspection (host only, no device work).
!
!   organ_meshes out.bin [x_num y_num]    per mesh: nn, ne (int64), NodCoord (nn x 3), ElemNod (ne x 8);
!                                         x_num y_num: cells of the Cylinder stem (default 10 10)
program organ_meshes
   use plantfem_gpu
   use iso_fortran_env, only: int32, int64, real64
   implicit none
   type(FEMDomain_) :: leaf, soy, stem
   character(len=256) :: outfile, arg
   integer :: u, cx, cy

   call get_command_argument(1, outfile)
   cx = 10
   cy = 10
   if (command_argument_count() >= 3) then
      call get_command_argument(2, arg)
      read (arg, *) cx
      call get_command_argument(3, arg)
      read (arg, *) cy
   end if
   call leaf%create("Leaf3D", x_num=8, y_num=2, z_num=20, x_len=0.2d0, y_len=0.01d0, z_len=1.0d0, shaperatio=0.3d0)
   call soy%create("Leaf3D", x_num=8, y_num=2, z_num=30, x_len=0.1d0, y_len=0.01d0, z_len=0.3d0, species=1, &
                   SoyWidthRatio=0.5d0)
   call stem%create("Cylinder", x_num=cx, y_num=cy, z_num=6, x_len=0.08d0, y_len=0.08d0, z_len=0.6d0)
   open (newunit=u, file=trim(outfile), access="stream", form="unformatted", status="replace")
   call put(leaf)
   call put(soy)
   call put(stem)
   close (u)
   print '(a,3(1x,i0))', "nodes", leaf%nn(), soy%nn(), stem%nn()
contains
   subroutine put(d)
      type(FEMDomain_), intent(in) :: d
      write (u) int(size(d%mesh%NodCoord, 1), int64), int(size(d%mesh%ElemNod, 1), int64), d%mesh%NodCoord, &
         d%mesh%ElemNod
   end subroutine
end program
