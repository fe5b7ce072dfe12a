! lh_single_rank.f90 -- the plane and canopy-list set-up of the reference host for one MPI rank.
!
! The GPU solver holds every plane and every spectral column of the channel in one process,
! so the Fortran host runs with np = 1.  Two routines of start.f90 cannot do that:
!
!   proc_lims_planes (start.f90:1268-1486) gives floor(r np / (2r+1)) = 0 ranks to each wall
!     band and then divides the band's planes by that count (start.f90:1316-1339), an
!     integer division by zero;
!   getbounds (start.f90:706-998) splits the ranks into lower and upper walls by
!     iproc < np/2 (start.f90:769), so the single rank gets the upper-wall plane range and
!     then keeps the lower-wall points inside it, i.e. none.
!
! These replacements give rank 0 every plane of the three bands and the canopy points of both
! walls.  The host links them in place of the two routines of its start.f90 object (whose
! symbols are made weak at link time); the rest of start() runs unchanged.  They stop with a
! message, as the reference does, when np /= 1.

subroutine proc_lims_planes(myid)
   use declaration
   implicit none
   integer :: myid

   if (np /= 1 .or. myid /= 0) then
      write(*,*) 'proc_lims_planes: the single-rank set-up needs np = 1'
      stop
   end if
   allocate(procs(0:nband), procs_b(0:nband), procs_c(0:nband))
   allocate(planelim(3,2,0:0), limPL_incw(3,2,0:0), limPL_excw(3,2,0:0))
   allocate(bandPL(0:0), bandPL_FFT(0:0))
   ! ranks up to and including each band (the cumulative counts of start.f90:1373-1460):
   ! the one rank is counted in the middle band, whose resolution it works at
   procs(0) = 0
   procs(botband) = 0
   procs(midband) = 1
   procs(topband) = 1
   procs_b = procs
   procs_c = procs
   ! planes without the walls (the convention of planelim): u/w and v from the first interior
   ! plane to the last; p from the first to the last plane the band-wise split gives
   ! (start.f90:1400-1458 with one rank per band: N(4,0)+2 .. N(4,nband)-1)
   planelim(ugrid,1,0) = N(4,0) + 1
   planelim(ugrid,2,0) = N(4,nband)
   planelim(vgrid,1,0) = N(3,0) + 1
   planelim(vgrid,2,0) = N(3,nband)
   planelim(pgrid,1,0) = N(4,0) + 2
   planelim(pgrid,2,0) = N(4,nband) - 1
   limPL_incw = planelim
   limPL_incw(:,1,0) = planelim(:,1,0) - 1
   limPL_incw(:,2,0) = planelim(:,2,0) + 1
   limPL_excw = planelim
   bandPL(0) = midband
   bandPL_FFT(0) = midband
   igal = Ngal(1,midband) + 2
   kgal = Ngal(2,midband)
   jgal(ugrid,1) = limPL_excw(ugrid,1,0)
   jgal(ugrid,2) = limPL_excw(ugrid,2,0)
   jgal(vgrid,1) = limPL_excw(vgrid,1,0)
   jgal(vgrid,2) = limPL_excw(vgrid,2,0)
   jgal(pgrid,1) = limPL_excw(pgrid,1,0)
   jgal(pgrid,2) = limPL_excw(pgrid,2,0)
end subroutine

subroutine getbounds(myid,status,ierr)
   use declaration
   implicit none
   include 'mpif.h'
   integer status(MPI_STATUS_SIZE),ierr,myid
   integer :: nlist_ib1, nlist_ib2, grd, ilst, nkept
   integer, allocatable :: list_ib1(:,:,:), list_ib2(:,:,:)

   if (np /= 1 .or. myid /= 0) then
      write(*,*) 'getbounds: the single-rank set-up needs np = 1'
      stop
   end if
   allocate(nlist_ib(2))
   allocate(nyuIB1(0:0), nyuIB2(0:0), nyvIB1(0:0), nyvIB2(0:0))
   if (geometry_type /= 3) then
      ! smooth channel: no canopy planes (start.f90:732-737)
      nlist_ib = 0
      nyuIB1 = -88
      nyuIB2 = -99
      nyvIB1 = -88
      nyvIB2 = -99
      allocate(list_ib(3,0,2))
      return
   end if
   ! the lists of both walls, as the master reads them back from boundary_canopies' file
   ! (start.f90:745-754)
   call boundary_canopies
   open(10,file=boundfname,form='unformatted',access='stream')
   read(10) Lx,Ly,Lz
   read(10) Ngal,nlist_ib1,nlist_ib2,nyu11,nyu21,nyu12,nyu22,nyv11,nyv21,nyv12,nyv22
   allocate(list_ib1(3,nlist_ib1,2),list_ib2(3,nlist_ib2,2))
   read(10) list_ib1,list_ib2
   close(10)
   ! the canopy planes of the rank: from the lower wall's first to the upper wall's last, cut
   ! to the planes it solves (the per-rank rule of start.f90:767-781, both walls at once)
   nyuIB1(0) = max(nyu11, planelim(ugrid,1,0))
   nyuIB2(0) = min(nyu22, planelim(ugrid,2,0))
   nyvIB1(0) = max(nyv11, planelim(vgrid,1,0))
   nyvIB2(0) = min(nyv22, planelim(vgrid,2,0))
   ! the points of both walls, lower wall first, in list order, each wall's list cut to that
   ! wall's canopy planes as the per-rank filter does (start.f90:767-781, 786-815): the
   ! upper list of boundary_canopies can hold v points below nyv12, which it drops
   nkept = 0
   do grd = 1,2
      nlist_ib(grd) = count_in(list_ib1, nlist_ib1, grd, 1) + count_in(list_ib2, nlist_ib2, grd, 2)
   end do
   nkept = max(nlist_ib(1), nlist_ib(2))
   allocate(list_ib(3,nkept,2))
   list_ib = 0
   do grd = 1,2
      nkept = 0
      call keep(list_ib1, nlist_ib1, grd, 1)
      call keep(list_ib2, nlist_ib2, grd, 2)
   end do
   deallocate(list_ib1, list_ib2)

contains

   ! plane j of grid grd inside the canopy planes of wall 1 (lower) or 2 (upper) that the
   ! rank solves
   logical function inside(j, grd, wall)
      integer :: j, grd, wall
      integer :: lo, hi
      if (grd == ugrid) then
         lo = merge(nyu11, nyu12, wall == 1)
         hi = merge(nyu21, nyu22, wall == 1)
         inside = j >= max(lo, planelim(ugrid,1,0)) .and. j <= min(hi, planelim(ugrid,2,0))
      else
         lo = merge(nyv11, nyv12, wall == 1)
         hi = merge(nyv21, nyv22, wall == 1)
         inside = j >= max(lo, planelim(vgrid,1,0)) .and. j <= min(hi, planelim(vgrid,2,0))
      end if
   end function

   integer function count_in(lst, nl, grd, wall)
      integer :: nl, grd, wall
      integer :: lst(3,nl,2)
      count_in = 0
      do ilst = 1,nl
         if (inside(lst(3,ilst,grd), grd, wall)) count_in = count_in + 1
      end do
   end function

   subroutine keep(lst, nl, grd, wall)
      integer :: nl, grd, wall
      integer :: lst(3,nl,2)
      do ilst = 1,nl
         if (inside(lst(3,ilst,grd), grd, wall)) then
            nkept = nkept + 1
            list_ib(:,nkept,grd) = lst(:,ilst,grd)
         end if
      end do
   end subroutine
end subroutine
