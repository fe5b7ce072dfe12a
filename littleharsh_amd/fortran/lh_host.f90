! lh_host.f90 -- the reference host's hand-over to the GPU solver.
!
! After call start(myid,status,ierr) (littleharsh.f90:98) has read input.in and built the
! grid, the wavenumbers and the canopy lists in the module `declaration`, lh_config_from_start
! fills an lh_config with them; lh_create then uploads everything (it keeps no pointer).  One
! resolution is assumed (every column in the middle band, start.f90:1038-1040) and one rank
! (the single-rank proc_lims_planes / getbounds of lh_single_rank.f90).
module lh_host
  use iso_c_binding
  use declaration
  use lh_amd
  implicit none
  ! copies lh_config points into; they live until the next call
  real(c_double), allocatable, target, private :: h_yu(:), h_yv(:), h_du(:), h_dv(:)
  real(c_double), allocatable, target, private :: h_d2u(:), h_d2v(:)
  real(c_double), allocatable, target, private :: h_k1x(:), h_k2x(:), h_k1z(:), h_k2z(:)
  real(c_double), target, private :: h_gw(2), h_rk(15)
  integer(c_int), allocatable, target, private :: h_lu(:), h_lv(:)
contains

  subroutine lh_config_from_start(cfg, precision)
    type(lh_config), intent(out) :: cfg
    integer(c_int), intent(in) :: precision
    integer :: i, k, nu, nv
    if (Ngal(1,botband) /= Ngal(1,midband) .or. Ngal(2,botband) /= Ngal(2,midband)) then
      write(*,*) 'lh_config_from_start: the GPU solver runs one resolution in every band'
      stop
    end if
    cfg%nx = N(1,midband);     cfg%nz = N(2,midband)
    cfg%px = Ngal(1,midband);  cfg%pz = Ngal(2,midband)
    cfg%ny = N(3,nband) - dsty + 2          ! input ny: N(3,nband) = ny-2+dsty (y_grid_canopy.f90)
    cfg%dsty = dsty
    cfg%re = Re;  cfg%alp = alp;  cfg%bet = bet
    cfg%flag_ctpress = flag_ctpress;  cfg%mpgx = mpgx;  cfg%qxt = QxT;  cfg%ddthetavi = ddthetavi
    nu = N(4,nband) + 1 - N(4,0) + 1
    nv = N(3,nband) + 1 - N(3,0) + 1
    h_yu = yu(N(4,0):N(4,nband)+1);        h_du = dthdyu(N(4,0):N(4,nband)+1)
    h_yv = yv(N(3,0):N(3,nband)+1);        h_dv = dthdyv(N(3,0):N(3,nband)+1)
    h_d2u = reshape(dyu2i(:,N(4,0):N(4,nband)+1), [3*nu])
    h_d2v = reshape(dyv2i(:,N(3,0):N(3,nband)+1), [3*nv])
    cfg%yu = c_loc(h_yu);  cfg%yv = c_loc(h_yv);  cfg%dthdyu = c_loc(h_du);  cfg%dthdyv = c_loc(h_dv)
    cfg%dyu2i = c_loc(h_d2u);  cfg%dyv2i = c_loc(h_d2v)
    h_gw = gridweighting(midband,1:2);  cfg%gridweighting = c_loc(h_gw)
    h_rk = [aRK, bRK, gRK, cRK, dRK];   cfg%rk = c_loc(h_rk)
    ! def_k (start.f90:655-704): Im k1F and k2F, x modes 0..N/2 and z modes in FFT order
    h_k1x = [(dimag(k1F_x(i)), i = 0, N(1,midband)/2)]
    h_k2x = [(k2F_x(i), i = 0, N(1,midband)/2)]
    h_k1z = [(dimag(k1F_z(k)), k = 1, N(2,midband))]
    h_k2z = [(k2F_z(k), k = 1, N(2,midband))]
    cfg%k1x = c_loc(h_k1x);  cfg%k2x = c_loc(h_k2x);  cfg%k1z = c_loc(h_k1z);  cfg%k2z = c_loc(h_k2z)
    ! the canopy points after getbounds, list_ib(1:3, n, grid)
    h_lu = reshape(list_ib(:, 1:nlist_ib(ugrid), ugrid), [3*nlist_ib(ugrid)])
    h_lv = reshape(list_ib(:, 1:nlist_ib(vgrid), vgrid), [3*nlist_ib(vgrid)])
    cfg%nlist_u = nlist_ib(ugrid);  cfg%list_u = c_loc(h_lu)
    cfg%nlist_v = nlist_ib(vgrid);  cfg%list_v = c_loc(h_lv)
    cfg%tip_u = [0, Ngal(4,3)-dsty+1]       ! tridLU_3D_IB.f90:237
    cfg%tip_v = [-1, Ngal(3,3)-dsty+2]      ! tridLU_3D_IB.f90:317
    cfg%ntilex = ntilex;  cfg%ntilez = ntilez;  cfg%band_v = [N(3,1), N(3,2)]
    cfg%precision = precision
  end subroutine

  ! the lh_config as a stream file: the layout of littleharsh_amd.fortran_io.write_case up to
  ! the band limits (no state)
  subroutine lh_config_write(cfg, fname)
    type(lh_config), intent(in) :: cfg
    character(len=*), intent(in) :: fname
    integer :: un
    open(newunit=un, file=fname, access='stream', form='unformatted', status='replace')
    write(un) cfg%nx, cfg%nz, cfg%px, cfg%pz, cfg%ny, cfg%dsty, cfg%flag_ctpress, cfg%precision
    write(un) cfg%re, cfg%alp, cfg%bet, cfg%mpgx, cfg%qxt, cfg%ddthetavi
    write(un) h_yu, h_yv, h_du, h_dv, h_d2u, h_d2v, h_gw, h_rk, h_k1x, h_k2x, h_k1z, h_k2z
    write(un) cfg%nlist_u, cfg%nlist_v
    write(un) h_lu, h_lv
    write(un) cfg%tip_u, cfg%tip_v, cfg%ntilex, cfg%ntilez, cfg%band_v
    close(un)
  end subroutine
end module
