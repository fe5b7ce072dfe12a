! lh_demo.f90 -- a Fortran host driving the GPU solver through the lh_amd module.
!
! It reads a channel set-up and initial state from a stream file (the layout written by
! littleharsh_amd.fortran_io.write_case), runs `nsteps` RK3 steps with the substep loop of
! the reference host (littleharsh.f90:164-190, one lh_substep per kRK), and writes the
! final state, max|div| and the flow rate to a second stream file.
!
! usage: lh_demo case.bin result.bin
program lh_demo
  use iso_c_binding
  use lh_amd
  implicit none
  type(lh_config) :: cfg
  type(c_ptr) :: h
  integer(c_int) :: st, krk, n, nsteps
  integer(c_int) :: hdr(8), ilist(2), tips(4), tiles(2), band(2)
  real(c_double) :: phys(6), dt, err, mpgx, qx, dtl, cmax
  real(c_double), allocatable, target :: yu(:), yv(:), dthdyu(:), dthdyv(:), dyu2i(:), dyv2i(:)
  real(c_double), allocatable, target :: gw(:), rk(:), k1x(:), k2x(:), k1z(:), k2z(:)
  integer(c_int), allocatable, target :: list_u(:), list_v(:)
  complex(c_double_complex), allocatable, target :: u1(:,:), u2(:,:), u3(:,:), p(:,:)
  integer :: nyu, nyv, nyp, mx, ncol, un
  character(len=512) :: fin, fout

  call get_command_argument(1, fin)
  call get_command_argument(2, fout)
  open(newunit=un, file=trim(fin), access='stream', form='unformatted', status='old')
  read(un) hdr
  read(un) phys
  cfg%nx = hdr(1); cfg%nz = hdr(2); cfg%px = hdr(3); cfg%pz = hdr(4)
  cfg%ny = hdr(5); cfg%dsty = hdr(6); cfg%flag_ctpress = hdr(7); cfg%precision = hdr(8)
  cfg%re = phys(1); cfg%alp = phys(2); cfg%bet = phys(3)
  cfg%mpgx = phys(4); cfg%qxt = phys(5); cfg%ddthetavi = phys(6)
  nyu = cfg%ny + 2*cfg%dsty + 1
  nyv = cfg%ny + 2*cfg%dsty
  nyp = cfg%ny + 2*cfg%dsty - 1
  mx = cfg%nx/2 + 1
  ncol = mx*cfg%nz
  allocate(yu(nyu), yv(nyv), dthdyu(nyu), dthdyv(nyv), dyu2i(3*nyu), dyv2i(3*nyv))
  allocate(gw(2), rk(15), k1x(mx), k2x(mx), k1z(cfg%nz), k2z(cfg%nz))
  read(un) yu, yv, dthdyu, dthdyv, dyu2i, dyv2i, gw, rk, k1x, k2x, k1z, k2z
  read(un) ilist
  allocate(list_u(max(1, 3*ilist(1))), list_v(max(1, 3*ilist(2))))
  if (ilist(1) > 0) read(un) list_u(1:3*ilist(1))
  if (ilist(2) > 0) read(un) list_v(1:3*ilist(2))
  read(un) tips, tiles, band
  ! the reference's cfield storage: f(j, column)
  allocate(u1(nyu, ncol), u2(nyv, ncol), u3(nyu, ncol), p(nyp, ncol))
  read(un) u1, u2, u3, p
  read(un) nsteps, dt
  close(un)

  cfg%yu = c_loc(yu); cfg%yv = c_loc(yv); cfg%dthdyu = c_loc(dthdyu); cfg%dthdyv = c_loc(dthdyv)
  cfg%dyu2i = c_loc(dyu2i); cfg%dyv2i = c_loc(dyv2i)
  cfg%gridweighting = c_loc(gw); cfg%rk = c_loc(rk)
  cfg%k1x = c_loc(k1x); cfg%k2x = c_loc(k2x); cfg%k1z = c_loc(k1z); cfg%k2z = c_loc(k2z)
  cfg%nlist_u = ilist(1); cfg%nlist_v = ilist(2)
  cfg%list_u = c_loc(list_u); cfg%list_v = c_loc(list_v)
  cfg%tip_u = tips(1:2); cfg%tip_v = tips(3:4)
  cfg%ntilex = tiles(1); cfg%ntilez = tiles(2); cfg%band_v = band

  st = lh_create(cfg, h)
  call lh_check(h, st, 'lh_create')
  st = lh_set_state(h, c_loc(u1), c_loc(u2), c_loc(u3), c_loc(p))
  call lh_check(h, st, 'lh_set_state')
  do n = 1, nsteps
    do krk = 1, 3                       ! littleharsh.f90:164
      st = lh_substep(h, krk, dt)       ! replaces littleharsh.f90:166-188
      call lh_check(h, st, 'lh_substep')
    end do
  end do
  st = lh_get_state(h, c_loc(u1), c_loc(u2), c_loc(u3), c_loc(p))
  call lh_check(h, st, 'lh_get_state')
  st = lh_max_div(h, err)
  call lh_check(h, st, 'lh_max_div')
  st = lh_get_scalars(h, mpgx, qx, dtl, cmax)
  call lh_check(h, st, 'lh_get_scalars')
  st = lh_destroy(h)
  write(*,'(a,i0,a,es12.4,a,es22.14)') 'lh_demo: ', nsteps, ' steps, max|div| = ', err, ', mpgx = ', mpgx
  open(newunit=un, file=trim(fout), access='stream', form='unformatted', status='replace')
  write(un) u1, u2, u3, p, err, mpgx
  close(un)
end program
