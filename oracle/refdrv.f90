! refdrv.f90 -- test driver (project code) that calls the reference's own routines, compiled
! from their sources in the reference tree, to produce golden vectors for the oracle.
! It is checker infrastructure only: it sets the state of the reference module
! `declaration` for one process and one resolution band, then calls the routines with plain
! arrays handed over from C.  Nothing here is part of the solver.
!
! Every entry point is bind(C).  Complex arrays are interleaved (re, im) doubles; the
! column layout is the reference's f(j, column).
module refdrv
  use iso_c_binding
  use declaration
  implicit none
contains

  ! Set up sizes, grid (y_grid_canopy), wavenumbers (def_k), RK tables, column maps and
  ! band limits for one rank with every column in the middle band.
  subroutine ref_setup(nx, nz, ny_in, dsty_in, posth_in, dyq_in, re_in, alp_in, bet_in, &
                       mpgx_in, dirlen, dirc) bind(C, name="ref_setup")
    integer(c_int), value :: nx, nz, ny_in, dsty_in, dirlen
    real(c_double), value :: posth_in, dyq_in, re_in, alp_in, bet_in, mpgx_in
    character(kind=c_char), intent(in) :: dirc(dirlen)
    integer :: i, k, c, iband, ngx, ngz

    np = 1
    Re = re_in
    alp = alp_in
    bet = bet_in
    mpgx = mpgx_in
    Lx = 2d0*pi/alp
    Lz = 2d0*pi/bet
    Ly = 2d0
    dsty = dsty_in
    posth = posth_in
    dyq = dyq_in
    ppp = 5
    dirout = ''
    do i = 1, dirlen
      dirout(i:i) = dirc(i)
    end do
    ngx = 3*nx/2
    ngz = 3*nz/2
    ! start.f90:93-118 with both bands at one resolution
    allocate(Ngal(4,0:nband+1), N(4,0:nband+1), Ny(3,0:nband+1))
    Ngal = 0
    Ngal(1,1:nband) = ngx
    Ngal(1,nband+1) = -2
    Ngal(2,1:nband) = ngz
    Ngal(3,nband) = ny_in - 2
    N = Ngal
    do iband = 1, nband
      N(1,iband) = 2*(Ngal(1,iband)/3)
      N(2,iband) = 2*(Ngal(2,iband)/3)
    end do
    allocate(h_ny(botband))
    h_ny(1) = 1.5d0*posth
    call y_grid_canopy
    call def_k
    ! RK coefficients as start.f90:413-427
    aRK = (/ 4d0/15d0, 1d0/15d0, 1d0/6d0 /)
    bRK = aRK
    gRK = (/ 8d0/15d0, 5d0/12d0, 3d0/4d0 /)
    cRK = (/ 0d0, -17d0/60d0, -5d0/12d0 /)
    dRK = (/ 8d0/15d0, 2d0/15d0, 1d0/3d0 /)
    ! columns: every column in the middle band, k outer (start.f90:1105-1119)
    allocate(columns_num(nband,0:0))
    columns_num = 0
    columns_num(midband,0) = (N(1,midband)/2+1)*N(2,midband)
    allocate(columns_i(columns_num(midband,0),nband,0:0), columns_k(columns_num(midband,0),nband,0:0))
    c = 0
    do k = 1, N(2,midband)
      do i = 0, N(1,midband)/2
        c = c + 1
        columns_i(c,midband,0) = i
        columns_k(c,midband,0) = k
      end do
    end do
    allocate(jlim(2,3,nband))
    jlim = 0
    jlim(1,ugrid,2) = N(4,0);   jlim(2,ugrid,2) = N(4,3)+1
    jlim(1,vgrid,2) = N(3,0);   jlim(2,vgrid,2) = N(3,3)+1
    jlim(1,pgrid,2) = N(4,0)+1; jlim(2,pgrid,2) = N(4,3)
    jlim(:,:,1) = jlim(:,:,2)
    jlim(:,:,3) = jlim(:,:,2)
    ! gridweighting, start.f90:497-503
    allocate(gridweighting(nband,2))
    gridweighting(:,1) = -(yv(N(3,0))-yu(N(4,0)))/(yv(N(3,0))-yu(N(4,0)+1))
    gridweighting(:,2) = (yu(N(4,3)+1)-yv(N(3,3)+1))/(yv(N(3,3)+1)-yu(N(4,3)))
    allocate(bandPL(0:0))
    bandPL(0) = midband
    jgal = 0
    allocate(u11(0:nn+2))
    allocate(DG(sband:eband))
    do iband = sband, eband
      allocate(DG(iband)%f_dg(3,jlim(1,pgrid,iband):jlim(2,pgrid,iband),columns_num(iband,0)))
      DG(iband)%f_dg = 0d0
    end do
    call LU_buildP(jlim(1,pgrid,midband), jlim(2,pgrid,midband), 0, midband, DG)
  end subroutine

  subroutine ref_sizes(out) bind(C, name="ref_sizes")
    integer(c_int), intent(out) :: out(8)
    out(1) = N(4,0); out(2) = N(4,3); out(3) = N(3,0); out(4) = N(3,3)
    out(5) = nn; out(6) = columns_num(midband,0); out(7) = N(3,1); out(8) = N(3,2)
  end subroutine

  ! metric arrays in the reference index ranges
  subroutine ref_grid(yu_o, yv_o, du_o, dv_o, d2u_o, d2v_o, gw_o) bind(C, name="ref_grid")
    real(c_double), intent(out) :: yu_o(*), yv_o(*), du_o(*), dv_o(*), d2u_o(*), d2v_o(*), gw_o(2)
    integer :: nu, nv
    nu = N(4,3) + 1 - N(4,0) + 1
    nv = N(3,3) + 1 - N(3,0) + 1
    yu_o(1:nu) = yu(N(4,0):N(4,3)+1)
    du_o(1:nu) = dthdyu(N(4,0):N(4,3)+1)
    yv_o(1:nv) = yv(N(3,0):N(3,3)+1)
    dv_o(1:nv) = dthdyv(N(3,0):N(3,3)+1)
    d2u_o(1:3*nu) = reshape(dyu2i(:,N(4,0):N(4,3)+1), (/3*nu/))
    d2v_o(1:3*nv) = reshape(dyv2i(:,N(3,0):N(3,3)+1), (/3*nv/))
    gw_o(1) = gridweighting(midband,1)
    gw_o(2) = gridweighting(midband,2)
  end subroutine

  subroutine ref_wavenumbers(k1x, k2x, k1z, k2z) bind(C, name="ref_wavenumbers")
    real(c_double), intent(out) :: k1x(*), k2x(*), k1z(*), k2z(*)
    integer :: i, k
    do i = 0, N(1,1)/2
      k1x(i+1) = dimag(k1F_x(i))
      k2x(i+1) = k2F_x(i)
    end do
    do k = 1, N(2,1)
      k1z(k) = dimag(k1F_z(k))
      k2z(k) = k2F_z(k)
    end do
  end subroutine

  ! canopy lists: boundary_canopies writes them to a stream file in dirout
  subroutine ref_canopies(ntx, ntz, dx, dz) bind(C, name="ref_canopies")
    integer(c_int), value :: ntx, ntz, dx, dz
    ntilex = ntx
    ntilez = ntz
    dstx = dx
    dstz = dz
    write(ext1,'(i4.4)') N(1,1)
    write(ext2,'(i4.4)') N(2,1)
    write(ext3,'(i4.4)') nn+2
    call boundary_canopies
  end subroutine

  ! RHS0_u1/u2/u3 (littleharsh.f90:614-731) on state arrays; du and Lu returned
  subroutine ref_rhs0(krk_in, u1, u2, u3, p, nu1, nu2, nu3, du1, du2, du3, lu1, lu2, lu3) &
      bind(C, name="ref_rhs0")
    integer(c_int), value :: krk_in
    complex(c_double_complex), intent(in) :: u1(*), u2(*), u3(*), p(*), nu1(*), nu2(*), nu3(*)
    complex(c_double_complex), intent(out) :: du1(*), du2(*), du3(*), lu1(*), lu2(*), lu3(*)
    type(cfield) :: fu1(sband:eband), fu2(sband:eband), fu3(sband:eband), fp(sband:eband)
    type(cfield) :: fn1(sband:eband), fn2(sband:eband), fn3(sband:eband)
    type(cfield) :: fd1(sband:eband), fd2(sband:eband), fd3(sband:eband)
    type(cfield) :: fl1(sband:eband), fl2(sband:eband), fl3(sband:eband)
    kRK = krk_in
    call mk(fu1, ugrid, 0, u1); call mk(fu2, vgrid, 0, u2); call mk(fu3, ugrid, 0, u3)
    call mk(fp, pgrid, 0, p)
    call mk(fn1, ugrid, 1, nu1); call mk(fn2, vgrid, 1, nu2); call mk(fn3, ugrid, 1, nu3)
    call mkz(fd1, ugrid); call mkz(fd2, vgrid); call mkz(fd3, ugrid)
    call mkz(fl1, ugrid); call mkz(fl2, vgrid); call mkz(fl3, ugrid)
    call RHS0_u1(fd1, fu1, fn1, fp, fl1, 0)
    call RHS0_u2(fd2, fu2, fn2, fp, fl2, 0)
    call RHS0_u3(fd3, fu3, fn3, fp, fl3, 0)
    call out(fd1, du1); call out(fd2, du2); call out(fd3, du3)
    call out(fl1, lu1); call out(fl2, lu2); call out(fl3, lu3)
  end subroutine

  ! solveP + v_corr (littleharsh.f90:557-612, 826-937) with dt, kRK
  subroutine ref_solvep_vcorr(krk_in, dt_in, u1, u2, u3, p, psi_o, div_o) bind(C, name="ref_solvep_vcorr")
    integer(c_int), value :: krk_in
    real(c_double), value :: dt_in
    complex(c_double_complex), intent(inout) :: u1(*), u2(*), u3(*), p(*)
    complex(c_double_complex), intent(out) :: psi_o(*), div_o(*)
    type(cfield) :: fu1(sband:eband), fu2(sband:eband), fu3(sband:eband), fp(sband:eband)
    type(cfield) :: fpsi(sband:eband), fdiv(sband:eband)
    integer :: status(1), ierr
    kRK = krk_in
    dt = dt_in
    dti = 1d0/dt
    call mk(fu1, ugrid, 0, u1); call mk(fu2, vgrid, 0, u2); call mk(fu3, ugrid, 0, u3)
    call mk(fp, pgrid, 0, p)
    call mkz(fpsi, pgrid); call mkz(fdiv, pgrid)
    call solveP(fp, fpsi, fu1, fu2, fu3, fdiv, 0)
    call out(fpsi, psi_o)
    call v_corr(fu1, fu2, fu3, fpsi, fdiv, 0, status, ierr)
    call out(fu1, u1); call out(fu2, u2); call out(fu3, u3); call out(fp, p)
    call out(fdiv, div_o)
  end subroutine

  ! LU_build + LU_dec of the velocity matrices (tridLU_3D_IB.f90:427-493, 601-640)
  subroutine ref_lu_vel(krk_in, dt_in, grid, ay_o, axz_o) bind(C, name="ref_lu_vel")
    integer(c_int), value :: krk_in, grid
    real(c_double), value :: dt_in
    real(c_double), intent(out) :: ay_o(*), axz_o(*)
    real(8), allocatable :: ay(:,:,:), axz(:,:)
    integer :: nj
    kRK = krk_in
    dt = dt_in
    allocate(ay(3,jlim(1,grid,2):jlim(2,grid,2),nband))
    allocate(axz(maxval(columns_num),nband))
    ay = 0d0
    axz = 0d0
    call LU_build(jlim(1,grid,midband), jlim(2,grid,midband), grid, 0, midband, axz, ay)
    call LU_dec(jlim(1,grid,midband), jlim(2,grid,midband), grid, 0, midband, ay)
    nj = jlim(2,grid,2) - jlim(1,grid,2) + 1
    ay_o(1:3*nj) = reshape(ay(:,:,midband), (/3*nj/))
    axz_o(1:columns_num(midband,0)) = axz(1:columns_num(midband,0),midband)
    deallocate(ay, axz)
  end subroutine

  ! LU_build alone (tridLU_3D_IB.f90:427-493): the unfactored rows of the velocity matrix
  ! (lower, diagonal, upper per plane) and axz per column
  subroutine ref_lu_build(krk_in, dt_in, grid, ay_o, axz_o) bind(C, name="ref_lu_build")
    integer(c_int), value :: krk_in, grid
    real(c_double), value :: dt_in
    real(c_double), intent(out) :: ay_o(*), axz_o(*)
    real(8), allocatable :: ay(:,:,:), axz(:,:)
    integer :: nj
    kRK = krk_in
    dt = dt_in
    allocate(ay(3,jlim(1,grid,2):jlim(2,grid,2),nband))
    allocate(axz(maxval(columns_num),nband))
    ay = 0d0
    axz = 0d0
    call LU_build(jlim(1,grid,midband), jlim(2,grid,midband), grid, 0, midband, axz, ay)
    nj = jlim(2,grid,2) - jlim(1,grid,2) + 1
    ay_o(1:3*nj) = reshape(ay(:,:,midband), (/3*nj/))
    axz_o(1:columns_num(midband,0)) = axz(1:columns_num(midband,0),midband)
    deallocate(ay, axz)
  end subroutine

  ! DG factors after LU_buildP/LU_decP (tridLU_3D_IB.f90:497-571, 643-684)
  subroutine ref_dg(dg_o) bind(C, name="ref_dg")
    real(c_double), intent(out) :: dg_o(*)
    integer :: n
    n = size(DG(midband)%f_dg)
    dg_o(1:n) = reshape(DG(midband)%f_dg, (/n/))
  end subroutine

  ! flowrate_corr + LUsol0 (littleharsh.f90:453-485) on the (0,0) column
  subroutine ref_flowrate_corr(krk_in, dt_in, qxt_in, mpg, u, g) bind(C, name="ref_flowrate_corr")
    integer(c_int), value :: krk_in
    real(c_double), value :: dt_in, qxt_in
    real(c_double), intent(inout) :: mpg
    real(c_double), intent(out) :: g
    complex(c_double_complex), intent(inout) :: u(*)
    kRK = krk_in
    dt = dt_in
    QxT = qxt_in
    call flowrate_corr(u, mpg, g)
  end subroutine

  ! flowrateIm (littleharsh.f90:412-431)
  subroutine ref_flowrate(u, q) bind(C, name="ref_flowrate")
    complex(c_double_complex), intent(in) :: u(*)
    real(c_double), intent(out) :: q
    call flowrateIm(q, u)
  end subroutine

  ! buildsp / buildspUV (spectra.f90:18-56) accumulated once from zero
  subroutine ref_buildsp(grid, u, v, sp_o) bind(C, name="ref_buildsp")
    integer(c_int), value :: grid
    complex(c_double_complex), intent(in) :: u(*), v(*)
    real(c_double), intent(out) :: sp_o(*)
    real(8), allocatable :: sp(:,:)
    integer :: n, g2
    g2 = grid
    if (grid == 0) g2 = ugrid
    allocate(sp(jlim(1,g2,midband):jlim(2,g2,midband), columns_num(midband,0)))
    sp = 0d0
    if (grid == 0) then
      call buildspUV(sp, u, v, 0)
    else
      call buildsp(sp, u, grid, 0)
    end if
    n = size(sp)
    sp_o(1:n) = reshape(sp, (/n/))
    deallocate(sp)
  end subroutine

  ! stats (stats.f90:1-108) with one rank owning every plane, and the conditional folds
  ! (conddown1/2, condup1/2) called as stats does for its bottom and top band ranks.
  ! Physical planes arrive as (px, pz, nplanes) over the full plane range of each grid;
  ! st_o gets the 13 rows in write_stats order (each over the full range of its grid,
  ! padded to nyu); cb_o/ct_o the conditional arrays of the bottom and top bands.
  subroutine ref_stats(ntx, ntz, pu, pv, pw, pp, pvi, pwi, pwx, st_o, cb_o, ct_o) &
      bind(C, name="ref_stats")
    integer(c_int), value :: ntx, ntz
    real(c_double), intent(in) :: pu(*), pv(*), pw(*), pp(*), pvi(*), pwi(*), pwx(*)
    real(c_double), intent(out) :: st_o(*), cb_o(*), ct_o(*)
    integer :: status(64), ierr, gx, gz, lo(3), hi(3), nyu, q, a, g, nb(3)
    real(8), allocatable :: c(:,:,:,:)
    gx = Ngal(1,1)
    gz = Ngal(2,1)
    igal = gx + 2
    kgal = gz
    ntilex = ntx
    ntilez = ntz
    dnx = gx/ntilex
    dnz = gz/ntilez
    lo(ugrid) = N(4,0);   hi(ugrid) = N(4,3)+1
    lo(vgrid) = N(3,0);   hi(vgrid) = N(3,3)+1
    lo(pgrid) = N(4,0)+1; hi(pgrid) = N(4,3)
    nyu = hi(ugrid) - lo(ugrid) + 1
    call plane_in(u1PL, pu, lo(ugrid), hi(ugrid))
    call plane_in(u3PL, pw, lo(ugrid), hi(ugrid))
    call plane_in(u2PL_itp, pvi, lo(ugrid), hi(ugrid))
    call plane_in(u2PL, pv, lo(vgrid), hi(vgrid))
    call plane_in(u3PL_itp, pwi, lo(vgrid), hi(vgrid))
    call plane_in(wx, pwx, lo(vgrid), hi(vgrid))
    call plane_in(ppPL, pp, lo(pgrid), hi(pgrid))
    if (associated(limPL_incw)) deallocate(limPL_incw)
    allocate(limPL_incw(3,2,0:0))
    limPL_incw(:,1,0) = lo
    limPL_incw(:,2,0) = hi
    bandPL(0) = midband
    allocate(Um(lo(ugrid):hi(ugrid)), U2m(lo(ugrid):hi(ugrid)), Wm(lo(ugrid):hi(ugrid)), &
             W2m(lo(ugrid):hi(ugrid)), UVm(lo(ugrid):hi(ugrid)), UWm(lo(ugrid):hi(ugrid)))
    allocate(Vm(lo(vgrid):hi(vgrid)), V2m(lo(vgrid):hi(vgrid)), VWm(lo(vgrid):hi(vgrid)), &
             wxm(lo(vgrid):hi(vgrid)), wx2m(lo(vgrid):hi(vgrid)))
    allocate(Pm(lo(pgrid):hi(pgrid)), P2m(lo(pgrid):hi(pgrid)))
    Um = 0d0; U2m = 0d0; Wm = 0d0; W2m = 0d0; UVm = 0d0; UWm = 0d0
    Vm = 0d0; V2m = 0d0; VWm = 0d0; wxm = 0d0; wx2m = 0d0; Pm = 0d0; P2m = 0d0
    istat = 0
    call stats(0, status, ierr)
    st_o(1:13*nyu) = 0d0
    call row(1, Um); call row(2, U2m); call row(3, Vm); call row(4, V2m)
    call row(5, Wm); call row(6, W2m); call row(7, Pm); call row(8, P2m)
    call row(9, UVm); call row(10, UWm); call row(11, VWm); call row(12, wxm); call row(13, wx2m)
    deallocate(Um, U2m, Wm, W2m, UVm, UWm, Vm, V2m, VWm, wxm, wx2m, Pm, P2m)
    ! conditional: band planes as limPL_incw of the first band-1 rank / last band-3 rank
    nb(ugrid) = N(4,1) - N(4,0) + 1
    nb(vgrid) = N(3,1) - N(3,0) + 1
    nb(pgrid) = N(4,1) - N(4,0) - 1
    allocate(c(dnx, dnz, 0:maxval(nb)-1, 13))
    c = 0d0
    limPL_incw(ugrid,1,0) = N(4,0);   limPL_incw(ugrid,2,0) = N(4,1)
    limPL_incw(vgrid,1,0) = N(3,0);   limPL_incw(vgrid,2,0) = N(3,1)
    limPL_incw(pgrid,1,0) = N(4,0)+1; limPL_incw(pgrid,2,0) = N(4,1)-1
    call conddown1(c(:,:,0:nb(ugrid)-1,1), c(:,:,0:nb(ugrid)-1,2), u1PL(1,1,limPL_incw(ugrid,1,0)), ugrid, 0)
    call conddown1(c(:,:,0:nb(vgrid)-1,3), c(:,:,0:nb(vgrid)-1,4), u2PL(1,1,limPL_incw(vgrid,1,0)), vgrid, 0)
    call conddown1(c(:,:,0:nb(ugrid)-1,5), c(:,:,0:nb(ugrid)-1,6), u3PL(1,1,limPL_incw(ugrid,1,0)), ugrid, 0)
    call conddown1(c(:,:,0:nb(pgrid)-1,7), c(:,:,0:nb(pgrid)-1,8), ppPL(1,1,limPL_incw(pgrid,1,0)), pgrid, 0)
    call conddown1(c(:,:,0:nb(vgrid)-1,12), c(:,:,0:nb(vgrid)-1,13), wx(1,1,limPL_incw(vgrid,1,0)), vgrid, 0)
    call conddown2(c(:,:,0:nb(ugrid)-1,9), u1PL(1,1,limPL_incw(ugrid,1,0)), &
                   u2PL_itp(1,1,limPL_incw(ugrid,1,0)), ugrid, 0)
    call conddown2(c(:,:,0:nb(ugrid)-1,10), u1PL(1,1,limPL_incw(ugrid,1,0)), &
                   u3PL(1,1,limPL_incw(ugrid,1,0)), ugrid, 0)
    call conddown2(c(:,:,0:nb(vgrid)-1,11), u2PL(1,1,limPL_incw(vgrid,1,0)), &
                   u3PL_itp(1,1,limPL_incw(vgrid,1,0)), vgrid, 0)
    call cond_out(cb_o)
    c = 0d0
    limPL_incw(ugrid,1,0) = N(4,2)+1; limPL_incw(ugrid,2,0) = N(4,3)+1
    limPL_incw(vgrid,1,0) = N(3,2)+1; limPL_incw(vgrid,2,0) = N(3,3)+1
    limPL_incw(pgrid,1,0) = N(4,2)+2; limPL_incw(pgrid,2,0) = N(4,3)
    call condup1(c(:,:,0:nb(ugrid)-1,1), c(:,:,0:nb(ugrid)-1,2), u1PL(1,1,limPL_incw(ugrid,1,0)), 1d0, ugrid, 0)
    call condup1(c(:,:,0:nb(vgrid)-1,3), c(:,:,0:nb(vgrid)-1,4), u2PL(1,1,limPL_incw(vgrid,1,0)), -1d0, vgrid, 0)
    call condup1(c(:,:,0:nb(ugrid)-1,5), c(:,:,0:nb(ugrid)-1,6), u3PL(1,1,limPL_incw(ugrid,1,0)), 1d0, ugrid, 0)
    call condup1(c(:,:,0:nb(pgrid)-1,7), c(:,:,0:nb(pgrid)-1,8), ppPL(1,1,limPL_incw(pgrid,1,0)), 1d0, pgrid, 0)
    call condup1(c(:,:,0:nb(vgrid)-1,12), c(:,:,0:nb(vgrid)-1,13), wx(1,1,limPL_incw(vgrid,1,0)), 1d0, vgrid, 0)
    call condup2(c(:,:,0:nb(ugrid)-1,9), u1PL(1,1,limPL_incw(ugrid,1,0)), &
                 u2PL_itp(1,1,limPL_incw(ugrid,1,0)), 1d0, -1d0, ugrid, 0)
    call condup2(c(:,:,0:nb(ugrid)-1,10), u1PL(1,1,limPL_incw(ugrid,1,0)), &
                 u3PL(1,1,limPL_incw(ugrid,1,0)), 1d0, 1d0, ugrid, 0)
    call condup2(c(:,:,0:nb(vgrid)-1,11), u2PL(1,1,limPL_incw(vgrid,1,0)), &
                 u3PL_itp(1,1,limPL_incw(vgrid,1,0)), -1d0, 1d0, vgrid, 0)
    call cond_out(ct_o)
    deallocate(c)
    deallocate(u1PL, u3PL, u2PL_itp, u2PL, u3PL_itp, wx, ppPL)
  contains
    subroutine row(r, x)
      integer :: r
      real(8) :: x(:)
      st_o((r-1)*nyu+1:(r-1)*nyu+size(x)) = x
    end subroutine
    ! the quantities in write_stats order, each (dnx, dnz, nb(grid of q))
    subroutine cond_out(o)
      real(c_double) :: o(*)
      integer, parameter :: gq(13) = (/ ugrid, ugrid, vgrid, vgrid, ugrid, ugrid, pgrid, &
                                        pgrid, ugrid, ugrid, vgrid, vgrid, vgrid /)
      integer :: n
      a = 0
      do q = 1, 13
        g = gq(q)
        n = dnx*dnz*nb(g)
        o(a+1:a+n) = reshape(c(:,:,0:nb(g)-1,q), (/n/))
        a = a + n
      end do
    end subroutine
  end subroutine

  ! spectra writer recvwrspec (spectra.f90:125-182) for one rank: sp(j, column) over the
  ! jlim range of `grid`, written to the file named by fname (length n)
  subroutine ref_write_spectra(grid, sp_in, istat_in, t_in, iter_in, mpgx_in, flen, fname) &
      bind(C, name="ref_write_spectra")
    integer(c_int), value :: grid, istat_in, iter_in, flen
    real(c_double), value :: t_in, mpgx_in
    real(c_double), intent(in) :: sp_in(*)
    character(kind=c_char), intent(in) :: fname(flen)
    real(8), allocatable :: sp(:,:), buff(:,:,:)
    integer :: status(64), ierr, i, nj, nc
    t = t_in; iter = iter_in; istat = istat_in; mpgx = mpgx_in
    fnameimb = ''
    do i = 1, flen
      fnameimb(i:i) = fname(i)
    end do
    nj = jlim(2,grid,midband) - jlim(1,grid,midband) + 1
    nc = columns_num(midband,0)
    allocate(sp(jlim(1,grid,midband):jlim(2,grid,midband), nc))
    sp = reshape(sp_in(1:nj*nc), (/nj, nc/))
    allocate(buff(0:N(1,midband)/2, 1:N(2,midband)/2+1, jlim(1,grid,midband):jlim(2,grid,midband)))
    call recvwrspec(sp, buff, grid, 0, status, ierr)
    deallocate(sp, buff)
  end subroutine

  ! write_stats (stats.f90:255-416) for one rank owning the bottom-band planes: the plain
  ! rows over those planes, the conditional arrays of that band.  Inputs: rows(13, nyu)
  ! (slot-indexed over the full range of each grid) and the 13 conditional arrays one after
  ! another in the order of write_stats, each (dnx, dnz, nb).
  subroutine ref_write_stats(ntx, ntz, rows, cond, istat_in, t_in, iter_in, mpgx_in, dlen, dirc) &
      bind(C, name="ref_write_stats")
    integer(c_int), value :: ntx, ntz, istat_in, iter_in, dlen
    real(c_double), value :: t_in, mpgx_in
    real(c_double), intent(in) :: rows(*), cond(*)
    character(kind=c_char), intent(in) :: dirc(dlen)
    integer :: status(64), ierr, i, lo(3), hi(3), nyu, a, q, g, nb(3)
    integer, parameter :: gq(13) = (/ ugrid, ugrid, vgrid, vgrid, ugrid, ugrid, pgrid, &
                                      pgrid, ugrid, ugrid, vgrid, vgrid, vgrid /)
    t = t_in; iter = iter_in; istat = istat_in; mpgx = mpgx_in
    dirout = ''
    do i = 1, dlen
      dirout(i:i) = dirc(i)
    end do
    ntilex = ntx; ntilez = ntz
    dnx = Ngal(1,1)/ntilex
    dnz = Ngal(2,1)/ntilez
    write(ext1,'(i4.4)') N(1,1)
    write(ext2,'(i4.4)') N(2,1)
    write(ext3,'(i4.4)') nn+2
    write(ext4,'(i5.5)') int(10d0*t)
    lo(ugrid) = N(4,0);   hi(ugrid) = N(4,1)
    lo(vgrid) = N(3,0);   hi(vgrid) = N(3,1)
    lo(pgrid) = N(4,0)+1; hi(pgrid) = N(4,1)-1
    nb = hi - lo + 1
    nyu = N(4,3) + 1 - N(4,0) + 1
    if (associated(limPL_incw)) deallocate(limPL_incw)
    allocate(limPL_incw(3,2,0:0))
    limPL_incw(:,1,0) = lo
    limPL_incw(:,2,0) = hi
    if (associated(procs)) deallocate(procs)
    allocate(procs(0:nband))
    procs = (/ 0, 0, 1, 1 /)
    call rowin(Um, 1, ugrid);  call rowin(U2m, 2, ugrid)
    call rowin(Vm, 3, vgrid);  call rowin(V2m, 4, vgrid)
    call rowin(Wm, 5, ugrid);  call rowin(W2m, 6, ugrid)
    call rowin(Pm, 7, pgrid);  call rowin(P2m, 8, pgrid)
    call rowin(UVm, 9, ugrid); call rowin(UWm, 10, ugrid); call rowin(VWm, 11, vgrid)
    call rowin(wxm, 12, vgrid); call rowin(wx2m, 13, vgrid)
    a = 0
    call cin(UmC, 1); call cin(U2mC, 2); call cin(VmC, 3); call cin(V2mC, 4)
    call cin(WmC, 5); call cin(W2mC, 6); call cin(PmC, 7); call cin(P2mC, 8)
    call cin(UVmC, 9); call cin(UWmC, 10); call cin(VWmC, 11); call cin(wxmC, 12); call cin(wx2mC, 13)
    call write_stats(0, status, ierr)
  contains
    subroutine rowin(x, r, grid)
      real(8), pointer :: x(:)
      integer :: r, grid, j
      if (associated(x)) deallocate(x)
      allocate(x(lo(grid):hi(grid)))
      do j = lo(grid), hi(grid)
        x(j) = rows((r-1)*nyu + (j - lo(grid)) + 1)
      end do
    end subroutine
    subroutine cin(x, q1)
      real(8), pointer :: x(:,:,:)
      integer :: q1, m
      if (associated(x)) deallocate(x)
      allocate(x(dnx, dnz, 0:nb(gq(q1))-1))
      m = dnx*dnz*nb(gq(q1))
      x = reshape(cond(a+1:a+m), (/dnx, dnz, nb(gq(q1))/))
      a = a + m
    end subroutine
  end subroutine

  subroutine plane_in(x, src, lo, hi)
    real(8), pointer :: x(:,:,:)
    real(c_double), intent(in) :: src(*)
    integer :: lo, hi, j, k, a, gx, gz
    gx = igal - 2
    gz = kgal
    allocate(x(igal, kgal, lo:hi))
    x = 0d0
    a = 0
    do j = lo, hi
      do k = 1, gz
        x(1:gx, k, j) = src(a+1:a+gx)
        a = a + gx
      end do
    end do
  end subroutine

  ! ---- helpers: cfield band arrays holding copies of plain arrays (middle band only)
  subroutine mk(f, grid, inner, a)
    type(cfield) :: f(sband:eband)
    integer :: grid, inner
    complex(c_double_complex), intent(in) :: a(*)
    integer :: b, j0, j1, nj, nc
    do b = sband, eband
      j0 = jlim(1,grid,b) + inner
      j1 = jlim(2,grid,b) - inner
      nc = columns_num(b,0)
      allocate(f(b)%f(j0:j1, nc))
      f(b)%f = (0d0, 0d0)
    end do
    j0 = jlim(1,grid,midband) + inner
    j1 = jlim(2,grid,midband) - inner
    nj = j1 - j0 + 1
    nc = columns_num(midband,0)
    f(midband)%f = reshape(a(1:nj*nc), (/nj, nc/))
  end subroutine

  subroutine mkz(f, grid)
    type(cfield) :: f(sband:eband)
    integer :: grid, b
    do b = sband, eband
      allocate(f(b)%f(jlim(1,grid,b):jlim(2,grid,b), columns_num(b,0)))
      f(b)%f = (0d0, 0d0)
    end do
  end subroutine

  subroutine out(f, a)
    type(cfield) :: f(sband:eband)
    complex(c_double_complex), intent(out) :: a(*)
    integer :: n
    n = size(f(midband)%f)
    a(1:n) = reshape(f(midband)%f, (/n/))
  end subroutine

end module refdrv
