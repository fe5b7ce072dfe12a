! lh_host_setup.f90 -- runs the reference host's set-up on one rank and writes what it hands
! to lh_create: MPI start-up as the reference program does it (littleharsh.f90:66-99), then
! start() with input.in from the working directory, then lh_config_from_start.  Needs no GPU.
!
! usage: lh_host_setup config.bin
program lh_host_setup
  use iso_c_binding
  use declaration
  use lh_amd
  use lh_host
  implicit none
  include 'mpif.h'
  integer :: status(MPI_STATUS_SIZE), ierr, myid
  type(lh_config) :: cfg
  character(len=512) :: fout
  call get_command_argument(1, fout)
  call MPI_INIT(ierr)
  call MPI_COMM_RANK(MPI_COMM_WORLD, myid, ierr)
  call MPI_COMM_SIZE(MPI_COMM_WORLD, np, ierr)
  call start(myid, status, ierr)
  call lh_config_from_start(cfg, LH_PREC_FP64)
  call lh_config_write(cfg, trim(fout))
  write(*,'(a,i0,a,i0,a,i0,a,i0)') 'lh_host_setup: nx ', cfg%nx, ' nz ', cfg%nz, ' ny ', cfg%ny, &
      ' canopy points ', cfg%nlist_u
  call MPI_FINALIZE(ierr)
end program
