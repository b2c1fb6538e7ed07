! ref_harness.F90 -- test infrastructure only.
! Drives the reference's own Fortran modules (built from /root/reference by
! oracle/build_ref.sh) through a list of substep phases on a state read from a
! binary file, then writes the resulting state.  Used to check the C oracle
! phase by phase.  Phases: 0 ucoeff, 1 coesys, 2 hdnbc, 3 hdnl1, 4 hdnl2,
! 5 hdnl3, 6 step, 7 divq, 9 upsol, 10 uppres (the Poisson solve, phase 8,
! needs FFTW and is not built).
program ref_harness
  use run_tools
  use cons_tools
  use para_tools
  use comm_tools
  use mesh_tools
  use solu_tools
  use solp_tools
  use momentum_coef
  use momentum_terms
  use inverse
  use divergence
  implicit none
  type(run) :: var
  integer :: ierr, nt, nr, nz, nph, i, ph, irk, hasmesh
  integer, allocatable :: phases(:)
  real(kind=8) :: rmax, zmax, Re, dt, slip, ntime
  real(kind=8), allocatable :: mrm(:), mrc(:), mdmr(:), mdcr(:)
  character(len=512) :: fin, fout

  call get_command_argument(1, fin)
  call get_command_argument(2, fout)
  open(unit=10, file=trim(fin), access='stream', form='unformatted', status='old')
  read(10) nt, nr, nz, nph, hasmesh
  read(10) rmax, zmax, Re, dt, slip, ntime
  allocate(phases(2*nph))
  read(10) phases
  ! optional radial mesh (rm, rc, dmr, dcr over j = 0..nr+1) written into
  ! the mesh module's arrays after create_mesh_a
  allocate(mrm(0:nr+1), mrc(0:nr+1), mdmr(0:nr+1), mdcr(0:nr+1))
  if (hasmesh /= 0) read(10) mrm, mrc, mdmr, mdcr

  open(unit=11, file='inp.txt', form='formatted', status='replace')
  write(11,'(a)') '===Comm==='
  write(11,'(a)') 'dim 3'
  write(11,'(a)') 'npth 1 1 1'
  write(11,'(a)') ''
  write(11,'(a)') '===Mesh==='
  write(11,'(a,3i8)') 'nx ', nt, nr, nz
  write(11,'(a,es26.17e3)') 'maxr ', rmax
  write(11,'(a,es26.17e3)') 'maxz ', zmax
  write(11,'(a)') ''
  write(11,'(a)') '===Flag==='
  write(11,'(a)') 'poisson fftp'
  write(11,'(a)') ''
  write(11,'(a)') '===Para==='
  write(11,'(a,es26.17e3)') 'dt ', dt
  write(11,'(a)') 'totstp 1'
  write(11,'(a)') 'expstp 1'
  write(11,'(a)') 'anlstp 1'
  write(11,'(a)') 'restrt 0'
  write(11,'(a)') 'verbose 0'
  write(11,'(a)') 'u 0.0'
  write(11,'(a)') 'v 0.0'
  write(11,'(a)') 'w 0.0'
  write(11,'(a)') 'w 0.0 0.0 0.0 16.0'
  write(11,'(a)') 'rw.r0.rh 1.0 0.5 5.0'
  write(11,'(a)') 'p 1.0'
  write(11,'(a,es26.17e3)') 'Re ', Re
  write(11,'(a)') 'Pe 1.0'
  write(11,'(a)') 'nrkstep 3'
  write(11,'(a,es26.17e3)') 'slip ', slip
  close(11)

  call mpi_init(ierr)
  call mpi_comm_size(MPI_COMM_WORLD, var%com%np, ierr)
  call mpi_comm_rank(MPI_COMM_WORLD, var%com%ip, ierr)
  var%com%comm_0 = MPI_COMM_WORLD
  var%par%verbose = 0
  call create_cons(var%con)
  call create_comm(var%com, var%con)
  call create_para(var%par, var%com, var%con)
  call create_mesh_a(var%msh, var%par, var%com, var%con)
  if (hasmesh /= 0) then
    var%msh%rm(0:nr+1) = mrm
    var%msh%rc(0:nr+1) = mrc
    var%msh%dmr(0:nr+1) = mdmr
    var%msh%dcr(0:nr+1) = mdcr
  end if
  call alct_solp(var%sp, var%msh)
  call alct_solu(var%su, var%msh, var%com)
  call zero_all()
  var%par%ntime = ntime

  read(10) var%su%q1, var%su%q2, var%su%q3, var%su%dq1, var%su%dq2, var%su%dq3
  read(10) var%su%ru1, var%su%ru2, var%su%ru3, var%su%qrhs, var%su%qcap, var%su%pres
  read(10) var%su%prg, var%sp%rhs, var%sp%dph
  read(10) var%su%q1s, var%su%q2s, var%su%q3s, var%su%dq3no
  read(10) var%su%corr1s, var%su%corr2s, var%su%corr3s, var%su%corr1n, var%su%corr2n, &
           var%su%corr3n, var%su%csort, var%su%tconv
  close(10)

  do i = 1, nph
    ph = phases(2*i-1)
    irk = phases(2*i)
    var%par%irkstep = irk
    select case (ph)
    case (0); call ucoeff(var)
    case (1); call coesys(var)
    case (2); call hdnbc(var)
    case (3); call hdnl1(var)
    case (4); call hdnl2(var)
    case (5); call hdnl3(var)
    case (6); call step(var)
    case (7); call divq(var)
    case (9); call upsol(var)
    case (10); call uppres(var)
    case default
      write(*,*) 'unsupported phase', ph
      stop 2
    end select
  end do

  open(unit=12, file=trim(fout), access='stream', form='unformatted', status='replace')
  write(12) var%su%q1, var%su%q2, var%su%q3, var%su%dq1, var%su%dq2, var%su%dq3
  write(12) var%su%ru1, var%su%ru2, var%su%ru3, var%su%qrhs, var%su%qcap, var%su%pres
  write(12) var%su%prg, var%sp%rhs, var%sp%dph
  write(12) var%su%q1s, var%su%q2s, var%su%q3s, var%su%dq3no
  write(12) var%su%corr1s, var%su%corr2s, var%su%corr3s, var%su%corr1n, var%su%corr2n, &
            var%su%corr3n, var%su%csort, var%su%tconv
  write(12) var%par%ntime, var%par%unom0, var%par%unom1
  write(12) var%su%amj1(0:nr+1), var%su%acj1(0:nr+1), var%su%apj1(0:nr+1)
  write(12) var%su%amj2(0:nr+1), var%su%acj2(0:nr+1), var%su%apj2(0:nr+1)
  write(12) var%su%amj3(0:nr+1), var%su%acj3(0:nr+1), var%su%apj3(0:nr+1)
  write(12) var%su%amk1(0:nz+1), var%su%ack1(0:nz+1), var%su%apk1(0:nz+1)
  write(12) var%su%amk2(0:nz+1), var%su%ack2(0:nz+1), var%su%apk2(0:nz+1)
  write(12) var%su%amk3(0:nz+1), var%su%ack3(0:nz+1), var%su%apk3(0:nz+1)
  close(12)
  call mpi_finalize(ierr)

contains

  subroutine zero_all()
    var%su%u1 = 0.d0; var%su%u2 = 0.d0; var%su%u3 = 0.d0
    var%su%q1 = 0.d0; var%su%q2 = 0.d0; var%su%q3 = 0.d0
    var%su%dq1 = 0.d0; var%su%dq2 = 0.d0; var%su%dq3 = 0.d0
    var%su%ru1 = 0.d0; var%su%ru2 = 0.d0; var%su%ru3 = 0.d0
    var%su%qrhs = 0.d0; var%su%qcap = 0.d0; var%su%pres = 0.d0; var%su%prg = 0.d0
    var%su%q1s = 0.d0; var%su%q2s = 0.d0; var%su%q3s = 0.d0; var%su%dq3no = 0.d0
    var%su%h1d11 = 0.d0; var%su%h1d22d = 0.d0; var%su%h1d22n = 0.d0; var%su%h1dare = 0.d0
    var%su%h2d11 = 0.d0; var%su%h2d22m = 0.d0; var%su%h2d22p = 0.d0; var%su%h2dare = 0.d0
    var%su%s1t11 = 0.d0; var%su%s1t2n = 0.d0; var%su%s1t2s = 0.d0; var%su%s1t2c = 0.d0
    var%su%s2t11 = 0.d0; var%su%s2t2p = 0.d0; var%su%s2t2m = 0.d0; var%su%s2t2c = 0.d0
    var%su%s3t2n = 0.d0; var%su%s3t2s = 0.d0
    var%su%amj1 = 0.d0; var%su%acj1 = 0.d0; var%su%apj1 = 0.d0
    var%su%amj2 = 0.d0; var%su%acj2 = 0.d0; var%su%apj2 = 0.d0
    var%su%amj3 = 0.d0; var%su%acj3 = 0.d0; var%su%apj3 = 0.d0
    var%su%amjs = 0.d0; var%su%acjs = 0.d0; var%su%apjs = 0.d0
    var%su%amk1 = 0.d0; var%su%ack1 = 0.d0; var%su%apk1 = 0.d0
    var%su%amk2 = 0.d0; var%su%ack2 = 0.d0; var%su%apk2 = 0.d0
    var%su%amk3 = 0.d0; var%su%ack3 = 0.d0; var%su%apk3 = 0.d0
    var%su%amks = 0.d0; var%su%acks = 0.d0; var%su%apks = 0.d0
    var%su%atri = 0.d0; var%su%btri = 0.d0; var%su%ctri = 0.d0; var%su%rtri = 0.d0; var%su%stri = 0.d0
    var%su%ami = 0.d0; var%su%aci = 0.d0; var%su%api = 0.d0; var%su%fi = 0.d0; var%su%gam2 = 0.d0
    var%su%corr1s = 0.d0; var%su%corr2s = 0.d0; var%su%corr3s = 0.d0
    var%su%corr1n = 0.d0; var%su%corr2n = 0.d0; var%su%corr3n = 0.d0
    var%su%csort = 0.d0; var%su%tconv = 0.d0
    var%sp%rhs = 0.d0; var%sp%sol = 0.d0; var%sp%ref = 0.d0; var%sp%dph = 0.d0
    var%msh%qc = 0.d0; var%msh%da = 0.d0
  end subroutine zero_all

end program ref_harness
