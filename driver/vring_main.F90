! vring_main.F90 -- the vring driver: reads inp.txt, sets up the state, runs
! the time loop of loop.F90:47-153 with the RK3 substeps on the GPU through
! libvrh, and writes restart, field, analysis and particle files in the
! reference's formats.
program vring_main
  use iso_c_binding
  use vrh_f
  use vring_io
  implicit none
  type(deck) :: d
  type(vrh_grid), target :: grid
  type(vrh_params) :: params
  type(c_ptr) :: h
  real(c_double), allocatable, target :: q1(:, :, :), q2(:, :, :), q3(:, :, :), pres(:, :, :)
  real(c_double), allocatable, target :: pu(:, :), phs(:)
  integer(c_int), allocatable, target :: palive(:)
  integer(c_int), target :: choice(9)
  real(c_double) :: ntime, stat(20), dtcfl, st, astat(10), tsnap
  logical :: pending, panl, export, anl
  integer :: pstep
  real(c_double) :: t0, t1, t2, t3
  integer :: nt, nr, nz, istep, nstep, nstepi, rc, np, ip
  integer(c_long) :: capl(1)
  integer(c_int) :: nsubv(1)
  real(c_double) :: stv(1), x(3), rr
  integer, allocatable :: seedv(:)
  integer :: nseed
  character(len=32) :: arg
  ! inlet batches: droplet radii in micrometres and Euler sub-steps of the
  ! reference's eight batches (part_tools.F90:262-269, 493-502)
  real, parameter :: rad_um(8) = (/0.5, 1.0, 2.0, 5.0, 10.0, 20.0, 50.0, 100.0/)
  integer, parameter :: nsub_ref(8) = (/20, 10, 10, 5, 5, 5, 5, 5/)
  integer :: nb, b, nvarp
  integer(c_long), allocatable :: capb(:)
  real(c_double), allocatable :: stb(:), radb(:)
  integer(c_int), allocatable :: nsubb(:)
  integer(c_long) :: nlive
  real :: rs
  real(c_double) :: r8
  real(c_double), allocatable, target :: m1(:), m2(:), m3(:), m4(:)

  call read_deck('inp.txt', d)
  nt = d%nthetag; nr = d%nrg; nz = d%nzg
  arg = ' '
  if (command_argument_count() >= 1) call get_command_argument(1, arg)
  if (trim(arg) == '--echo') then   ! parse and echo the deck only (no GPU)
    call echo_deck(d)
    write(*, '(a,3i8,2es16.8)') 'mesh ', nt, nr, nz, d%rmax, d%zmax
    write(*, '(a,es16.8,4i8)') 'para ', d%dt, d%totstp, d%expstp, d%anlstp, d%restrt
    write(*, '(a,3es16.8,i4)') 'flow ', d%Re, d%Pe, d%slip, d%nrkstep
    write(*, '(a,i4,4es16.8)') 'ring ', d%nring, d%rings(:, 1)
    write(*, '(a,3i4,2f10.4)') 'gpu  ', d%precision, d%theta_eigen, d%autotune, d%rstretch, d%zstretch
    stop
  end if
  if (d%npth(1) * d%npth(2) * d%npth(3) /= 1) &
    write(*, '(a)') 'note: npth decomposition collapsed to one whole-domain layout on one GPU'
  if (d%nrkstep /= 3) stop 'nrkstep must be 3 (loop.F90 runs three substeps)'
  if (d%poisson /= 'fftp') stop 'only poisson fftp is available'
  call echo_deck(d)
  allocate(q1(nt, 0:nr+2, 0:nz+2), q2(nt, 0:nr+2, 0:nz+2), q3(nt, 0:nr+2, 0:nz+2), pres(nt, 0:nr+2, 0:nz+2))
  q1 = 0.d0; q2 = 0.d0; q3 = 0.d0; pres = 0.d0
  nstepi = 0
  ntime = 0.d0
  grid = vrh_grid(nt, nr, nz, d%rmax, d%zmax)
  if (d%restrt == 1) then
    call restart_load(nstepi, ntime, nt, nr, nz, q1, q2, q3)
  else if (d%nring > 0) then
    allocate(phs(max(1, d%pert_m1 - d%pert_m0 + 1)))
    phs = d%phase(1:size(phs))
    rc = vrh_ring_init(grid, d%nring, d%rings, d%pert_amp, d%pert_m0, d%pert_m1, c_loc(phs), c_loc(q1), &
                       c_loc(q2), c_loc(q3))
    if (rc /= 0) stop 'vrh_ring_init failed'
  end if
  ! a CFL number fixes dt from the initial ring; a restart keeps the deck's dt
  ! (the reference reads dt from inp.txt, para_tools.F90:68-86)
  if (d%cfl > 0.d0 .and. d%restrt /= 1) then
    rc = vrh_cfl_dt(grid, c_loc(q1), c_loc(q2), c_loc(q3), d%cfl, d%ujet, dtcfl)
    d%dt = dtcfl
    write(*, '(a,es16.8)') 'dt from CFL: ', d%dt
  end if
  params = vrh_params(d%dt, d%Re, d%slip, d%theta_eigen)
  call vrh_check(h, vrh_create(grid, params, d%precision, h), 'vrh_create')
  call vrh_check(h, vrh_upload_fields(h, c_loc(q1), c_loc(q2), c_loc(q3), c_loc(pres)), 'upload')
  call vrh_check(h, vrh_set_time(h, ntime), 'set_time')
  ! stretched meshes (the library's hooks; particles need uniform meshes)
  if (d%rstretch > 0.d0 .or. d%zstretch > 0.d0) then
    if (d%npart > 0 .or. d%pinlet > 0) stop 'particles need the uniform meshes (mesh 0 0)'
  end if
  if (d%rstretch > 0.d0) then
    allocate(m1(0:nr+1), m2(0:nr+1), m3(0:nr+1), m4(0:nr+1), mesh_rm(0:nr+1))
    call sinh_radial_mesh(nr, d%rmax, d%rstretch, m1, m2, m3, m4)
    call vrh_check(h, vrh_set_radial_mesh(h, m1, m2, m3, m4), 'radial mesh')
    mesh_rm = m1
    deallocate(m1, m2, m3, m4)
    write(*, '(a,f8.4)') 'radial mesh: sinh stretch ', d%rstretch
  end if
  if (d%zstretch > 0.d0) then
    allocate(m1(0:nz+1), m2(0:nz+1), m3(0:nz+1), m4(0:nz+1), mesh_zm(0:nz+1))
    call tanh_axial_mesh(nz, d%zmax, d%zstretch, m1, m2, m3, m4)
    call vrh_check(h, vrh_set_axial_mesh(h, m1, m2, m3, m4), 'axial mesh')
    mesh_zm = m1
    deallocate(m1, m2, m3, m4)
    write(*, '(a,f8.4)') 'axial mesh: tanh stretch ', d%zstretch
  end if
  if (d%npart > 0) then
    np = d%npart
    allocate(pu(merge(6, 117, d%pscheme == VRH_PART_TRILINEAR_RK3), np), palive(np))
    pu = 0.d0
    palive = 1
    call random_seed(size=nseed)
    allocate(seedv(nseed))
    seedv = 12345 + d%pseed
    call random_seed(put=seedv)
    ip = 0
    do while (ip < np)
      call random_number(x)
      x(1:2) = (2.d0 * x(1:2) - 1.d0) * (d%ptorus(1) + d%ptorus(3))
      x(3) = d%ptorus(2) + (2.d0 * x(3) - 1.d0) * d%ptorus(3)
      rr = sqrt(x(1)**2 + x(2)**2)
      if ((rr - d%ptorus(1))**2 + (x(3) - d%ptorus(2))**2 <= d%ptorus(3)**2) then
        ip = ip + 1
        pu(1:3, ip) = x
        if (size(pu, 1) == 117) then
          pu(7, ip) = 1.d0; pu(11, ip) = 1.d0; pu(15, ip) = 1.d0
        end if
      end if
    end do
    ! Stokes number as part_tools.F90:273 (default-real constants)
    st = real(1.0 / 18.0, c_double) * real(998.0 / 1.2, c_double) * (d%pradius / 0.02d0)**2 * d%Re
    capl(1) = np; stv(1) = st; nsubv(1) = d%nsub
    call vrh_check(h, vrh_particles_create(h, 1, capl, stv, nsubv, d%pscheme), 'particles_create')
    call vrh_check(h, vrh_particles_upload(h, 0, int(np, c_long), c_loc(pu), c_loc(palive)), 'particles_upload')
  end if
  nb = 0
  if (d%pinlet > 0) then
    if (d%npart > 0) stop 'inlet batches and torus particles are exclusive'
    nb = min(d%pinlet, 8)
    nvarp = merge(6, 117, d%pscheme == VRH_PART_TRILINEAR_RK3)
    allocate(capb(nb), stb(nb), radb(nb), nsubb(nb), pu(nvarp, d%pcap), palive(d%pcap))
    do b = 1, nb
      ! prt%r and prt%st as init_part evaluates them (part_tools.F90:262-273):
      ! the radius in default real, the Stokes number's constants too
      rs = rad_um(b) * (10.0**(-6.0)) / 0.02
      r8 = rs
      radb(b) = r8
      stb(b) = (1.0 / 18.0) * (998.0 / 1.2) * (r8**2) * d%Re
      capb(b) = d%pcap
      nsubb(b) = nsub_ref(b)
    end do
    call vrh_check(h, vrh_particles_create(h, nb, capb, stb, nsubb, d%pscheme), 'particles_create')
    write(*, '(a,i2,a,i8,a,i4,a)') 'inlet particles: ', nb, ' batches of ', d%pcap, &
        ' slots, injection every ', d%pinj, ' steps'
  end if
  if (d%autotune /= 0) then
    call vrh_check(h, vrh_autotune(h, 3, c_loc(choice)), 'autotune')
    write(*, '(a,9i5)') 'autotune choice: ', choice
  end if
  stat = 0.d0
  pending = .false.
  call cpu_time(t0)
  do istep = 0, d%totstp
    nstep = istep + nstepi
    call cpu_time(t1)
    ! inlet injection of every batch (loop.F90:54-59: mod(nstep, inj) == 0,
    ! from the inlet velocity before the step's substeps); positions from
    ! the library's counter-based stream, seeded per batch
    if (nb > 0) then
      if (mod(nstep, d%pinj) == 0) then
        do b = 1, nb
          call vrh_check(h, vrh_particles_inject(h, b - 1, d%r0, d%pinj, int(8 * d%pseed + b, c_long_long), &
                                                 c_null_ptr), 'inject')
        end do
      end if
    end if
    ! the step runs on the GPU while the host writes the files of the last
    ! field export (taken as a device snapshot at that step)
    call vrh_check(h, vrh_step_async(h, 1), 'step')
    if (pending) call flush_export()
    call vrh_check(h, vrh_synchronize(h), 'step')
    call cpu_time(t2)
    call vrh_check(h, vrh_get_time(h, ntime), 'get_time')
    export = .false.
    if (d%expstp > 0) export = mod(nstep, d%expstp) == 0
    anl = .false.
    if (d%anlstp > 0) anl = mod(nstep, d%anlstp) == 0
    if (export) then
      ! restart file, and (on analysis steps) the host analysis pass with the
      ! Tecplot files, from the snapshot during the next step
      call vrh_check(h, vrh_snapshot_begin(h), 'snapshot')
      pending = .true.
      pstep = nstep
      panl = anl
      if (d%npart > 0) then
        call vrh_check(h, vrh_particles_download(h, 0, int(np, c_long), c_loc(pu), c_loc(palive)), 'particles')
        call part_export(nstep, 1, d%pradius / 0.02d0, st, np, size(pu, 1), pu, palive)
      end if
      do b = 1, nb   ! partNNNNNsizeBB.dat per batch (part_tools.F90:444-467)
        call vrh_check(h, vrh_particles_compact(h, b - 1), 'compact')
        call vrh_check(h, vrh_particles_count(h, b - 1, nlive), 'count')
        if (nlive > 0) call vrh_check(h, vrh_particles_download(h, b - 1, nlive, c_loc(pu), c_loc(palive)), &
                                      'particles')
        call part_export(nstep, b, radb(b), stb(b), int(nlive), nvarp, pu, palive)
      end do
    else if (anl) then
      ! statistics only: reduced on the device, ten numbers come back
      call vrh_check(h, vrh_analysis(h, astat), 'analysis')
      call analysis_record(d, nstep, ntime, astat, stat)
    end if
    call cpu_time(t3)
    ! loop.F90:144-151 (times in minutes)
    write(*, "(2a)") "__________________________________________"
    write(*, "(2a)") "::Step________::Time________::Residual____::Performance_::Partitioning"
    write(*, "(a,I10,3(a,F10.6),a,I10)") " I: ", nstepi, " I: ", 0.d0, " M1:", 0.d0, " TM:", (t2 - t1) / 60.d0, " TN:", 1
    write(*, "(a,I10,3(a,F10.6),a,I10)") " R: ", nstep, " R: ", ntime, " M2:", 0.d0, " TP:", (t3 - t2) / 60.d0, " TT:", 1
    write(*, "(a,I10,3(a,F10.6),a,I10)") " F: ", nstepi + d%totstp, " F: ", d%totstp * d%dt, " P1:", 0.d0, &
        " TS:", (t3 - t1) / 60.d0, " TR:", 1
    write(*, "(a,I10,3(a,F10.6),a,I10)") " E: ", d%expstp, " Dt:", d%dt, " P2:", 0.d0, " TT:", &
        (t3 - t0) / 60.d0 / real(max(istep, 1), c_double), " TZ:", 1
  end do
  if (pending) call flush_export()
  call vrh_check(h, vrh_download_fields(h, c_loc(q1), c_loc(q2), c_loc(q3), c_loc(pres)), 'download')
  rc = vrh_destroy(h)

contains

  ! files of the pending field export: restart_save (run.F90:390-466), then
  ! the analysis pass that writes the Tecplot files (run.F90:238-388)
  subroutine flush_export()
    call vrh_check(h, vrh_snapshot_wait(h, c_loc(q1), c_loc(q2), c_loc(q3), c_loc(pres), tsnap), 'snapshot')
    call restart_save(pstep, 3, tsnap, nt, nr, nz, q1, q2, q3)
    if (panl) call analysis(d, pstep, 3, tsnap, nt, nr, nz, q1, q2, q3, pres, stat)
    pending = .false.
  end subroutine flush_export

end program vring_main
