! substep_check.F90 -- a reference-style RK loop through the C ABI against
! vrh_step.  The reference advances ntime (loop.F90:53), calls iter_part at
! irk == 1 (loop.F90:67) and runs the three substeps (loop.F90:63-131); a
! driver that keeps that loop calls vrh_set_time(t + dt), then
! vrh_particles_step and vrh_substep for irk = 1, 2, 3.  Handle A takes
! nstep vrh_step calls, handle B the same steps as that loop; the fields
! (q1, q2, q3, pres, prg) and two LSQ-Euler particle batches must agree bit
! for bit, in both precisions.  Prints one line per precision and stops
! with status 1 on a difference.
program substep_check
  use iso_c_binding
  use vrh_f
  implicit none
  integer, parameter :: nt = 32, nr = 32, nz = 64, nstep = 3, np = 300, nvar = 117
  type(vrh_grid) :: grid
  type(vrh_params) :: prm
  type(c_ptr) :: ha, hb
  real(c_double), allocatable, target :: q1(:, :, :), q2(:, :, :), q3(:, :, :), fa(:), fb(:)
  real(c_double), allocatable, target :: pu(:, :), pa(:, :), pb(:, :)
  integer(c_int), allocatable, target :: al(:), ala(:), alb(:)
  real(c_double), target :: rings(4), phs(3)
  real(c_double) :: dt, t, th, rr
  integer(c_long) :: cap(2), n
  real(c_double) :: st(2)
  integer(c_int) :: nsub(2), prec, irk, rc
  integer :: i, istep, b, k, ndiff
  character(len=8) :: names(5) = (/'q1      ', 'q2      ', 'q3      ', 'pres    ', 'prg     '/)

  grid = vrh_grid(nt, nr, nz, 4.d0, 12.d0)
  rings = (/1.d0, 3.d0, 0.3d0, 1.d0/)
  phs = (/0.3d0, 1.1d0, 2.0d0/)
  allocate(q1(nt, 0:nr+2, 0:nz+2), q2(nt, 0:nr+2, 0:nz+2), q3(nt, 0:nr+2, 0:nz+2))
  rc = vrh_ring_init(grid, 1, rings, 2.d-2, 2, 4, c_loc(phs), c_loc(q1), c_loc(q2), c_loc(q3))
  if (rc /= 0) stop 'vrh_ring_init failed'
  rc = vrh_cfl_dt(grid, c_loc(q1), c_loc(q2), c_loc(q3), 0.3d0, 1.d0, dt)
  ! particles in the ring's torus, at rest, identity Jacobian (deterministic)
  allocate(pu(nvar, np), pa(nvar, np), pb(nvar, np), al(np), ala(np), alb(np))
  pu = 0.d0
  do i = 1, np
    th = 2.d0 * acos(-1.d0) * (i - 0.5d0) / np
    rr = 1.d0 + 0.4d0 * cos(7.d0 * th)
    pu(1, i) = rr * cos(th); pu(2, i) = rr * sin(th); pu(3, i) = 3.d0 + 0.4d0 * sin(5.d0 * th)
    pu(7, i) = 1.d0; pu(11, i) = 1.d0; pu(15, i) = 1.d0
  end do
  al = 1
  cap = (/int(np, c_long), int(np, c_long)/)
  st = (/0.05d0, 0.5d0/)
  nsub = (/20, 5/)
  ndiff = 0
  do prec = VRH_FP64, VRH_MIXED
    prm = vrh_params(dt, 1000.d0, 0.d0, VRH_THETA_EXACT_M2)
    call vrh_check(ha, vrh_create(grid, prm, prec, ha), 'create A')
    call vrh_check(hb, vrh_create(grid, prm, prec, hb), 'create B')
    t = 0.15d0
    call vrh_check(ha, vrh_upload_fields(ha, c_loc(q1), c_loc(q2), c_loc(q3), c_null_ptr), 'upload A')
    call vrh_check(hb, vrh_upload_fields(hb, c_loc(q1), c_loc(q2), c_loc(q3), c_null_ptr), 'upload B')
    call vrh_check(ha, vrh_set_time(ha, t), 'time A')
    call vrh_check(ha, vrh_particles_create(ha, 2, cap, st, nsub, VRH_PART_LSQ_EULER), 'particles A')
    call vrh_check(hb, vrh_particles_create(hb, 2, cap, st, nsub, VRH_PART_LSQ_EULER), 'particles B')
    do b = 0, 1
      call vrh_check(ha, vrh_particles_upload(ha, b, int(np, c_long), c_loc(pu), c_loc(al)), 'upload pA')
      call vrh_check(hb, vrh_particles_upload(hb, b, int(np, c_long), c_loc(pu), c_loc(al)), 'upload pB')
    end do
    call vrh_check(ha, vrh_step(ha, nstep), 'step A')
    do istep = 1, nstep
      t = dt + t                                              ! loop.F90:53
      call vrh_check(hb, vrh_set_time(hb, t), 'time B')
      do irk = 1, 3
        call vrh_check(hb, vrh_particles_step(hb, irk), 'particles B')   ! loop.F90:66-70
        call vrh_check(hb, vrh_substep(hb, irk), 'substep B')             ! loop.F90:71-127
      end do
    end do
    call vrh_check(hb, vrh_synchronize(hb), 'sync B')
    do k = 1, 5
      call vrh_check(ha, vrh_state_size(grid, trim(names(k)) // c_null_char, n), 'size')
      allocate(fa(n), fb(n))
      call vrh_check(ha, vrh_download_state(ha, trim(names(k)) // c_null_char, c_loc(fa)), 'download A')
      call vrh_check(hb, vrh_download_state(hb, trim(names(k)) // c_null_char, c_loc(fb)), 'download B')
      if (any(fa /= fb)) then
        ndiff = ndiff + 1
        write(*, '(a,i2,2a,es12.4)') 'precision ', prec, ': differs in ', trim(names(k)), maxval(abs(fa - fb))
      end if
      deallocate(fa, fb)
    end do
    do b = 0, 1
      call vrh_check(ha, vrh_particles_download(ha, b, int(np, c_long), c_loc(pa), c_loc(ala)), 'particles A')
      call vrh_check(hb, vrh_particles_download(hb, b, int(np, c_long), c_loc(pb), c_loc(alb)), 'particles B')
      if (any(pa /= pb) .or. any(ala /= alb)) then
        ndiff = ndiff + 1
        write(*, '(a,i2,a,i2)') 'precision ', prec, ': particles differ in batch ', b
      end if
      if (all(pa(1:3, :) == pu(1:3, :))) then
        ndiff = ndiff + 1
        write(*, '(a)') 'particles did not move'
      end if
    end do
    call vrh_check(ha, vrh_get_time(ha, th), 'time A')
    if (th /= t) then
      ndiff = ndiff + 1
      write(*, '(a,2es24.16)') 'times differ ', th, t
    end if
    rc = vrh_destroy(ha)
    rc = vrh_destroy(hb)
    write(*, '(a,i2,a,i3,a)') 'precision ', prec, ': ', nstep, ' vrh_step calls against the substep loop'
  end do
  if (ndiff /= 0) then
    write(*, '(a,i3,a)') 'substep check FAILED (', ndiff, ' differences)'
    stop 1
  end if
  write(*, '(a)') 'substep check: bit-identical'
end program substep_check
