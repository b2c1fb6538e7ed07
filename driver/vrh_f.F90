! vrh_f.F90 -- iso_c_binding interface to libvrh: every function of include/vrh.h.
module vrh_f
  use iso_c_binding
  implicit none

  integer(c_int), parameter :: VRH_FP64 = 0, VRH_MIXED = 1
  integer(c_int), parameter :: VRH_THETA_EXACT_M2 = 0, VRH_THETA_CONSISTENT = 1
  integer(c_int), parameter :: VRH_PART_LSQ_EULER = 0, VRH_PART_LSQ_RK3 = 1, VRH_PART_TRILINEAR_RK3 = 2
  integer(c_int), parameter :: VRH_PH_HDNBC = 2, VRH_PH_EXPLICIT = 3, VRH_PH_ADI = 6, VRH_PH_POISSON = 8, &
                               VRH_PH_PROJECT = 9

  type, bind(C) :: vrh_grid
    integer(c_int) :: ntheta, nr, nz
    real(c_double) :: rmax, zmax
  end type vrh_grid

  type, bind(C) :: vrh_params
    real(c_double) :: dt, Re, slip
    integer(c_int) :: theta_eigen
  end type vrh_params

  interface
    integer(c_int) function vrh_create(grid, params, precision, h) bind(C, name='vrh_create')
      import :: c_int, c_ptr, vrh_grid, vrh_params
      type(vrh_grid), intent(in) :: grid
      type(vrh_params), intent(in) :: params
      integer(c_int), value :: precision
      type(c_ptr), intent(out) :: h
    end function vrh_create

    integer(c_int) function vrh_destroy(h) bind(C, name='vrh_destroy')
      import :: c_int, c_ptr
      type(c_ptr), value :: h
    end function vrh_destroy

    integer(c_int) function vrh_last_error(h, buf, blen) bind(C, name='vrh_last_error')
      import :: c_int, c_ptr, c_char, c_size_t
      type(c_ptr), value :: h
      character(kind=c_char), intent(out) :: buf(*)
      integer(c_size_t), value :: blen
    end function vrh_last_error

    integer(c_int) function vrh_upload_fields(h, q1, q2, q3, pres) bind(C, name='vrh_upload_fields')
      import :: c_int, c_ptr
      type(c_ptr), value :: h, q1, q2, q3, pres
    end function vrh_upload_fields

    integer(c_int) function vrh_download_fields(h, q1, q2, q3, pres) bind(C, name='vrh_download_fields')
      import :: c_int, c_ptr
      type(c_ptr), value :: h, q1, q2, q3, pres
    end function vrh_download_fields

    integer(c_int) function vrh_set_time(h, t) bind(C, name='vrh_set_time')
      import :: c_int, c_ptr, c_double
      type(c_ptr), value :: h
      real(c_double), value :: t
    end function vrh_set_time

    ! rm, rc, dmr, dcr over j = 0..nr+1 (the mesh module's radial arrays)
    integer(c_int) function vrh_set_radial_mesh(h, rm, rc, dmr, dcr) bind(C, name='vrh_set_radial_mesh')
      import :: c_int, c_ptr, c_double
      type(c_ptr), value :: h
      real(c_double), intent(in) :: rm(*), rc(*), dmr(*), dcr(*)
    end function vrh_set_radial_mesh

    ! zm, zc, dmz, dcz over k = 0..nz+1 (the axial counterpart)
    integer(c_int) function vrh_set_axial_mesh(h, zm, zc, dmz, dcz) bind(C, name='vrh_set_axial_mesh')
      import :: c_int, c_ptr, c_double
      type(c_ptr), value :: h
      real(c_double), intent(in) :: zm(*), zc(*), dmz(*), dcz(*)
    end function vrh_set_axial_mesh

    integer(c_int) function vrh_get_time(h, t) bind(C, name='vrh_get_time')
      import :: c_int, c_ptr, c_double
      type(c_ptr), value :: h
      real(c_double), intent(out) :: t
    end function vrh_get_time

    integer(c_int) function vrh_step(h, nsteps) bind(C, name='vrh_step')
      import :: c_int, c_ptr
      type(c_ptr), value :: h
      integer(c_int), value :: nsteps
    end function vrh_step

    integer(c_int) function vrh_synchronize(h) bind(C, name='vrh_synchronize')
      import :: c_int, c_ptr
      type(c_ptr), value :: h
    end function vrh_synchronize

    integer(c_int) function vrh_autotune(h, reps, choice) bind(C, name='vrh_autotune')
      import :: c_int, c_ptr
      type(c_ptr), value :: h
      integer(c_int), value :: reps
      type(c_ptr), value :: choice
    end function vrh_autotune

    integer(c_int) function vrh_ring_init(grid, nrings, rings, pert_amp, m0, m1, phase, q1, q2, q3) &
        bind(C, name='vrh_ring_init')
      import :: c_int, c_ptr, c_double, vrh_grid
      type(vrh_grid), intent(in) :: grid
      integer(c_int), value :: nrings, m0, m1
      real(c_double), intent(in) :: rings(*)
      real(c_double), value :: pert_amp
      type(c_ptr), value :: phase
      type(c_ptr), value :: q1, q2, q3
    end function vrh_ring_init

    integer(c_int) function vrh_cfl_dt(grid, q1, q2, q3, cfl, ujet, dt) bind(C, name='vrh_cfl_dt')
      import :: c_int, c_ptr, c_double, vrh_grid
      type(vrh_grid), intent(in) :: grid
      type(c_ptr), value :: q1, q2, q3
      real(c_double), value :: cfl, ujet
      real(c_double), intent(out) :: dt
    end function vrh_cfl_dt

    integer(c_int) function vrh_particles_create(h, nbatch, cap, st, nsub, scheme) &
        bind(C, name='vrh_particles_create')
      import :: c_int, c_ptr, c_long, c_double
      type(c_ptr), value :: h
      integer(c_int), value :: nbatch, scheme
      integer(c_long), intent(in) :: cap(*)
      real(c_double), intent(in) :: st(*)
      integer(c_int), intent(in) :: nsub(*)
    end function vrh_particles_create

    integer(c_int) function vrh_particles_upload(h, batch, n, u, alive) bind(C, name='vrh_particles_upload')
      import :: c_int, c_ptr, c_long
      type(c_ptr), value :: h, u, alive
      integer(c_int), value :: batch
      integer(c_long), value :: n
    end function vrh_particles_upload

    integer(c_int) function vrh_particles_download(h, batch, n, u, alive) bind(C, name='vrh_particles_download')
      import :: c_int, c_ptr, c_long
      type(c_ptr), value :: h, u, alive
      integer(c_int), value :: batch
      integer(c_long), value :: n
    end function vrh_particles_download
    integer(c_int) function vrh_step_async(h, nsteps) bind(C, name='vrh_step_async')
      import :: c_int, c_ptr
      type(c_ptr), value :: h
      integer(c_int), value :: nsteps
    end function vrh_step_async

    integer(c_int) function vrh_snapshot_begin(h) bind(C, name='vrh_snapshot_begin')
      import :: c_int, c_ptr
      type(c_ptr), value :: h
    end function vrh_snapshot_begin

    integer(c_int) function vrh_snapshot_wait(h, q1, q2, q3, pres, t) bind(C, name='vrh_snapshot_wait')
      import :: c_int, c_ptr, c_double
      type(c_ptr), value :: h, q1, q2, q3, pres
      real(c_double), intent(out) :: t
    end function vrh_snapshot_wait

    integer(c_int) function vrh_analysis(h, out) bind(C, name='vrh_analysis')
      import :: c_int, c_ptr, c_double
      type(c_ptr), value :: h
      real(c_double), intent(out) :: out(10)
    end function vrh_analysis

    ! any named state array (include/vrh.h lists names and sizes)
    integer(c_int) function vrh_upload_state(h, name, host) bind(C, name='vrh_upload_state')
      import :: c_int, c_ptr, c_char
      type(c_ptr), value :: h, host
      character(kind=c_char), intent(in) :: name(*)
    end function vrh_upload_state

    integer(c_int) function vrh_download_state(h, name, host) bind(C, name='vrh_download_state')
      import :: c_int, c_ptr, c_char
      type(c_ptr), value :: h, host
      character(kind=c_char), intent(in) :: name(*)
    end function vrh_download_state

    integer(c_int) function vrh_state_size(grid, name, n) bind(C, name='vrh_state_size')
      import :: c_int, c_long, c_char, vrh_grid
      type(vrh_grid), intent(in) :: grid
      character(kind=c_char), intent(in) :: name(*)
      integer(c_long), intent(out) :: n
    end function vrh_state_size

    ! one substep of the step that has reached the current time (loop.F90:71-127)
    integer(c_int) function vrh_substep(h, irk) bind(C, name='vrh_substep')
      import :: c_int, c_ptr
      type(c_ptr), value :: h
      integer(c_int), value :: irk
    end function vrh_substep

    integer(c_int) function vrh_phase(h, phase, irk) bind(C, name='vrh_phase')
      import :: c_int, c_ptr
      type(c_ptr), value :: h
      integer(c_int), value :: phase, irk
    end function vrh_phase

    integer(c_int) function vrh_set_graph(h, on) bind(C, name='vrh_set_graph')
      import :: c_int, c_ptr
      type(c_ptr), value :: h
      integer(c_int), value :: on
    end function vrh_set_graph

    ! rhs, phi: (ntheta, nr, nz) host arrays (solve_lap_tripar)
    integer(c_int) function vrh_poisson_solve(h, rhs, phi) bind(C, name='vrh_poisson_solve')
      import :: c_int, c_ptr
      type(c_ptr), value :: h, rhs, phi
    end function vrh_poisson_solve

    integer(c_int) function vrh_ktimer_enable(h, on) bind(C, name='vrh_ktimer_enable')
      import :: c_int, c_ptr
      type(c_ptr), value :: h
      integer(c_int), value :: on
    end function vrh_ktimer_enable

    integer(c_int) function vrh_ktimer_read(h, idx, name, nlen, total_ms, cnt) bind(C, name='vrh_ktimer_read')
      import :: c_int, c_ptr, c_size_t
      type(c_ptr), value :: h, name, total_ms, cnt
      integer(c_int), value :: idx
      integer(c_size_t), value :: nlen
    end function vrh_ktimer_read

    integer(c_int) function vrh_set_tuning(h, choice) bind(C, name='vrh_set_tuning')
      import :: c_int, c_ptr
      type(c_ptr), value :: h
      integer(c_int), intent(in) :: choice(9)
    end function vrh_set_tuning

    integer(c_int) function vrh_device_ptr(h, name, ptr) bind(C, name='vrh_device_ptr')
      import :: c_int, c_ptr, c_char
      type(c_ptr), value :: h
      character(kind=c_char), intent(in) :: name(*)
      type(c_ptr), intent(out) :: ptr
    end function vrh_device_ptr

    ! dev_out: device memory of 10 doubles (e.g. from vrh_device_ptr)
    integer(c_int) function vrh_analysis_noalloc(h, dev_out) bind(C, name='vrh_analysis_noalloc')
      import :: c_int, c_ptr
      type(c_ptr), value :: h, dev_out
    end function vrh_analysis_noalloc

    integer(c_int) function vrh_particles_step(h, irk) bind(C, name='vrh_particles_step')
      import :: c_int, c_ptr
      type(c_ptr), value :: h
      integer(c_int), value :: irk
    end function vrh_particles_step

    integer(c_int) function vrh_particles_compact(h, batch) bind(C, name='vrh_particles_compact')
      import :: c_int, c_ptr
      type(c_ptr), value :: h
      integer(c_int), value :: batch
    end function vrh_particles_compact

    integer(c_int) function vrh_particles_sort(h, batch) bind(C, name='vrh_particles_sort')
      import :: c_int, c_ptr
      type(c_ptr), value :: h
      integer(c_int), value :: batch
    end function vrh_particles_sort

    integer(c_int) function vrh_particles_count(h, batch, n) bind(C, name='vrh_particles_count')
      import :: c_int, c_ptr, c_long
      type(c_ptr), value :: h
      integer(c_int), value :: batch
      integer(c_long), intent(out) :: n
    end function vrh_particles_count

    ! ninj: c_loc of an integer(c_long) (the call then waits), or c_null_ptr
    integer(c_int) function vrh_particles_inject(h, batch, r0, inj, seed, ninj) bind(C, name='vrh_particles_inject')
      import :: c_int, c_ptr, c_double, c_long_long
      type(c_ptr), value :: h, ninj
      integer(c_int), value :: batch, inj
      real(c_double), value :: r0
      integer(c_long_long), value :: seed
    end function vrh_particles_inject

    integer(c_int) function vrh_ensemble_create(members, n, e) bind(C, name='vrh_ensemble_create')
      import :: c_int, c_ptr
      type(c_ptr), intent(in) :: members(*)
      integer(c_int), value :: n
      type(c_ptr), intent(out) :: e
    end function vrh_ensemble_create

    integer(c_int) function vrh_ensemble_step(e, nsteps) bind(C, name='vrh_ensemble_step')
      import :: c_int, c_ptr
      type(c_ptr), value :: e
      integer(c_int), value :: nsteps
    end function vrh_ensemble_step

    integer(c_int) function vrh_ensemble_set_batch(e, steps_per_launch) bind(C, name='vrh_ensemble_set_batch')
      import :: c_int, c_ptr
      type(c_ptr), value :: e
      integer(c_int), value :: steps_per_launch
    end function vrh_ensemble_set_batch

    integer(c_int) function vrh_ensemble_last_error(e, buf, blen) bind(C, name='vrh_ensemble_last_error')
      import :: c_int, c_ptr, c_char, c_size_t
      type(c_ptr), value :: e
      character(kind=c_char), intent(out) :: buf(*)
      integer(c_size_t), value :: blen
    end function vrh_ensemble_last_error

    integer(c_int) function vrh_ensemble_destroy(e) bind(C, name='vrh_ensemble_destroy')
      import :: c_int, c_ptr
      type(c_ptr), value :: e
    end function vrh_ensemble_destroy
  end interface

contains

  subroutine vrh_check(h, rc, what)
    type(c_ptr), intent(in) :: h
    integer(c_int), intent(in) :: rc
    character(len=*), intent(in) :: what
    character(kind=c_char) :: buf(512)
    character(len=512) :: msg
    integer :: i, ierr
    if (rc == 0) return
    buf = c_null_char
    ierr = vrh_last_error(h, buf, int(512, c_size_t))
    msg = ' '
    do i = 1, 512
      if (buf(i) == c_null_char) exit
      msg(i:i) = buf(i)
    end do
    write (*, '(a,a,a,i0,a,a)') 'error: ', what, ' failed (', rc, '): ', trim(msg)
    stop 1
  end subroutine vrh_check

end module vrh_f
