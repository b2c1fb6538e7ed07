! vrh_f.F90 -- iso_c_binding interface to libvrh (include/vrh.h).
module vrh_f
  use iso_c_binding
  implicit none

  integer(c_int), parameter :: VRH_FP64 = 0, VRH_MIXED = 1
  integer(c_int), parameter :: VRH_THETA_EXACT_M2 = 0, VRH_THETA_CONSISTENT = 1
  integer(c_int), parameter :: VRH_PART_LSQ_EULER = 0, VRH_PART_LSQ_RK3 = 1, VRH_PART_TRILINEAR_RK3 = 2

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
