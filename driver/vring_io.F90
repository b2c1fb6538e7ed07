! vring_io.F90 -- input deck, restart files and Tecplot/analysis output of the
! vring driver, in the reference's file formats:
!   inp.txt sections ===Comm===, ===Mesh===, ===Flag===, ===Para===
!     (comm_tools.F90:36-47, mesh_tools.F90:84-93, para_tools.F90:66-96)
!   out.txt echo (run.F90:29-69)
!   restartNNNNNRKkk.dat / RESTART (run.F90:390-539)
!   2DfldNNNNNRKkk.dat, 3DfldNNNNNRKkkPR00.dat, analysis.dat (run.F90:238-388)
!   partNNNNNsizeBB.dat (part_tools.F90:444-467)
! Optional sections of this driver (ignored by the reference): ===Ring===
! (Gaussian-ring initial condition, CFL time step), ===GPU=== (precision,
! Poisson theta eigenvalue, autotuning, and optional stretched meshes:
! `mesh rs zs`, a sinh radial mesh of stretch rs and a tanh axial mesh of
! stretch zs, 0 for uniform), ===Part=== (particles seeded in a torus, or
! the reference's inlet batches).
module vring_io
  use iso_c_binding
  implicit none

  type :: deck
    integer :: npth(3) = 1
    integer :: nthetag, nrg, nzg
    real(c_double) :: rmax, zmax
    character(len=4) :: poisson = 'fftp'
    real(c_double) :: dt, u1, u2, u3, uring, uswrl, ufluc, kfluc, rw, r0, t0, p, Re, Pe, slip
    integer :: totstp, expstp, anlstp, restrt, verbose, nrkstep
    ! ===Ring===
    integer :: nring = 0
    real(c_double) :: rings(4, 8)
    real(c_double) :: pert_amp = 0.d0
    integer :: pert_m0 = 0, pert_m1 = -1
    real(c_double) :: phase(64) = 0.d0
    real(c_double) :: cfl = 0.d0, ujet = 1.d0
    ! ===GPU===
    integer :: precision = 0, theta_eigen = 0, autotune = 0
    real(c_double) :: rstretch = 0.d0, zstretch = 0.d0
    ! ===Part===
    integer :: npart = 0, nsub = 5, pscheme = 0, pseed = 0
    real(c_double) :: ptorus(3) = (/1.d0, 3.d0, 0.6d0/), pradius = 5.0d-6
    ! inlet batches (loop.F90:39-59): the first pinlet of the reference's
    ! eight droplet sizes, pcap particles each, injected every pinj steps
    integer :: pinlet = 0, pinj = 10
    integer(c_long) :: pcap = 50000
  end type deck

  ! cell centres of a stretched mesh (set by the driver after
  ! vrh_set_radial_mesh / vrh_set_axial_mesh); unallocated: uniform
  real(c_double), allocatable :: mesh_rm(:), mesh_zm(:)

contains

  ! Stretched meshes as the library takes them (over index 0..n+1): radial
  ! r(xi) = R sinh(s xi/R)/sinh(s) with centres at xi = (j - 1/2) R/nr and
  ! faces at (j - 1) R/nr (rc(1) = 0), dr/dxi at both; axial faces
  ! z_k = Z (1 + tanh(s (k/nz - 1))/tanh(s)), centres midway, mirrored
  ! ghost cells, dz/dzeta as widths and centre spacings over Z/nz.
  subroutine sinh_radial_mesh(nr, rmax, s, rm, rc, dmr, dcr)
    integer, intent(in) :: nr
    real(c_double), intent(in) :: rmax, s
    real(c_double), intent(out) :: rm(0:nr+1), rc(0:nr+1), dmr(0:nr+1), dcr(0:nr+1)
    integer :: j
    real(c_double) :: dxi, xm, xc
    dxi = rmax / nr
    do j = 0, nr + 1
      xm = (j - 0.5d0) * dxi
      xc = (j - 1.d0) * dxi
      rm(j) = rmax * sinh(s * xm / rmax) / sinh(s)
      rc(j) = rmax * sinh(s * xc / rmax) / sinh(s)
      dmr(j) = s * cosh(s * xm / rmax) / sinh(s)
      dcr(j) = s * cosh(s * xc / rmax) / sinh(s)
    end do
    rc(1) = 0.d0
  end subroutine sinh_radial_mesh

  subroutine tanh_axial_mesh(nz, zmax, s, zm, zc, dmz, dcz)
    integer, intent(in) :: nz
    real(c_double), intent(in) :: zmax, s
    real(c_double), intent(out) :: zm(0:nz+1), zc(0:nz+1), dmz(0:nz+1), dcz(0:nz+1)
    real(c_double) :: zf(0:nz), dz
    integer :: k
    do k = 0, nz
      zf(k) = zmax * (1.d0 + tanh(s * (dble(k) / nz - 1.d0)) / tanh(s))
    end do
    zf(0) = 0.d0
    zf(nz) = zmax
    zc(1:nz+1) = zf(0:nz)
    zc(0) = -zc(2)
    do k = 1, nz
      zm(k) = 0.5d0 * (zf(k) + zf(k - 1))
    end do
    zm(0) = -zm(1)
    zm(nz + 1) = 2.d0 * zmax - zm(nz)
    dz = zmax / nz
    do k = 1, nz
      dmz(k) = (zf(k) - zf(k - 1)) / dz
    end do
    dmz(0) = dmz(1)
    dmz(nz + 1) = dmz(nz)
    do k = 1, nz + 1
      dcz(k) = (zm(k) - zm(k - 1)) / dz
    end do
    dcz(0) = dcz(1)
  end subroutine tanh_axial_mesh

  logical function find_section(u, name)
    integer, intent(in) :: u
    character(len=*), intent(in) :: name
    character(len=128) :: line
    integer :: ios
    rewind(u)
    find_section = .false.
    do
      read(u, '(a)', iostat=ios) line
      if (ios /= 0) return
      if (trim(adjustl(line)) == trim(adjustl(name))) then
        find_section = .true.
        return
      end if
    end do
  end function find_section

  subroutine read_deck(fname, d)
    character(len=*), intent(in) :: fname
    type(deck), intent(inout) :: d
    integer :: u, i, ios
    character(len=16) :: a
    open(newunit=u, file=fname, form='formatted', status='old')
    if (find_section(u, '===Comm===')) then
      read(u, *) a
      read(u, *) a, d%npth
    end if
    if (.not. find_section(u, '===Mesh===')) stop 'inp.txt: no ===Mesh=== section'
    read(u, *) a, d%nthetag, d%nrg, d%nzg
    read(u, *) a, d%rmax
    read(u, *) a, d%zmax
    if (find_section(u, '===Flag===')) read(u, *) a, d%poisson
    if (.not. find_section(u, '===Para===')) stop 'inp.txt: no ===Para=== section'
    read(u, *) a, d%dt
    read(u, *) a, d%totstp
    read(u, *) a, d%expstp
    read(u, *) a, d%anlstp
    read(u, *) a, d%restrt
    read(u, *) a, d%verbose
    read(u, *) a, d%u1
    read(u, *) a, d%u2
    read(u, *) a, d%u3
    read(u, *) a, d%uring, d%uswrl, d%ufluc, d%kfluc
    read(u, *) a, d%rw, d%r0, d%t0
    read(u, *) a, d%p
    read(u, *) a, d%Re
    read(u, *) a, d%Pe
    read(u, *) a, d%nrkstep
    read(u, *) a, d%slip
    if (find_section(u, '===Ring===')) then
      read(u, *) a, d%nring
      do i = 1, d%nring
        read(u, *) a, d%rings(:, i)
      end do
      read(u, *, iostat=ios) a, d%pert_amp, d%pert_m0, d%pert_m1
      if (ios == 0 .and. d%pert_m1 >= d%pert_m0) read(u, *) a, d%phase(1:d%pert_m1 - d%pert_m0 + 1)
      read(u, *, iostat=ios) a, d%cfl, d%ujet
      if (ios /= 0) d%cfl = 0.d0
    end if
    if (find_section(u, '===GPU===')) then
      read(u, *, iostat=ios) a, d%precision
      read(u, *, iostat=ios) a, d%theta_eigen
      read(u, *, iostat=ios) a, d%autotune
      read(u, *, iostat=ios) a, d%rstretch, d%zstretch
      if (ios /= 0) then
        d%rstretch = 0.d0; d%zstretch = 0.d0
      end if
    end if
    if (find_section(u, '===Part===')) then
      read(u, *) a, d%npart
      read(u, *) a, d%ptorus
      read(u, *) a, d%pradius
      read(u, *) a, d%nsub
      read(u, *) a, d%pscheme
      read(u, *, iostat=ios) a, d%pseed
      read(u, *, iostat=ios) a, d%pinlet, d%pcap
      if (ios /= 0) d%pinlet = 0
    end if
    close(u)
  end subroutine read_deck

  ! out.txt, run.F90:29-69 (one domain: local sizes equal global sizes)
  subroutine echo_deck(d)
    type(deck), intent(in) :: d
    integer :: u
    real(c_double) :: dr, dz, dth
    dth = 2.d0 * acos(-1.d0) / d%nthetag
    dr = d%rmax / d%nrg
    dz = d%zmax / d%nzg
    open(newunit=u, file='out.txt', form='formatted', status='replace')
    write(u, "(A,3(I16.8))") 'Domain decomposition__(comm)_'
    write(u, "(A,3(I16.8))") 'n. processors (th,r,z):', 1, 1, 1
    write(u, "(A,3(I16.8))") 'Mesh characteristics__(mesh)_'
    write(u, "(A,3(I16.8))") 'nth,nr,nz   local     :', d%nthetag, d%nrg, d%nzg
    write(u, "(A,3(I16.8))") 'nth,nr,nz   global    :', d%nthetag, d%nrg, d%nzg
    write(u, "(A,3(F16.8))") 'dth,dr,dz             :', dth, dr, dz
    write(u, "(A,3(F16.8))") 'rmax                  :', d%rmax, 0.5d0 * dr, (d%nrg - 0.5d0) * dr
    write(u, "(A,3(F16.8))") 'zmax                  :', d%zmax, 0.5d0 * dz, (d%nzg - 0.5d0) * dz
    write(u, "(A,3(I16.8))") 'Problem constants_____(cons)_'
    write(u, "(A,3(I16.8))") 'solver                :'
    write(u, "(A,3(F16.8))") 'Problem parameters____(para)_'
    write(u, "(A,3(I16.8))") 'total steps           :', d%totstp
    write(u, "(A,3(I16.8))") 'export every          :', d%expstp
    write(u, "(A,3(I16.8))") 'analys every          :', d%anlstp
    write(u, "(A,3(I16.8))") 'restart               :', d%restrt
    write(u, "(A,3(F16.8))") 'u1,v2,w3    nominal   :', d%u1, d%u2, d%u3
    write(u, "(A,4(F16.8))") 'ur,wr,uf,kf nominal   :', d%uring, d%uswrl, d%ufluc, d%kfluc
    write(u, "(A,4(F16.8))") 'rw,r0,t0              :', d%rw, d%r0, d%t0
    write(u, "(A,3(F16.8))") 'p           nominal   :', d%p
    write(u, "(A,3(F16.8))") 'Re                    :', d%Re
    write(u, "(A,3(F16.8))") 'Pe                    :', d%Pe
    write(u, "(A,3(F16.8))") 'dt                    :', d%dt
    write(u, "(A,3(I16.8))") 'rkstep                :', d%nrkstep
    write(u, "(A,3(F16.8))") 'slip                  :', d%slip
    close(u)
  end subroutine echo_deck

  ! restart_save, run.F90:390-466: header record, then one record
  ! (q1,q2,q3) per theta point, iz outer, ir middle, ith inner
  subroutine restart_save(nstep, irk, ntime, nt, nr, nz, q1, q2, q3)
    integer, intent(in) :: nstep, irk, nt, nr, nz
    real(c_double), intent(in) :: ntime
    real(c_double), intent(in) :: q1(nt, 0:nr+2, 0:nz+2), q2(nt, 0:nr+2, 0:nz+2), q3(nt, 0:nr+2, 0:nz+2)
    character(len=5) :: tc
    character(len=2) :: rc
    integer :: u, iz, ir, ith
    write(tc, '(I5.5)') nstep
    write(rc, '(I2.2)') irk
    open(newunit=u, file='restart'//tc//'RK'//rc//'.dat', form='unformatted', status='replace')
    write(u) nstep, ntime, nt, nr, nz
    do iz = 1, nz
      do ir = 1, nr
        do ith = 1, nt
          write(u) q1(ith, ir, iz), q2(ith, ir, iz), q3(ith, ir, iz)
        end do
      end do
    end do
    close(u)
  end subroutine restart_save

  ! restart_load, run.F90:469-539 (reads ./RESTART)
  subroutine restart_load(nstepi, ntime, nt, nr, nz, q1, q2, q3)
    integer, intent(out) :: nstepi
    real(c_double), intent(out) :: ntime
    integer, intent(in) :: nt, nr, nz
    real(c_double), intent(inout) :: q1(nt, 0:nr+2, 0:nz+2), q2(nt, 0:nr+2, 0:nz+2), q3(nt, 0:nr+2, 0:nz+2)
    integer :: u, iz, ir, ith, a, b, c
    open(newunit=u, file='RESTART', form='unformatted', status='old')
    read(u) nstepi, ntime, a, b, c
    if (a /= nt .or. b /= nr .or. c /= nz) stop 'RESTART: grid does not match inp.txt'
    do iz = 1, nz
      do ir = 1, nr
        do ith = 1, nt
          read(u) q1(ith, ir, iz), q2(ith, ir, iz), q3(ith, ir, iz)
        end do
      end do
    end do
    close(u)
  end subroutine restart_load

  ! analysis, run.F90:238-388 on one domain: 2D/3D Tecplot exports when
  ! mod(nstep,expstp)==0 and the vortex-core statistics of analysis.dat
  subroutine analysis(d, nstep, irk, ntime, nt, nr, nz, q1, q2, q3, pres, stat)
    type(deck), intent(in) :: d
    integer, intent(in) :: nstep, irk, nt, nr, nz
    real(c_double), intent(in) :: ntime
    real(c_double), intent(inout) :: q1(nt, 0:nr+2, 0:nz+2), q2(nt, 0:nr+2, 0:nz+2), q3(nt, 0:nr+2, 0:nz+2)
    real(c_double), intent(in) :: pres(nt, 0:nr+2, 0:nz+2)
    real(c_double), intent(inout) :: stat(20)
    character(len=5) :: tc
    character(len=2) :: rc
    integer :: ipass, iz, ir, ith, ithp1, u21, u22
    real(c_double) :: dth, dr, dz, dx1, dx2, dx3, rm(0:nr+1), zm(0:nz+1), thm(nt)
    real(c_double) :: xc, yc, q1m, q2m, q3m, dwdr, dvdz, vort, swrl, ux, uy
    real(c_double) :: vortmax, swrlmax, vortmaxall, swrlmaxall
    real(c_double) :: xvort, yvort, vvort, uvort, xswrl, yswrl, vswrl, uswrl, w
    logical :: export
    dth = 2.d0 * acos(-1.d0) / nt
    dr = d%rmax / nr
    dz = d%zmax / nz
    dx1 = 1.d0 / dth; dx2 = 1.d0 / dr; dx3 = 1.d0 / dz
    do ir = 0, nr + 1
      rm(ir) = (ir - 0.5d0) * dr
    end do
    do iz = 0, nz + 1
      zm(iz) = (iz - 0.5d0) * dz
    end do
    if (allocated(mesh_rm)) rm = mesh_rm   ! msh%rm of a stretched mesh
    if (allocated(mesh_zm)) zm = mesh_zm
    do ith = 1, nt
      thm(ith) = (ith - 0.5d0) * dth
    end do
    write(tc, '(I5.5)') nstep
    write(rc, '(I2.2)') irk
    export = mod(nstep, d%expstp) == 0
    vortmaxall = 0.d0; swrlmaxall = 0.d0
    do ipass = 1, 2
      if (export .and. ipass == 2) then
        open(newunit=u21, file='2Dfld'//tc//'RK'//rc//'.dat', form='formatted', status='replace')
        open(newunit=u22, file='3Dfld'//tc//'RK'//rc//'PR00.dat', form='formatted', status='replace')
        write(u22, "(a)") 'title = "sample mesh"'
        write(u22, "(a)") 'variables = "x", "y", "z", "u", "v", "w", "p", "ut", "ur", "vr", "sr0", "vr0" '
        write(u22, "(a,I5,a,I5,a,I5,a)") 'zone i=', nt, ', j=', nr, ',k=', nz, ' f=point'
      end if
      vortmax = -10000.d0; swrlmax = -10000.d0
      xvort = 0; yvort = 0; vvort = 0; uvort = 0; xswrl = 0; yswrl = 0; vswrl = 0; uswrl = 0
      do iz = 1, nz
        do ir = 1, nr
          do ith = 1, nt
            ithp1 = ith + 1
            if (ithp1 > nt) ithp1 = 1
            if (ir == 1) then
              q3(ith, 0, iz) = q3(ith, 1, iz)
              q2(ith, 0, iz) = q2(ith, 1, iz)
              q1(ith, 0, iz) = q1(ith, 1, iz)
            end if
            xc = rm(ir) * cos(thm(ith))
            yc = rm(ir) * sin(thm(ith))
            if (nt > 1) q1m = 0.5 * (q1(ith, ir, iz) + q1(ithp1, ir, iz))
            if (nt == 1) q1m = q1(ith, ir, iz)
            q2m = 0.5 * (q2(ith, ir, iz) + q2(ith, ir + 1, iz))
            q3m = 0.5 * (q3(ith, ir, iz) + q3(ith, ir, iz + 1))
            dwdr = 0.5 * ((q3(ith, ir + 1, iz) + q3(ith, ir + 1, iz + 1)) - (q3(ith, ir - 1, iz) + &
                   q3(ith, ir - 1, iz + 1))) / (rm(ir + 1) - rm(ir - 1))
            dvdz = 0.5 * ((q2(ith, ir, iz + 1) / rm(ir) + q2(ith, ir + 1, iz + 1) / rm(ir + 1)) - &
                   (q2(ith, ir + 1, iz - 1) / rm(ir + 1) + q2(ith, ir, iz - 1) / rm(ir))) / (zm(iz + 1) - zm(iz - 1))
            vort = +(dvdz - dwdr)
            swrl = q1m
            ux = q2m * cos(thm(ith)) / rm(ir) - q1m * sin(thm(ith))
            uy = q2m * sin(thm(ith)) / rm(ir) + q1m * cos(thm(ith))
            if (zm(iz) < 0.90 * d%zmax .and. zm(iz) > 0.10 * d%zmax .and. rm(ir) < 0.80 * d%rmax .and. &
                rm(ir) > d%rmax * 0.05 .and. ir /= 1) then
              if (ipass == 1) vortmax = max(vort, vortmax)
              if (ipass == 1) swrlmax = max(swrl, swrlmax)
              if (ipass == 2) then
                w = dx1 * dx2 * dx3
                if (vort > 0.00 * vortmaxall) then
                  xvort = xvort + zm(iz) * vort**2 * rm(ir) * w
                  vvort = vvort + vort**2 * rm(ir) * w
                  yvort = yvort + vort**2 * rm(ir) * w
                  uvort = uvort + vort**2 * w
                end if
                if (swrl > 0.0 * swrlmaxall) then
                  xswrl = xswrl + zm(iz) * swrl**2 * rm(ir) * w
                  vswrl = vswrl + swrl**2 * rm(ir) * w
                  yswrl = yswrl + swrl**2 * rm(ir) * w
                  uswrl = uswrl + swrl**2 * w
                end if
              end if
            end if
            if (export .and. ipass == 2) then
              write(u22, "(20(E16.8))") xc, yc, zm(iz), ux, uy, q3(ith, ir, iz), pres(ith, ir, iz), q1m, &
                   q2m / rm(ir), vort, q1m / (swrlmaxall + 0.000001), vort / (vortmaxall + 0.000001)
              if (ith == 1) write(u21, "(20(E16.8))") xc, yc, zm(iz), ux, uy, q3(ith, ir, iz), &
                   pres(ith, ir, iz), q1m, q2m / rm(ir), vort, q1m / (swrlmaxall + 0.000001), &
                   vort / (vortmaxall + 0.000001)
            end if
          end do
        end do
        if (export .and. ipass == 2) write(u21, "(a)")
      end do
      if (export .and. ipass == 2) then
        write(u21, "(a)")
        close(u21)
        close(u22)
      end if
      if (ipass == 1) then
        vortmaxall = vortmax
        swrlmaxall = swrlmax
      end if
    end do
    call analysis_record(d, nstep, ntime, (/vortmaxall, swrlmaxall, xvort, yvort, vvort, uvort, &
                         xswrl, yswrl, vswrl, uswrl/), stat)
  end subroutine analysis

  ! Second half of analysis (run.F90:362-384): the statistics update and the
  ! analysis.dat record, from a(1:10) = vortmax, swrlmax, stat(1:8) -- the
  ! host loop above or the device reduction vrh_analysis.
  subroutine analysis_record(d, nstep, ntime, a, stat)
    type(deck), intent(in) :: d
    integer, intent(in) :: nstep
    real(c_double), intent(in) :: ntime, a(10)
    real(c_double), intent(inout) :: stat(20)
    real(c_double) :: vortmaxall, swrlmaxall, xvortold, xswrlold, xvortnew, xswrlnew
    integer :: u23, i
    vortmaxall = a(1)
    swrlmaxall = a(2)
    xvortold = stat(1) / stat(3)
    xswrlold = stat(5) / stat(7)
    stat(1:8) = a(3:10)
    xvortnew = stat(1) / stat(3)
    xswrlnew = stat(5) / stat(7)
    stat(9) = (xvortnew - xvortold) / (d%dt * d%anlstp)
    stat(10) = (xswrlnew - xswrlold) / (d%dt * d%anlstp)
    if (nstep == 1) then
      open(newunit=u23, file='analysis.dat', form='formatted', status='replace')
    else
      open(newunit=u23, file='analysis.dat', form='formatted', position='append')
    end if
    write(u23, "(I5,20(E16.8))") nstep, ntime, vortmaxall, swrlmaxall, &
        (stat(i)**(1.0 / float(i)) / stat(i + 2)**(1.0 / float(i)), i = 1, 2), (stat(i + 2), i = 1, 2), &
        ((stat(i + 4)**(1.0 / float(i))) / stat(i + 6)**(1.0 / float(i)), i = 1, 2), &
        (stat(i + 6)**(1.0 / float(i)), i = 1, 2), stat(9), stat(10)
    close(u23)
  end subroutine analysis_record

  ! particle export, part_tools.F90:444-467: header, then per live particle
  ! (time, concentration, id), u(1:nvar), d(1:10) with d(1) = det(J)
  subroutine part_export(nstep, ib, radius, st, n, nvar, u, alive)
    integer, intent(in) :: nstep, ib, n, nvar
    real(c_double), intent(in) :: radius, st
    real(c_double), intent(in) :: u(nvar, n)
    integer(c_int), intent(in) :: alive(n)
    character(len=5) :: tc
    character(len=2) :: bc
    integer :: uu, ip, ii
    real(c_double) :: v(3), dd(10)
    write(tc, '(I5.5)') nstep
    write(bc, '(I2.2)') ib
    open(newunit=uu, file='part'//tc//'size'//bc//'.dat', form='formatted', status='replace')
    write(uu, "(A32,E20.8,A12,E20.8,A2)") 'title = "particles diameter: ', radius, ' stokes: ', st, '"'
    write(uu, "(A32,200(A4,I3.3))") 'variables = "id", "time", "c",', ('u   ', ii, ii = 1, nvar), ('d   ', ii, ii = 1, 10)
    do ip = 1, n
      if (alive(ip) == 0) cycle
      dd = 0.d0
      if (nvar >= 15) then
        dd(1) = u(7, ip) * u(11, ip) * u(15, ip) + u(8, ip) * u(12, ip) * u(13, ip) + u(9, ip) * u(10, ip) * u(14, ip) &
              - u(13, ip) * u(11, ip) * u(9, ip) - u(10, ip) * u(8, ip) * u(15, ip) - u(14, ip) * u(12, ip) * u(7, ip)
      end if
      v = (/0.d0, 1.d0, dble(ip)/)
      write(uu, "(200(E20.8))") (v(ii), ii = 1, 3), (u(ii, ip), ii = 1, nvar), (dd(ii), ii = 1, 10)
    end do
    close(uu)
  end subroutine part_export

end module vring_io
