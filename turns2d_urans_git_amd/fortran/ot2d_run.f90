! ot2d_run -- the OVERTURNS-2D driver loop on libot2d.
!
! Reads the reference's own case files and writes its own output files, so a
! case directory prepared for the reference program runs unchanged:
!   stdin     &inputs namelist (same member names as src/modules.f:197-208)
!   bc.inp    one &bcinp group per mesh (src/setup.f:911-931)
!   fort.1    unformatted vertex grids: [nmesh], (jmax,kmax) per mesh, then
!             xv,yv per mesh (src/setup.f:1253-1255, src/io.f:343-376)
!   ihcbc.<n> one &ihcbcinp group per mesh n = 0..num_grids-1 when
!             num_grids > 1 (ihc/src/ihc.f90:147-208)
!   fort.3 / fort_TS<n>.3  restart solution when iread > 0
!             (src/init.f:140-166, src/io.f:383-427)
! and writes
!   fort.7<m> per mesh m: istep0, rsum, resrho, resmax, totime, theta_col,
!             one line per time instance every npnorm steps (src/step.f:95-101)
!   fort.11   forces of mesh 1 (when it has a wall, bc type 5) every npnorm
!             steps, appended; steady: istep0 cl cd cm cpower; time-spectral:
!             istep0 nsp cl cd cm cpower (src/forces.f:40-112)
!   cp_TS<n>.dat  surface pressure of mesh 1, instance n, with the forces
!             (src/forces.f:165-231)
!   fort.8 / fort.9, or fort_TS<n>.8 / fort_TS<n>.9 per instance: restart
!             solution and cell-centre grid (with iblank when overset) at
!             istep0 = 2, every nrest steps and at the end; SA.dat /
!             komega.dat (or _TS<n>) beside them (src/io.f:2-340)
! The time loop mirrors src/Turns_overset.f:357-554: steps between two
! outputs run on the device without host round trips (ot2d_iterate), the
! step that reports runs through ot2d_step.  Overset meshes are connected
! once at the start (connect2d, static meshes) and exchange fringe values in
! every step (do_interpolations).
!
! Environment: OT2D_LU=fp32 selects the FP32 LU-SGS operator (default: the
! reference's FP64 arithmetic).
program ot2d_run
  use, intrinsic :: iso_c_binding
  use ot2d_capi
  implicit none

  ! &inputs, the reference deck
  integer :: iread, num_grids, nblade, jmax, kmax, nhalo, jtail1, half, nsteps, nrest, npnorm, nspec
  logical :: timespectral, invisc, lamin, iprecon, static_conn, axisymmetric, if_obj, if_objtot
  integer :: iturb, itrans, iunst, motiontype, ntac, itnmax, irhsy, ilim, ilhs, idual, idual_turb
  integer :: iconstant_cfl, jint, kint, nmovie, isin, ideform, iteflap
  real(8) :: fsmach, fmtip, rartio, alfa, rey, vnuinf, tuinf, dt, dtpseudo(100), cfltur, timeac, epse
  real(8) :: Mp, totime, rf, ang0, angmax, phase, fourbar_toggle
  namelist /inputs/ iread, num_grids, nblade, jmax, kmax, nhalo, jtail1, half, nsteps, nrest, npnorm, &
       timespectral, nspec, fsmach, alfa, rey, invisc, lamin, iturb, itrans, iunst, motiontype, &
       fmtip, rartio, vnuinf, tuinf, ntac, itnmax, dt, dtpseudo, cfltur, timeac, &
       epse, irhsy, ilim, ilhs, idual, idual_turb, iconstant_cfl, iprecon, Mp, &
       totime, rf, ang0, angmax, phase, jint, kint, static_conn, fourbar_toggle, axisymmetric, &
       nmovie, isin, ideform, iteflap, if_obj, if_objtot

  ! &bcinp and &ihcbcinp
  integer :: nbc, irot_dir, ibtyp(OT2D_MAXBC), ibdir(OT2D_MAXBC), jbcs(OT2D_MAXBC), jbce(OT2D_MAXBC)
  integer :: kbcs(OT2D_MAXBC), kbce(OT2D_MAXBC), ibproc(OT2D_MAXBC)
  real(8) :: xcen, ycen
  namelist /bcinp/ nbc, irot_dir, xcen, ycen, ibtyp, ibdir, jbcs, jbce, kbcs, kbce
  namelist /ihcbcinp/ nbc, ibtyp, ibdir, jbcs, jbce, kbcs, kbce, ibproc

  ! one mesh: its vertex grid, and host copies of its state, cell centres
  ! and iblank for the output files
  type mesh_t
    integer :: jm = 0, km = 0, jd = 0, kd = 0
    integer :: jtail = 1                  ! wall start (bc type 5), src/init.f:60-66
    logical :: body = .false.
    real(8), allocatable :: xv(:,:), yv(:,:)
    real(8), allocatable :: q(:,:,:,:)   ! (jd,kd,nspec,nq)
    real(8), allocatable :: x(:,:,:), y(:,:,:)
    integer(c_int32_t), allocatable :: ib(:,:,:)
  end type mesh_t

  type(ot2d_params) :: p
  type(ot2d_mesh_desc), allocatable :: md(:)
  type(ot2d_ihcbc), allocatable :: ihc(:)
  type(mesh_t), allocatable :: ms(:)
  type(c_ptr) :: case
  integer(c_int32_t) :: rc, jd, kd, nq, nv, done, istep0c, nfr
  real(8), allocatable :: norms(:), itnorms(:)
  integer(c_int32_t), allocatable :: argmax(:), itargmax(:)
  real(8) :: f(6), theta_col, rey_a
  integer :: im, ib, istep, istep0, nrc2, nxt, ios, nmesh, ns, nsp, i, nm1, itn
  character(len=16) :: luenv
  character(len=64) :: fname

  ! ---- defaults: the reference's data statements
  rc = ot2d_default_params(p)
  iread = p%iread; nhalo = p%nhalo; half = p%half
  invisc = p%invisc /= 0; lamin = p%lamin /= 0; iturb = p%iturb; itrans = p%itrans
  irhsy = p%irhsy; ilim = p%ilim; ilhs = p%ilhs; iconstant_cfl = p%iconstant_cfl; idual = p%idual
  iprecon = p%iprecon /= 0; ntac = p%ntac; iunst = p%iunst; itnmax = p%itnmax
  timespectral = p%timespectral /= 0; nspec = p%nspec; motiontype = p%motiontype
  nsteps = p%nsteps; npnorm = p%npnorm; nrest = p%nrest
  fsmach = p%fsmach; alfa = p%alfa; rey = p%rey; vnuinf = p%vnuinf; tuinf = p%tuinf
  dt = p%dt; timeac = p%timeac; epse = p%epse; Mp = p%Mp
  rf = p%rf; angmax = p%angmax; fmtip = p%fmtip; rartio = p%rartio; totime = p%totime
  num_grids = 1; nblade = 1; jmax = 109; kmax = 36; jtail1 = 19
  dtpseudo = 1.0d0; cfltur = 1.0d0; idual_turb = 0; ang0 = 0.d0; phase = 0.d0
  jint = 1; kint = 1; static_conn = .false.; fourbar_toggle = 0.d0; axisymmetric = .false.
  nmovie = 50; isin = 0; ideform = 0; iteflap = 0; if_obj = .false.; if_objtot = .false.

  read(5, inputs)
  if (if_obj .or. iteflap /= 0 .or. axisymmetric .or. ideform /= 0) &
       stop 'ot2d_run: objective, flap, axisymmetric and deforming options are not supported'
  ! src/setup.f:940-945
  if (timespectral) then
    iunst = 0
  else
    nspec = 1
  end if
  if (iunst /= 0 .and. .not. (iunst == 1 .and. motiontype == 1)) &
       stop 'ot2d_run: of moving grids, only pitching (iunst = 1, motiontype = 1) is supported'
  if (invisc) lamin = .true.
  if (lamin) iturb = -1
  nmesh = num_grids
  ns = nspec

  p%iread = iread; p%nhalo = nhalo; p%half = half
  p%invisc = merge(1, 0, invisc); p%lamin = merge(1, 0, lamin); p%iturb = iturb; p%itrans = itrans
  p%irhsy = irhsy; p%ilim = ilim; p%ilhs = ilhs; p%iconstant_cfl = iconstant_cfl; p%idual = idual
  p%iprecon = merge(1, 0, iprecon); p%ntac = ntac; p%iunst = iunst; p%itnmax = itnmax
  p%timespectral = merge(1, 0, timespectral); p%nspec = nspec; p%motiontype = motiontype
  p%nsteps = nsteps; p%npnorm = npnorm; p%nrest = nrest
  p%fsmach = fsmach; p%alfa = alfa; p%rey = rey; p%vnuinf = vnuinf; p%tuinf = tuinf
  p%dt = dt; p%timeac = timeac; p%epse = epse; p%Mp = Mp
  p%rf = rf; p%angmax = angmax; p%fmtip = fmtip; p%rartio = rartio; p%totime = totime
  p%static_conn = merge(1, 0, static_conn)
  call get_environment_variable('OT2D_LU', luenv)
  p%lu_fp32 = merge(1, 0, trim(luenv) == 'fp32')

  allocate(md(nmesh), ms(nmesh), ihc(nmesh))

  ! ---- bc.inp: one group per mesh
  open(10, file='bc.inp', status='old', form='formatted')
  do im = 1, nmesh
    nbc = 0; irot_dir = 1; xcen = 0.d0; ycen = 0.d0
    ibtyp = 0; ibdir = 0; jbcs = 0; jbce = 0; kbcs = 0; kbce = 0
    read(10, bcinp)
    md(im)%nbc = nbc; md(im)%irot_dir = irot_dir; md(im)%xcen = xcen; md(im)%ycen = ycen
    do ib = 1, OT2D_MAXBC
      md(im)%bc(ib) = ot2d_bcseg(ibtyp(ib), ibdir(ib), jbcs(ib), jbce(ib), kbcs(ib), kbce(ib))
    end do
    do ib = 1, nbc
      if (ibtyp(ib) == 5) then
        ms(im)%body = .true.
        ms(im)%jtail = jbcs(ib)
      end if
    end do
  end do
  close(10)

  ! ---- fort.1
  open(1, file='fort.1', status='old', form='unformatted')
  if (nmesh > 1) then
    read(1) nm1
    if (nm1 /= nmesh) stop 'ot2d_run: fort.1 holds a different number of meshes than num_grids'
  end if
  read(1) (ms(im)%jm, ms(im)%km, im = 1, nmesh)
  do im = 1, nmesh
    allocate(ms(im)%xv(ms(im)%jm, ms(im)%km), ms(im)%yv(ms(im)%jm, ms(im)%km))
    read(1) ms(im)%xv, ms(im)%yv
    md(im)%jmax = ms(im)%jm; md(im)%kmax = ms(im)%km
  end do
  close(1)

  call ot2d_check(ot2d_case_create(p, int(nmesh, c_int32_t), md, case), 'ot2d_case_create')
  do im = 1, nmesh
    call ot2d_check(ot2d_sizes(p, md(im), jd, kd, nq, nv), 'ot2d_sizes')
    ms(im)%jd = jd; ms(im)%kd = kd
    allocate(ms(im)%q(jd, kd, ns, nq), ms(im)%x(jd, kd, ns), ms(im)%y(jd, kd, ns), ms(im)%ib(jd, kd, ns))
    call ot2d_check(ot2d_set_grid(case, int(im - 1, c_int32_t), ms(im)%xv, ms(im)%yv), 'ot2d_set_grid')
    call ot2d_check(ot2d_get_field(case, int(im - 1, c_int32_t), OT2D_F_X, ms(im)%x), 'ot2d_get_field')
    call ot2d_check(ot2d_get_field(case, int(im - 1, c_int32_t), OT2D_F_Y, ms(im)%y), 'ot2d_get_field')
  end do
  allocate(norms(3 * nmesh * ns), argmax(3 * nmesh * ns))
  allocate(itnorms(3 * nmesh * ns * itnmax), itargmax(3 * nmesh * ns * itnmax))
  ! the dual time of each Newton sub-iteration (src/Turns_overset.f:478)
  if (itnmax > 1) call ot2d_check(ot2d_set_dtpseudo(case, int(min(itnmax, 16), c_int32_t), dtpseudo), &
       'ot2d_set_dtpseudo')

  rey_a = merge(rey / fsmach, rey / fmtip, fsmach /= 0.d0)
  istep0 = 0
  if (iread > 0) call read_restart()
  theta_col = alfa
  ! time-spectral pitching: the reference leaves theta_col at the angle of
  ! the last instance of the last mesh it set up (find_coord_TS,
  ! src/setup.f:364-369; tspec, src/init.f:71-78)
  if (timespectral .and. motiontype == 1 .and. ns > 1 .and. rf /= 0.d0) then
    block
      real(8) :: pi, tsp, rotf, theta
      pi = 4.d0 * atan(1.d0)
      tsp = 2 * pi * (ns - 1) / ns / abs(rf)
      rotf = md(nmesh)%irot_dir * rf
      theta = angmax * pi / 180.d0 * (1.d0 - cos(rotf * tsp))
      theta_col = alfa + theta * 180.d0 / pi
    end block
  end if

  ! ---- connectivity of overset meshes (src/Turns_overset.f:294-353)
  if (nmesh > 1) then
    do im = 1, nmesh
      write(fname, '(a,i0)') 'ihcbc.', im - 1
      open(21, file=trim(fname), status='old', form='formatted', iostat=ios)
      if (ios /= 0) stop 'ot2d_run: an ihcbc.<n> file is missing'
      nbc = 0; ibtyp = 0; ibdir = 0; jbcs = 0; jbce = 0; kbcs = 0; kbce = 0; ibproc = 0
      read(21, ihcbcinp)
      close(21)
      ihc(im)%nbc = nbc
      ihc(im)%ibtyp = ibtyp; ihc(im)%ibdir = ibdir; ihc(im)%jbcs = jbcs; ihc(im)%jbce = jbce
      ihc(im)%kbcs = kbcs; ihc(im)%kbce = kbce
    end do
    call ot2d_check(ot2d_connect(case, ihc, nfr), 'ot2d_connect')
  end if

  do im = 1, nmesh
    write(fname, '(a,i0)') 'fort.', 70 + im
    open(70 + im, file=trim(fname), status='replace', form='formatted')
  end do
  nrc2 = nsteps - istep0
  istep = 0
  totime = 0.d0
  do while (istep < nrc2)
    ! next step whose end needs the host: a norm line (istep0), a force
    ! line (istep), a restart write (istep0 = 2 or a multiple of nrest)
    nxt = istep + 1
    do while (nxt < nrc2)
      if (mod(istep0 + (nxt - istep), npnorm) == 0 .or. mod(nxt, npnorm) == 0 .or. &
          istep0 + (nxt - istep) == 2 .or. mod(istep0 + (nxt - istep), nrest) == 0) exit
      nxt = nxt + 1
    end do
    if (nxt - istep > 1) then
      rc = ot2d_iterate(case, int(nxt - istep - 1, c_int32_t), c_null_ptr, done)
      do i = 1, done
        call advance_time(istep + i)
      end do
      istep = istep + done; istep0 = istep0 + done
      call ot2d_check(rc, 'ot2d_iterate')
      if (rc /= OT2D_OK) then
        call stopped(rc)
        exit
      end if
    end if
    rc = ot2d_step(case, norms, argmax)
    call ot2d_check(rc, 'ot2d_step')
    istep = istep + 1; istep0 = istep0 + 1
    call advance_time(istep)
    if (mod(istep0, npnorm) == 0) then
      ! one line per Newton sub-iteration of the step (src/step.f:52-101)
      if (itnmax > 1) then
        call ot2d_check(ot2d_subiter_norms(case, itnorms, itargmax), 'ot2d_subiter_norms')
      else
        itnorms = norms
        itargmax = argmax
      end if
      do itn = 1, itnmax
        do im = 1, nmesh
          do nsp = 1, ns
            i = 3 * (((itn - 1) * nmesh + im - 1) * ns + nsp - 1)
            write(70 + im, '(i7,5(e18.10))') istep0, itnorms(i + 1), itnorms(i + 2), itnorms(i + 3), totime, &
                 theta_col
            write(6, '(a,3i4,3(x,e12.4))') '  j,k,n,rmax,l2rho,l2 =', itargmax(i + 1), itargmax(i + 2), &
                 itargmax(i + 3), itnorms(i + 3), itnorms(i + 2), itnorms(i + 1)
          end do
        end do
      end do
    end if
    if (rc /= OT2D_OK) then
      call stopped(rc)
      exit
    end if
    if (mod(istep0, nrest) == 0 .or. istep0 == 2) call astore()
    if (mod(istep, npnorm) == 0 .and. ms(1)%body) call forces()
  end do
  do im = 1, nmesh
    close(70 + im)
  end do
  write(6, *) 'finished iterations, storing solution files now..'
  call astore()
  rc = ot2d_case_destroy(case)
  stop 'ot2d_run completed successfully'

contains

  ! the cell centres of mesh m as they are now (a moving grid turns them)
  subroutine cell_centres(m)
    integer, intent(in) :: m
    call ot2d_check(ot2d_get_field(case, int(m - 1, c_int32_t), OT2D_F_X, ms(m)%x), 'ot2d_get_field')
    call ot2d_check(ot2d_get_field(case, int(m - 1, c_int32_t), OT2D_F_Y, ms(m)%y), 'ot2d_get_field')
  end subroutine cell_centres

  ! the time and pitch angle of a moving grid at time step n of the run
  ! (src/Turns_overset.f:362-379)
  subroutine advance_time(n)
    integer, intent(in) :: n
    if (iunst /= 0) then
      totime = totime + dt
      theta_col = alfa + angmax * (1.d0 - cos(md(1)%irot_dir * rf * n * dt))
    end if
  end subroutine advance_time

  ! negative speed of sound ends the loop (src/Turns_overset.f:497); an
  ! out-of-bounds norm stops the run (src/step.f:104-110)
  subroutine stopped(st)
    integer(c_int32_t), intent(in) :: st
    if (st == OT2D_NORM) then
      call ot2d_check(ot2d_get_istep0(case, istep0c), 'ot2d_get_istep0')
      write(6, '(a,i0)') ' norm is out of bounds, solution suspended at istep0 = ', istep0c
      stop 'norm'
    end if
    write(6, '(a)') ' negative speed of sound: stopping the iterations'
  end subroutine stopped

  ! compute_forces for mesh 1 (src/forces.f:2-112): one fort.11 line per
  ! instance, and the instance's surface pressure file from force2d
  subroutine forces()
    integer :: n
    call ot2d_check(ot2d_get_state(case, 0_c_int32_t, ms(1)%q), 'ot2d_get_state')
    if (iunst /= 0) call cell_centres(1)
    open(11, file='fort.11', position='append', form='formatted')
    do n = 1, ns
      call ot2d_check(ot2d_forces(case, 0_c_int32_t, int(n - 1, c_int32_t), f), 'ot2d_forces')
      if (.not. timespectral) then
        if (iunst == 0) then
          write(11, '(i6,4(x,e18.10))') istep0, f(1), f(2), f(3), f(4)
        else
          write(11, '(i6,5(x,e18.10))') istep0, f(1), f(2), f(3), f(4), theta_col
        end if
      else
        write(11, '(2i6,4(x,e18.10))') istep0, n, f(1), f(2), f(3), f(4)
      end if
      call surface_cp(n)
    end do
    close(11)
  end subroutine forces

  ! cp at the wall cells averaged with the halo row, written as force2d
  ! does (src/forces.f:165-231): x, cp, y per wall cell, list-directed
  subroutine surface_cp(n)
    integer, intent(in) :: n
    integer :: j, k, j1, j2, kk, logcp
    real(8) :: zz, rr, rho, u, v, e, vsq, pp, pp1, cp, fmm, gm1, pinf
    character(len=60) :: filecp, int_str
    k = nhalo + 1
    j1 = ms(1)%jtail + nhalo
    if (half == 0) then
      j2 = ms(1)%jd - j1 + 1
    else
      j2 = ms(1)%jd - 1
    end if
    if (fsmach /= 0.d0) then
      fmm = fsmach
    else if (fmtip /= 0.d0) then
      fmm = fmtip
    else
      fmm = 1.d0
    end if
    gm1 = p%gamma - 1.d0
    pinf = 1.d0 / p%gamma
    write(int_str, '(i7)') n
    filecp = 'cp_TS' // trim(adjustl(int_str)) // '.dat'
    logcp = 90000 + n
    open(unit=logcp, file=trim(filecp), status='unknown')
    do j = j1, j2
      zz = 0.d0
      do kk = k, k - 1, -1
        associate (qq => ms(1)%q)
          rr = 1.d0 / qq(j, kk, n, 1)
          rho = qq(j, kk, n, 1) * qq(j, kk, n, nq)
          u = qq(j, kk, n, 2) * rr
          v = qq(j, kk, n, 3) * rr
          e = qq(j, kk, n, 4) * qq(j, kk, n, nq)
        end associate
        vsq = u * u + v * v
        pp = gm1 * (e - rho * vsq / 2.d0)
        pp1 = pp / pinf
        cp = 2.d0 * (pp1 - 1.d0) / (p%gamma * fmm**2)
        if (kk == k) then
          zz = cp
        else
          zz = 0.5d0 * (zz + cp)
        end if
      end do
      write(logcp, *) ms(1)%x(j, k, n), zz, ms(1)%y(j, k, n)
    end do
    close(logcp)
  end subroutine surface_cp

  ! astore (src/io.f:2-170): per instance, fort.8 (solution) and fort.9
  ! (grid, with iblank when overset) as multizone plot3d files, the
  ! turbulence model's tecplot file, and istep0 last.  The reference scales
  ! its state to conserved variables for writing and back afterwards, in
  ! place, which this repeats (the state is set back from the rescaled copy).
  subroutine astore()
    integer :: i, im, j, k, n, jj, kk, lg(ns), lq(ns), lt(ns)
    character(len=40) :: gfile, qfile, tfile, istr
    do n = 1, ns
      if (timespectral) then
        write(istr, *) n
        gfile = 'fort_TS' // trim(adjustl(istr)) // '.9'
        qfile = 'fort_TS' // trim(adjustl(istr)) // '.8'
      else
        gfile = 'fort.9'
        qfile = 'fort.8'
      end if
      lg(n) = 1000 * n + 9
      lq(n) = 1000 * n + 8
      lt(n) = 1000 * n + 10
      open(unit=lg(n), file=trim(gfile), status='replace', form='unformatted')
      open(unit=lq(n), file=trim(qfile), status='replace', form='unformatted')
      if (nmesh > 1) write(lg(n)) nmesh
      write(lg(n)) (ms(i)%jd, ms(i)%kd, i = 1, nmesh)
      if (nmesh > 1) write(lq(n)) nmesh
      write(lq(n)) (ms(i)%jd, ms(i)%kd, i = 1, nmesh)
    end do
    ! store (src/io.f:278-340), mesh by mesh
    do im = 1, nmesh
      associate (m => ms(im))
        jj = m%jd; kk = m%kd
        call ot2d_check(ot2d_get_state(case, int(im - 1, c_int32_t), m%q), 'ot2d_get_state')
        call ot2d_check(ot2d_get_iblank(case, int(im - 1, c_int32_t), m%ib), 'ot2d_get_iblank')
        if (iunst /= 0) call cell_centres(im)
        do n = 1, ns
          do k = 1, kk
            do j = 1, jj
              m%q(j, k, n, 1) = m%q(j, k, n, 1) * m%q(j, k, n, nq)
              m%q(j, k, n, 2) = m%q(j, k, n, 2) * m%q(j, k, n, nq)
              m%q(j, k, n, 3) = m%q(j, k, n, 3) * m%q(j, k, n, nq)
              m%q(j, k, n, 4) = m%q(j, k, n, 4) * m%q(j, k, n, nq)
            end do
          end do
          ! the reference holds rey per unit speed of sound (src/setup.f:1200-1204)
          write(lq(n)) fsmach, alfa, rey_a * fsmach, totime
          write(lq(n)) (((m%q(j, k, n, i), j = 1, jj), k = 1, kk), i = 1, 4)
          if (iturb == 1 .or. iturb == 2) write(lq(n)) (((m%q(j, k, n, i), j = 1, jj), k = 1, kk), i = 5, nv)
          if (num_grids > 1) then
            write(lg(n)) ((m%x(j, k, n), j = 1, jj), k = 1, kk), ((m%y(j, k, n), j = 1, jj), k = 1, kk), &
                 ((m%ib(j, k, n), j = 1, jj), k = 1, kk)
          else
            write(lg(n)) ((m%x(j, k, n), j = 1, jj), k = 1, kk), ((m%y(j, k, n), j = 1, jj), k = 1, kk)
          end if
          do k = 1, kk
            do j = 1, jj
              m%q(j, k, n, 1) = m%q(j, k, n, 1) / m%q(j, k, n, nq)
              m%q(j, k, n, 2) = m%q(j, k, n, 2) / m%q(j, k, n, nq)
              m%q(j, k, n, 3) = m%q(j, k, n, 3) / m%q(j, k, n, nq)
              m%q(j, k, n, 4) = m%q(j, k, n, 4) / m%q(j, k, n, nq)
            end do
          end do
        end do
        call ot2d_check(ot2d_set_state(case, int(im - 1, c_int32_t), m%q), 'ot2d_set_state')
      end associate
    end do
    ! turbulence files: storevnu / storekomega (src/io.f:76-276)
    if (iturb == 1 .or. iturb == 2) then
      do n = 1, ns
        if (iturb == 1) then
          tfile = 'SA'
        else
          tfile = 'komega'
        end if
        if (timespectral) then
          write(istr, *) n
          tfile = trim(tfile) // '_TS' // trim(adjustl(istr))
        end if
        tfile = trim(adjustl(tfile)) // '.dat'
        open(unit=lt(n), file=trim(tfile), status='replace', form='formatted')
        write(lt(n), *) "TITLE = ""Turbulence Model solution File"""
        if (iturb == 1) then
          write(lt(n), *) "VARIABLES = ""X"" ""Y"" ""IBlank"" ""Vnu_t"""
        else
          ! the reference's literal runs to column 132 of its fixed-form line
          ! before the continuation (src/io.f:137-138)
          write(lt(n), *) 'VARIABLES = "X" "Y" "IBlank" "K"' // repeat(' ', 62) // ' "Omega"'
        end if
      end do
      do im = 1, nmesh
        associate (m => ms(im))
          jj = m%jd; kk = m%kd
          do n = 1, ns
            ! k-omega SST: the turbulence record once more (storekomega
            ! writes it after store has, src/io.f:255)
            if (iturb == 2) write(lq(n)) (((m%q(j, k, n, i), j = 1, jj), k = 1, kk), i = 5, nv)
            write(lt(n), *) "ZONE"
            write(lt(n), *) "I = ", jj, ", J = ", kk
            write(lt(n), *) "ZONETYPE = Ordered, DATAPACKING = BLOCK"
            write(lt(n), '(3F22.13)') ((m%x(j, k, n), j = 1, jj), k = 1, kk)
            write(lt(n), '(3F22.13)') ((m%y(j, k, n), j = 1, jj), k = 1, kk)
            write(lt(n), '(3I10)') ((m%ib(j, k, n), j = 1, jj), k = 1, kk)
            if (iturb == 1) then
              write(lt(n), '(3F22.13)') ((m%q(j, k, n, 5), j = 1, jj), k = 1, kk)
            else
              write(lt(n), '(3F22.13)') ((m%q(j, k, n, 5) / (m%q(j, k, n, 1) * m%q(j, k, n, nq)), &
                   j = 1, jj), k = 1, kk)
              write(lt(n), '(3F22.13)') ((m%q(j, k, n, 6) / (m%q(j, k, n, 1) * m%q(j, k, n, nq)), &
                   j = 1, jj), k = 1, kk)
            end if
          end do
        end associate
      end do
    end if
    do n = 1, ns
      write(lq(n)) istep0
      close(lg(n))
      close(lq(n))
      if (iturb == 1 .or. iturb == 2) close(lt(n))
    end do
  end subroutine astore

  ! fort.3 / fort_TS<n>.3 in the fort.8 layout (initia, src/init.f:140-166;
  ! restr2 / restr_turb, src/io.f:383-427), then q/J (qdivj,
  ! src/init.f:180-206).  Like the reference, istep0 is the first word of
  ! the record after the last mesh's records.
  subroutine read_restart()
    integer :: i, im, j, k, n, jt(nmesh), kt(nmesh), nm, lr
    real(8) :: fstip, alf, reypr
    character(len=40) :: rfile, istr
    do im = 1, nmesh
      call ot2d_check(ot2d_get_state(case, int(im - 1, c_int32_t), ms(im)%q), 'ot2d_get_state')
    end do
    do n = 1, ns
      if (timespectral) then
        write(istr, *) n
        rfile = 'fort_TS' // trim(adjustl(istr)) // '.3'
      else
        rfile = 'fort.3'
      end if
      lr = 1000 * n + 3
      open(lr, file=trim(rfile), status='old', form='unformatted', iostat=ios)
      if (ios /= 0) stop 'ot2d_run: iread > 0 but a restart file cannot be opened'
      if (nmesh > 1) read(lr) nm
      read(lr) (jt(im), kt(im), im = 1, nmesh)
      do im = 1, nmesh
        if (jt(im) /= ms(im)%jd .or. kt(im) /= ms(im)%kd) stop 'ot2d_run: the restart file does not match the grid'
        read(lr) fstip, alf, reypr, totime
        read(lr) (((ms(im)%q(j, k, n, i), j = 1, ms(im)%jd), k = 1, ms(im)%kd), i = 1, 4)
        if (iturb == 1 .or. iturb == 2) &
             read(lr) (((ms(im)%q(j, k, n, i), j = 1, ms(im)%jd), k = 1, ms(im)%kd), i = 5, nv)
      end do
      read(lr) istep0
      close(lr)
    end do
    do im = 1, nmesh
      associate (m => ms(im))
        do n = 1, ns
          do k = 1, m%kd
            do j = 1, m%jd
              m%q(j, k, n, 1) = m%q(j, k, n, 1) / m%q(j, k, n, nq)
              m%q(j, k, n, 2) = m%q(j, k, n, 2) / m%q(j, k, n, nq)
              m%q(j, k, n, 3) = m%q(j, k, n, 3) / m%q(j, k, n, nq)
              m%q(j, k, n, 4) = m%q(j, k, n, 4) / m%q(j, k, n, nq)
            end do
          end do
        end do
        call ot2d_check(ot2d_set_state(case, int(im - 1, c_int32_t), m%q), 'ot2d_set_state')
      end associate
    end do
    call ot2d_check(ot2d_set_istep0(case, int(istep0, c_int32_t)), 'ot2d_set_istep0')
  end subroutine read_restart
end program ot2d_run
