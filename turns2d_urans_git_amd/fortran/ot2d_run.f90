! ot2d_run -- the OVERTURNS-2D driver loop on libot2d.
!
! Reads the reference's own case files and writes its own output files, so a
! case directory prepared for the reference program runs unchanged:
!   stdin    &inputs namelist (same member names as src/modules.f:197-208)
!   bc.inp   one &bcinp group per mesh (src/setup.f:899-931)
!   fort.1   unformatted vertex grid (jmax,kmax; xv,yv) (src/setup.f:1253-1254,
!            src/io.f grid)
!   fort.3   restart solution when iread > 0 (src/init.f:140-166)
! and writes
!   fort.71  istep0, rsum, resrho, resmax, totime, theta_col every npnorm
!            steps (src/step.f:95-101)
!   fort.11  istep0, cl, cd, cm, cpower every npnorm steps, appended
!            (src/forces.f:46, 100, 123)
!   fort.8 / fort.9  restart solution and cell-centre grid at istep0 = 2,
!            every nrest steps and at the end (src/io.f:2-340)
! The time loop mirrors src/Turns_overset.f:357-554 for one steady mesh:
! steps between two outputs run on the device without host round trips
! (ot2d_iterate), the step that reports runs through ot2d_step.
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

  ! &bcinp
  integer :: nbc, irot_dir, ibtyp(OT2D_MAXBC), ibdir(OT2D_MAXBC), jbcs(OT2D_MAXBC), jbce(OT2D_MAXBC)
  integer :: kbcs(OT2D_MAXBC), kbce(OT2D_MAXBC)
  real(8) :: xcen, ycen
  namelist /bcinp/ nbc, irot_dir, xcen, ycen, ibtyp, ibdir, jbcs, jbce, kbcs, kbce

  type(ot2d_params) :: p
  type(ot2d_mesh_desc) :: md(1)
  type(c_ptr) :: case
  integer(c_int32_t) :: rc, jd, kd, nq, nv, done, istep0c, argmax(3)
  real(8), allocatable :: xv(:,:), yv(:,:), q(:,:,:), x(:,:), y(:,:)
  real(8) :: norms(3), f(6), theta_col, rey_a
  integer :: jm, km, ib, istep, istep0, nrc2, nxt, ios
  character(len=16) :: luenv

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
  if (num_grids /= 1) stop 'ot2d_run: overset grids (num_grids > 1) are not supported by this driver'
  if (if_obj .or. iteflap /= 0 .or. axisymmetric .or. ideform /= 0) &
       stop 'ot2d_run: objective, flap, axisymmetric and deforming options are not supported'

  p%iread = iread; p%nhalo = nhalo; p%half = half
  p%invisc = merge(1, 0, invisc); p%lamin = merge(1, 0, lamin); p%iturb = iturb; p%itrans = itrans
  p%irhsy = irhsy; p%ilim = ilim; p%ilhs = ilhs; p%iconstant_cfl = iconstant_cfl; p%idual = idual
  p%iprecon = merge(1, 0, iprecon); p%ntac = ntac; p%iunst = iunst; p%itnmax = itnmax
  p%timespectral = merge(1, 0, timespectral); p%nspec = nspec; p%motiontype = motiontype
  p%nsteps = nsteps; p%npnorm = npnorm; p%nrest = nrest
  p%fsmach = fsmach; p%alfa = alfa; p%rey = rey; p%vnuinf = vnuinf; p%tuinf = tuinf
  p%dt = dt; p%timeac = timeac; p%epse = epse; p%Mp = Mp
  p%rf = rf; p%angmax = angmax; p%fmtip = fmtip; p%rartio = rartio; p%totime = totime
  call get_environment_variable('OT2D_LU', luenv)
  p%lu_fp32 = merge(1, 0, trim(luenv) == 'fp32')
  if (timespectral .and. nspec > 1) stop 'ot2d_run: time-spectral decks need the multi-instance driver'

  ! ---- bc.inp
  open(10, file='bc.inp', status='old', form='formatted')
  nbc = 0; irot_dir = 1; xcen = 0.d0; ycen = 0.d0
  ibtyp = 0; ibdir = 0; jbcs = 0; jbce = 0; kbcs = 0; kbce = 0
  read(10, bcinp)
  close(10)

  ! ---- fort.1
  open(1, file='fort.1', status='old', form='unformatted')
  read(1) jm, km
  allocate(xv(jm, km), yv(jm, km))
  read(1) xv, yv
  close(1)

  md(1)%jmax = jm; md(1)%kmax = km; md(1)%nbc = nbc; md(1)%irot_dir = irot_dir
  md(1)%xcen = xcen; md(1)%ycen = ycen
  do ib = 1, OT2D_MAXBC
    md(1)%bc(ib) = ot2d_bcseg(ibtyp(ib), ibdir(ib), jbcs(ib), jbce(ib), kbcs(ib), kbce(ib))
  end do
  call ot2d_check(ot2d_sizes(p, md(1), jd, kd, nq, nv), 'ot2d_sizes')
  call ot2d_check(ot2d_case_create(p, 1_c_int32_t, md, case), 'ot2d_case_create')
  call ot2d_check(ot2d_set_grid(case, 0_c_int32_t, xv, yv), 'ot2d_set_grid')
  allocate(q(jd, kd, nq), x(jd, kd), y(jd, kd))
  call ot2d_check(ot2d_get_field(case, 0_c_int32_t, OT2D_F_X, x), 'ot2d_get_field')
  call ot2d_check(ot2d_get_field(case, 0_c_int32_t, OT2D_F_Y, y), 'ot2d_get_field')

  rey_a = merge(rey / fsmach, rey / fmtip, fsmach /= 0.d0)
  istep0 = 0
  if (iread > 0) call read_restart()
  theta_col = alfa

  open(71, file='fort.71', status='replace', form='formatted')
  nrc2 = nsteps - istep0
  istep = 0
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
    if (mod(istep0, npnorm) == 0) then
      write(71, '(i7,5(e18.10))') istep0, norms(1), norms(2), norms(3), totime, theta_col
      write(6, '(a,3i4,3(x,e12.4))') '  j,k,n,rmax,l2rho,l2 =', argmax(1), argmax(2), argmax(3), &
           norms(3), norms(2), norms(1)
    end if
    if (rc /= OT2D_OK) then
      call stopped(rc)
      exit
    end if
    if (mod(istep0, nrest) == 0 .or. istep0 == 2) call store()
    if (mod(istep, npnorm) == 0) then
      call ot2d_check(ot2d_forces(case, 0_c_int32_t, 0_c_int32_t, f), 'ot2d_forces')
      open(11, file='fort.11', position='append', form='formatted')
      write(11, '(i6,4(x,e18.10))') istep0, f(1), f(2), f(3), f(4)
      close(11)
    end if
  end do
  close(71)
  write(6, *) 'finished iterations, storing solution files now..'
  call store()
  rc = ot2d_case_destroy(case)
  stop 'ot2d_run completed successfully'

contains

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

  ! fort.8 (solution) and fort.9 (grid) as astore/store write them
  ! (src/io.f:2-130, 278-340); the reference scales its state to
  ! conserved variables for writing and back afterwards, which this repeats
  subroutine store()
    integer :: j, k, n
    call ot2d_check(ot2d_get_state(case, 0_c_int32_t, q), 'ot2d_get_state')
    do n = 1, 4
      do k = 1, kd
        do j = 1, jd
          q(j, k, n) = q(j, k, n) * q(j, k, nq)
        end do
      end do
    end do
    open(8, file='fort.8', status='replace', form='unformatted')
    write(8) jd, kd
    ! the reference holds rey per unit speed of sound (src/setup.f:1200-1204)
    write(8) fsmach, alfa, rey_a * fsmach, totime
    write(8) (((q(j, k, n), j = 1, jd), k = 1, kd), n = 1, 4)
    if (iturb == 1 .and. .not. lamin) write(8) (((q(j, k, n), j = 1, jd), k = 1, kd), n = 5, nv)
    write(8) istep0
    close(8)
    open(9, file='fort.9', status='replace', form='unformatted')
    write(9) jd, kd
    write(9) ((x(j, k), j = 1, jd), k = 1, kd), ((y(j, k), j = 1, jd), k = 1, kd)
    close(9)
    do n = 1, 4
      do k = 1, kd
        do j = 1, jd
          q(j, k, n) = q(j, k, n) / q(j, k, nq)
        end do
      end do
    end do
    call ot2d_check(ot2d_set_state(case, 0_c_int32_t, q), 'ot2d_set_state')
  end subroutine store

  ! fort.3 in the fort.8 layout (restr2 / restr_turb, src/io.f:383-427),
  ! then q/J (qdivj, src/init.f:180-206)
  subroutine read_restart()
    integer :: j, k, n, jt, kt
    real(8) :: fstip, alf, reypr
    call ot2d_check(ot2d_get_state(case, 0_c_int32_t, q), 'ot2d_get_state')
    open(3, file='fort.3', status='old', form='unformatted', iostat=ios)
    if (ios /= 0) stop 'ot2d_run: iread > 0 but fort.3 cannot be opened'
    read(3) jt, kt
    if (jt /= jd .or. kt /= kd) stop 'ot2d_run: fort.3 does not match the grid'
    read(3) fstip, alf, reypr, totime
    read(3) (((q(j, k, n), j = 1, jd), k = 1, kd), n = 1, 4)
    if (iturb == 1 .and. .not. lamin) read(3) (((q(j, k, n), j = 1, jd), k = 1, kd), n = 5, nv)
    read(3) istep0
    close(3)
    do n = 1, 4
      do k = 1, kd
        do j = 1, jd
          q(j, k, n) = q(j, k, n) / q(j, k, nq)
        end do
      end do
    end do
    call ot2d_check(ot2d_set_state(case, 0_c_int32_t, q), 'ot2d_set_state')
    call ot2d_check(ot2d_set_istep0(case, int(istep0, c_int32_t)), 'ot2d_set_istep0')
  end subroutine read_restart
end program ot2d_run
