! Test infrastructure: drives the reference OVERTURNS-2D subroutines, built
! from their own sources by oracle/Makefile, and dumps per-stage arrays so the
! C restatement (oracle/ot2d_oracle.c) and the HIP path can be checked
! against the reference itself.  Never part of the product path.
!
! The set-up follows the reference main program (src/Turns_overset.f:172-353):
! read_inputs, computeTScoefs, determine_size, set_pointers_globals, initia.
! One iteration replays the body of step() (src/step.f:38-215) call by call so
! that every intermediate array can be written out:
!   time_step (src/time.f:2) -> compute_residual (src/step.f:221)
!   -> residual norms (src/step.f:60-93) -> rbc (src/bc.f:88) -> s*tscale
!   -> ilu2d (src/lhs.f:470) -> solve_turbulence_equations (src/step.f:303)
!   -> q update (src/step.f:205-211) -> monitor (src/step.f:332)
! Mode "step" calls the reference step() itself instead, so a test can check
! that the replay is bit-identical to it.
!
! Control (environment):
!   OT2D_MODE    run | stages | step       (default run)
!   OT2D_NSTEPS  iterations to run         (default: nsteps of the deck)
!   OT2D_QIN     optional raw file holding the flat internal q array
!   OT2D_ISTEP0  istep0 to start from       (default 0)
!   OT2D_OUT     output directory           (default .)
!   OT2D_DUMPAT  in run mode also dump stage arrays at this iteration (0: none)
! With several meshes (num_grids > 1) the connectivity is found as the main
! program does (connect2d, update_halo_iblanks) and do_interpolations runs
! after every iteration.
! Outputs (raw little-endian float64 / int32, Fortran order):
!   norms.txt          it rsum resrho resmax jmax kmax nmax   (full precision)
!   forces.txt         it cl cd cm cpower (every npnorm iterations, mesh 1)
!   <stage>_m<im>.bin  per-stage arrays for mesh im
!   grid_m<im>.bin     x,y,xx,xy,yx,yy,q(:,:,nq) after initialisation
!   conn_s<nsp>_m<im>.txt, imesh/ibc/idonor/frac_s<nsp>_m<im>.bin  overset lists
program ref_harness
  use params_global
  use ihc
  implicit none
  real, allocatable :: s(:), q(:), qtn(:), qtnm1(:), qnewt(:), ss(:)
  real, allocatable :: x(:), y(:), xv(:), yv(:), tscale(:), bt(:)
  integer, allocatable :: iblank(:)
  real, allocatable :: xx(:), xy(:), yx(:), yy(:), ug(:), vg(:), ugv(:), vgv(:), turmu(:)
  real, allocatable :: xole(:), yole(:), xold(:), yold(:)
  integer, allocatable :: jgmx(:), kgmx(:), ipointer(:,:)
  real, allocatable :: tspec(:), Ds(:,:)
  integer :: nmesh, im, ig, igq, igs, igv, igb, jd, kd, it, nit, mstop, istart
  integer :: dumpat, ios
  character(len=256) :: mode, outdir, qin, buf
  character(len=40) :: bcfile, fprefix
  logical :: dump
  real :: cfx, cfy, cmt, clt, cdt, cpw
  ! overset connectivity, as the main program holds it (src/Turns_overset.f:209-237)
  real, allocatable :: xgl(:,:,:), ygl(:,:,:), volg(:,:,:), xglv(:,:,:), yglv(:,:,:), frac(:,:,:,:)
  integer, allocatable :: ibgl(:,:,:), imesh(:,:,:,:), idonor(:,:,:,:), ibc(:,:,:)
  integer, allocatable :: nfringe(:,:), ndonor(:,:), iisptr(:,:), iieptr(:,:), jgmxv(:), kgmxv(:)
  character(len=40), allocatable :: ihcbcfile(:)
  integer :: Nj, Nk, idsize, nsp, j, k

  nmesh = 6
  allocate(jgmx(nmesh), kgmx(nmesh))
  jgmx = 0; kgmx = 0
  bcfile = "bc.inp"
  fprefix = "fort"
  call read_inputs(jgmx, kgmx, nmesh, bcfile)
  allocate(tspec(nspec), Ds(nspec, nspec))
  call computeTScoefs(nspec, Ds)
  allocate(ipointer(nmesh, 5))
  call determine_size(jgmx, kgmx, nspec, nq, nv, ipointer, nmesh, &
       igrd, igrdv, igrdb, iqdim, isdim, mdim, nadd)
  allocate(s(isdim), ss(isdim), q(iqdim), qtn(isdim), qtnm1(isdim), qnewt(isdim))
  allocate(x(igrd), y(igrd), xv(igrdv), yv(igrdv), tscale(igrd), bt(igrd), iblank(igrd))
  allocate(xx(igrd), xy(igrd), yx(igrd), yy(igrd), ug(igrd), vg(igrd), turmu(igrd))
  allocate(ugv(igrdv), vgv(igrdv), xole(igrdv), yole(igrdv), xold(igrdv), yold(igrdv))
  s = 0.; ss = 0.; qtn = 0.; qtnm1 = 0.; qnewt = 0.; tscale = 0.; bt = 1.

  do im = 1, nmesh
    call set_pointers_globals(im, ipointer, ig, igq, igs, igv, igb, jd, kd, jgmx, kgmx, nmesh)
    call initia(q(igq), s(igs), x(ig), y(ig), xv(igv), yv(igv), xx(ig), xy(ig), yx(ig), &
         yy(ig), ug(ig), vg(ig), ugv(igv), vgv(igv), turmu(ig), &
         xold(igv), yold(igv), xole(igv), yole(igv), iblank(ig), tspec, im, jd, kd)
  end do

  call getenv_str("OT2D_MODE", "run", mode)
  call getenv_str("OT2D_OUT", ".", outdir)
  call getenv_str("OT2D_QIN", "", qin)
  call getenv_str("OT2D_ISTEP0", "0", buf); read(buf, *) istart
  call getenv_str("OT2D_NSTEPS", "-1", buf); read(buf, *) nit
  call getenv_str("OT2D_DUMPAT", "0", buf); read(buf, *) dumpat
  if (nit < 0) nit = nsteps
  if (num_grids > 1) call connectivity()

  if (len_trim(qin) > 0) then
    open(unit=91, file=trim(qin), access="stream", form="unformatted", status="old", iostat=ios)
    if (ios /= 0) stop "cannot open OT2D_QIN"
    read(91) q
    close(91)
  end if
  istep0 = istart

  do im = 1, nmesh
    call set_pointers_globals(im, ipointer, ig, igq, igs, igv, igb, jd, kd, jgmx, kgmx, nmesh)
    call dump_grid(im)
  end do

  open(unit=81, file=trim(outdir)//"/norms.txt", status="replace")
  open(unit=82, file=trim(outdir)//"/forces.txt", status="replace")
  totime = 0.0
  do it = 1, nit
    istep = it
    istep0 = istep0 + 1
    dump = (trim(mode) == "stages" .and. it == 1) .or. (dumpat > 0 .and. it == dumpat)
    do itn = 1, itnmax
      do im = 1, nmesh
        call set_pointers_globals(im, ipointer, ig, igq, igs, igv, igb, jd, kd, jgmx, kgmx, nmesh)
        if (.not. iprecon) Mp = 1.0
        dualtime = dtpseudo(itn)
        if (dump) call dump_r("qin", im, q(igq), jd*kd*nspec*nq)
        if (dump) call dump_r("turmuin", im, turmu(ig), jd*kd*nspec)
        call time_step(q(igq), xx(ig), xy(ig), yx(ig), yy(ig), ug(ig), vg(ig), &
             tscale(ig), bt(ig), iblank(ig), jd, kd)
        if (dump) call dump_r("tscale", im, tscale(ig), jd*kd*nspec)
        if (trim(mode) == "step") then
          call step(q(igq), qtn(igs), qtnm1(igs), qnewt(igs), s(igs), &
               x(ig), y(ig), xv(igv), yv(igv), iblank(ig), &
               xx(ig), xy(ig), yx(ig), yy(ig), ug(ig), vg(ig), ugv(igv), vgv(igv), &
               xold(igv), yold(igv), xole(igv), yole(igv), &
               turmu(ig), tscale(ig), bt(ig), Ds, im, jd, kd)
        else
          call replay_step(im, q(igq), s(igs), ss(igs), qtn(igs), qtnm1(igs), qnewt(igs), &
               x(ig), y(ig), xv(igv), yv(igv), iblank(ig), xx(ig), xy(ig), yx(ig), yy(ig), &
               ug(ig), vg(ig), ugv(igv), vgv(igv), xold(igv), yold(igv), xole(igv), yole(igv), &
               turmu(ig), tscale(ig), bt(ig), jd, kd, dump, it)
        end if
        call monitor(mstop, q(igq), jd, kd)
        if (dump) call dump_r("qout", im, q(igq), jd*kd*nspec*nq)
        if (dump) call dump_r("turmu", im, turmu(ig), jd*kd*nspec)
        if (mstop > 0) then
          write(81, '(a,i8)') "# mstop at iteration ", it
          goto 19
        end if
      end do
      ! src/Turns_overset.f:500-503
      if (nmesh > 1) call do_interpolations(q, jgmx, kgmx, ibc, imesh, idonor, frac, &
           nfringe, ndonor, iisptr, iieptr, idsize, iqdim, nmesh)
    end do
    if (mod(it, npnorm) == 0) then
      im = 1
      call set_pointers_globals(im, ipointer, ig, igq, igs, igv, igb, jd, kd, jgmx, kgmx, nmesh)
      if (bodyflag(im)) then
        call compute_forces(jd, kd, x(ig), y(ig), xv(igv), yv(igv), q(igq), &
             xx(ig), xy(ig), yx(ig), yy(ig), cfx, cfy, cmt, clt, cdt, cpw, im, fprefix)
        write(82, '(i8,5es25.16e3)') it, clt, cdt, cmt, cpw, cfx
      end if
    end if
  end do
19 continue
  close(81)
  close(82)
  do im = 1, nmesh
    call set_pointers_globals(im, ipointer, ig, igq, igs, igv, igb, jd, kd, jgmx, kgmx, nmesh)
    call dump_r("qfinal", im, q(igq), jd*kd*nspec*nq)
    call dump_r("turmufinal", im, turmu(ig), jd*kd*nspec)
  end do

contains

  ! connectivity set-up of the main program (src/Turns_overset.f:209-237,
  ! 245-256, 296-353), then the lists of every mesh are written out:
  ! conn_m<im>.txt holds nfringe ndonor iisptr iieptr, the .bin files
  ! imesh(nfringe,2), ibc(nfringe), idonor(ndonor,2), frac(ndonor,2)
  subroutine connectivity()
    integer :: jj
    character(len=40) :: istr
    Nj = jgmx(1); Nk = kgmx(1); idsize = jgmx(1)*kgmx(1)
    do im = 2, nmesh
      if (Nj <= jgmx(im)) Nj = jgmx(im)
      if (Nk <= kgmx(im)) Nk = kgmx(im)
      if (idsize <= jgmx(im)*kgmx(im)) idsize = jgmx(im)*kgmx(im)
    end do
    allocate(xgl(Nj,Nk,nmesh), ygl(Nj,Nk,nmesh), volg(Nj,Nk,nmesh), ibgl(Nj,Nk,nmesh))
    allocate(imesh(idsize,2,nmesh,nspec), idonor(idsize,2,nmesh,nspec))
    allocate(frac(idsize,2,nmesh,nspec), ibc(idsize,nmesh,nspec))
    allocate(nfringe(nmesh,nspec), ndonor(nmesh,nspec), iisptr(nmesh,nspec), iieptr(nmesh,nspec))
    allocate(ihcbcfile(nmesh), jgmxv(nmesh), kgmxv(nmesh))
    Nj = jgmx(1)-nadd; Nk = kgmx(1)-nadd
    do im = 2, nmesh
      if (Nj <= jgmx(im)-nadd) Nj = jgmx(im)-nadd
      if (Nk <= kgmx(im)-nadd) Nk = kgmx(im)-nadd
    end do
    allocate(xglv(Nj,Nk,nmesh), yglv(Nj,Nk,nmesh))
    do im = 1, nmesh
      call set_pointers_globals(im, ipointer, ig, igq, igs, igv, igb, jd, kd, jgmx, kgmx, nmesh)
      jgmxv(im) = jmax
      kgmxv(im) = kmax
      write(istr, *) im-1
      istr = adjustl(istr)
      ihcbcfile(im) = 'ihcbc.'//trim(istr)
    end do
    do nsp = 1, nspec
      do im = 1, nmesh
        call set_pointers_globals(im, ipointer, ig, igq, igs, igv, igb, jd, kd, jgmx, kgmx, nmesh)
        do k = 1, kd
          do j = 1, jd
            xgl(j,k,im) = x(ig-1 + jd*kd*(nsp-1) + jd*(k-1) + j)
            ygl(j,k,im) = y(ig-1 + jd*kd*(nsp-1) + jd*(k-1) + j)
            volg(j,k,im) = 1./q(igq - 1 + jd*kd*nspec*(nq-1) + jd*kd*(nsp-1) + jd*(k-1) + j)
          end do
        end do
        do k = 1, kmax
          do j = 1, jmax
            xglv(j,k,im) = xv(igv-1 + jmax*kmax*(nsp-1) + jmax*(k-1) + j)
            yglv(j,k,im) = yv(igv-1 + jmax*kmax*(nsp-1) + jmax*(k-1) + j)
          end do
        end do
      end do
      call connect2d(xglv, yglv, xgl, ygl, volg, ibgl, &
           idonor(:,:,:,nsp), frac(:,:,:,nsp), imesh(:,:,:,nsp), &
           ibc(:,:,nsp), iisptr(:,nsp), iieptr(:,nsp), ndonor(:,nsp), &
           nfringe(:,nsp), jgmxv, kgmxv, nhalo, nmesh, ihcbcfile)
      do im = 1, nmesh
        call set_pointers_globals(im, ipointer, ig, igq, igs, igv, igb, jd, kd, jgmx, kgmx, nmesh)
        do k = 1, kd
          do j = 1, jd
            iblank(ig-1 + jd*kd*(nsp-1) + jd*(k-1) + j) = ibgl(j,k,im)
          end do
        end do
      end do
    end do
    do im = 1, nmesh
      call set_pointers_globals(im, ipointer, ig, igq, igs, igv, igb, jd, kd, jgmx, kgmx, nmesh)
      call update_halo_iblanks(iblank(ig), jd, kd)
    end do
    do im = 1, nmesh
      do nsp = 1, nspec
        write(istr, '(a,i0,a,i0,a)') "_s", nsp, "_m", im, ".txt"
        open(unit=93, file=trim(outdir)//"/conn"//trim(istr), status="replace")
        write(93, '(4i10)') nfringe(im,nsp), ndonor(im,nsp), iisptr(im,nsp), iieptr(im,nsp)
        close(93)
        jj = nfringe(im,nsp)
        write(istr, '(a,i0)') "_s", nsp
        call dump_i("imesh"//trim(istr), im, reshape(imesh(1:jj,1:2,im,nsp), (/2*jj/)), 2*jj)
        call dump_i("ibc"//trim(istr), im, ibc(1:jj,im,nsp), jj)
        jj = ndonor(im,nsp)
        call dump_i("idonor"//trim(istr), im, reshape(idonor(1:jj,1:2,im,nsp), (/2*jj/)), 2*jj)
        call dump_r("frac"//trim(istr), im, reshape(frac(1:jj,1:2,im,nsp), (/2*jj/)), 2*jj)
      end do
    end do
  end subroutine

  subroutine getenv_str(name, default, val)
    character(len=*), intent(in) :: name, default
    character(len=*), intent(out) :: val
    integer :: l, st
    call get_environment_variable(name, val, l, st)
    if (st /= 0 .or. l == 0) val = default
  end subroutine

  subroutine dump_r(stage, im, a, n)
    character(len=*), intent(in) :: stage
    integer, intent(in) :: im, n
    real, intent(in) :: a(n)
    character(len=300) :: fn
    write(fn, '(a,"/",a,"_m",i0,".bin")') trim(outdir), trim(stage), im
    open(unit=92, file=trim(fn), access="stream", form="unformatted", status="replace")
    write(92) a
    close(92)
  end subroutine

  subroutine dump_i(stage, im, a, n)
    character(len=*), intent(in) :: stage
    integer, intent(in) :: im, n
    integer, intent(in) :: a(n)
    character(len=300) :: fn
    write(fn, '(a,"/",a,"_m",i0,".bin")') trim(outdir), trim(stage), im
    open(unit=92, file=trim(fn), access="stream", form="unformatted", status="replace")
    write(92) a
    close(92)
  end subroutine

  subroutine dump_grid(im)
    integer, intent(in) :: im
    integer :: n
    n = jd*kd*nspec
    call dump_r("x", im, x(ig), n)
    call dump_r("y", im, y(ig), n)
    call dump_r("xx", im, xx(ig), n)
    call dump_r("xy", im, xy(ig), n)
    call dump_r("yx", im, yx(ig), n)
    call dump_r("yy", im, yy(ig), n)
    call dump_r("ug", im, ug(ig), n)
    call dump_r("vg", im, vg(ig), n)
    call dump_r("ugv", im, ugv(igv), jmax*kmax*nspec)
    call dump_r("vgv", im, vgv(igv), jmax*kmax*nspec)
    call dump_r("qinit", im, q(igq), n*nq)
    call dump_i("iblank", im, iblank(ig), n)
    call dump_r("Ds", im, Ds, nspec*nspec)
  end subroutine

  subroutine replay_step(im, q, s, ss, qtn, qtnm1, qnewt, x, y, xv, yv, iblank, xx, xy, yx, yy, &
       ug, vg, ugv, vgv, xold, yold, xole, yole, turmu, tscale, bt, jd, kd, dump, it)
    integer, intent(in) :: im, jd, kd, it
    logical, intent(in) :: dump
    real :: q(jd,kd,nspec,nq), s(jd,kd,nspec,nv), ss(jd,kd,nspec,nv)
    real :: qtn(jd,kd,nspec,nv), qtnm1(jd,kd,nspec,nv), qnewt(jd,kd,nspec,nv)
    real :: x(jd,kd,nspec), y(jd,kd,nspec), xv(jmax,kmax,nspec), yv(jmax,kmax,nspec)
    integer :: iblank(jd,kd,nspec)
    real :: xx(jd,kd,nspec), xy(jd,kd,nspec), yx(jd,kd,nspec), yy(jd,kd,nspec)
    real :: ug(jd,kd,nspec), vg(jd,kd,nspec), ugv(jmax,kmax,nspec), vgv(jmax,kmax,nspec)
    real :: xold(jmax,kmax,nspec), yold(jmax,kmax,nspec), xole(jmax,kmax,nspec), yole(jmax,kmax,nspec)
    real :: turmu(jd,kd,nspec), tscale(jd,kd,nspec), bt(jd,kd,nspec)
    integer :: j, k, n, nsp, jd1, kd1, nd1
    real :: rsum, resrho, resmax, smrs, smrs1, smrs2, smrs3, smrs4, volum

    s = 0.
    ss = 0.
    call compute_residual(q, qtn, qtnm1, qnewt, s, ss, x, y, xv, yv, iblank, xx, xy, yx, yy, &
         ug, vg, ugv, vgv, xold, yold, xole, yole, turmu, tscale, bt, Ds, im, jd, kd)
    if (dump) call dump_r("qbc", im, q, jd*kd*nspec*nq)
    if (dump) call dump_r("res", im, s, jd*kd*nspec*nv)
    if (.not. timespectral .or. itn == 1) then
      do nsp = 1, nspec
        rsum = 0.; resrho = 0.; resmax = 0.
        jd1 = 2; kd1 = 2; nd1 = 1
        do k = kbeg, kend
          do n = 1, nmv
            do j = jbeg, jend
              smrs = abs(s(j,k,nsp,n))*max(iblank(j,k,nsp),0)
              if (smrs > resmax) then
                jd1 = j; kd1 = k; nd1 = n; resmax = smrs
              end if
            end do
          end do
          do j = jbeg, jend
            smrs1 = s(j,k,nsp,1)*max(iblank(j,k,nsp),0)
            smrs2 = s(j,k,nsp,2)*max(iblank(j,k,nsp),0)
            smrs3 = s(j,k,nsp,3)*max(iblank(j,k,nsp),0)
            smrs4 = s(j,k,nsp,4)*max(iblank(j,k,nsp),0)
            resrho = resrho + s(j,k,nsp,1)**2
            rsum = rsum + smrs1*smrs1 + smrs2*smrs2 + smrs3*smrs3 + smrs4*smrs4
          end do
        end do
        volum = float(jd*kd)
        resrho = sqrt(resrho/volum)
        rsum = sqrt(rsum/volum)
        write(81, '(i8,i4,i4,3es25.16e3,3i6)') it, im, nsp, rsum, resrho, resmax, jd1, kd1, nd1
        if (rsum > 1000.0) then
          write(81, '(a)') "# norm out of bounds"
          stop 'norm'
        end if
      end do
    end if
    do nsp = 1, nspec
      do k = kbeg, kend
        do j = jbeg, jend
          do n = 1, nv
            s(j,k,nsp,n) = s(j,k,nsp,n) + ss(j,k,nsp,n)
          end do
        end do
      end do
      call rbc(q(:,:,nsp,:), s(:,:,nsp,:), x(:,:,nsp), y(:,:,nsp), &
           xx(:,:,nsp), xy(:,:,nsp), yx(:,:,nsp), yy(:,:,nsp), &
           ug(:,:,nsp), vg(:,:,nsp), im, jd, kd, bt(:,:,nsp))
      if (dump) call dump_r("rbc", im, s, jd*kd*nspec*nv)
      do k = 1, kd
        do j = 1, jd
          s(j,k,nsp,1) = s(j,k,nsp,1)*tscale(j,k,nsp)
          s(j,k,nsp,2) = s(j,k,nsp,2)*tscale(j,k,nsp)
          s(j,k,nsp,3) = s(j,k,nsp,3)*tscale(j,k,nsp)
          s(j,k,nsp,4) = s(j,k,nsp,4)*tscale(j,k,nsp)
        end do
      end do
      if (dump) call dump_r("sdt", im, s, jd*kd*nspec*nv)
      if (ilhs /= 1 .or. iprecon) stop "harness: only ilhs=1 without preconditioning"
      call ilu2d(q(:,:,nsp,:), s(:,:,nsp,:), jd, kd, jbeg, jend, kbeg, kend, &
           xx(:,:,nsp), xy(:,:,nsp), yx(:,:,nsp), yy(:,:,nsp), &
           ug(:,:,nsp), vg(:,:,nsp), turmu(:,:,nsp), tscale(:,:,nsp))
    end do
    if (dump) call dump_r("dq", im, s, jd*kd*nspec*nv)
    if (dump) call dump_r("qlu", im, q, jd*kd*nspec*nq)
    do nsp = 1, nspec
      if (.not. lamin) call solve_turbulence_equations(turmu(:,:,nsp), q(:,:,nsp,:), &
           s(:,:,nsp,:), x(:,:,nsp), y(:,:,nsp), xv(:,:,nsp), yv(:,:,nsp), &
           xx(:,:,nsp), xy(:,:,nsp), yx(:,:,nsp), yy(:,:,nsp), ug(:,:,nsp), vg(:,:,nsp), &
           tscale(:,:,nsp), iblank(:,:,nsp), jd, kd, im, nsp)
    end do
    if (dump) call dump_r("qturb", im, q, jd*kd*nspec*nq)
    do nsp = 1, nspec
      do k = kbeg, kend
        do j = jbeg, jend
          q(j,k,nsp,1) = q(j,k,nsp,1) + s(j,k,nsp,1)*max(iblank(j,k,nsp),0)
          q(j,k,nsp,2) = q(j,k,nsp,2) + s(j,k,nsp,2)*max(iblank(j,k,nsp),0)
          q(j,k,nsp,3) = q(j,k,nsp,3) + s(j,k,nsp,3)*max(iblank(j,k,nsp),0)
          q(j,k,nsp,4) = q(j,k,nsp,4) + s(j,k,nsp,4)*max(iblank(j,k,nsp),0)
        end do
      end do
    end do
  end subroutine

end program ref_harness
