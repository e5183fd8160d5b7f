! ISO_C_BINDING interface to libot2d (include/ot2d.h).
!
! The derived types mirror the C structs field for field; the functions are
! the C-ABI entry points that replace the reference driver's calls into
! time_step / step / monitor / compute_forces (src/Turns_overset.f:475-554)
! and its set-up (initia, src/init.f:2-177).  Case handles are opaque
! type(c_ptr).  Every function returns the int32 status of the C call:
! 0 ok, 1 negative speed of sound (mstop), 2 norm out of bounds, < 0 error.
module ot2d_capi
  use, intrinsic :: iso_c_binding
  implicit none

  integer, parameter :: OT2D_MAXBC = 25
  integer(c_int32_t), parameter :: OT2D_OK = 0, OT2D_MSTOP = 1, OT2D_NORM = 2

  type, bind(C) :: ot2d_bcseg
    integer(c_int32_t) :: ibtyp, ibdir, jbcs, jbce, kbcs, kbce
  end type ot2d_bcseg

  type, bind(C) :: ot2d_mesh_desc
    integer(c_int32_t) :: jmax, kmax, nbc, irot_dir
    real(c_double) :: xcen, ycen
    type(ot2d_bcseg) :: bc(OT2D_MAXBC)
  end type ot2d_mesh_desc

  type, bind(C) :: ot2d_params
    integer(c_int32_t) :: iread, nhalo, half
    integer(c_int32_t) :: invisc, lamin, iturb, itrans
    integer(c_int32_t) :: irhsy, ilim, ilhs, iconstant_cfl, idual, iprecon
    integer(c_int32_t) :: ntac, iunst, itnmax, timespectral, nspec, motiontype
    integer(c_int32_t) :: nsteps, npnorm, nrest
    real(c_double) :: fsmach, alfa, rey, vnuinf, tuinf
    real(c_double) :: dt, timeac, epse, dualtime, Mp
    real(c_double) :: rf, angmax, fmtip, rartio, totime
    real(c_double) :: gamma, pr, rmue, tinf
    integer(c_int32_t) :: lu_fp32
    integer(c_int32_t) :: static_conn
    integer(c_int32_t) :: reserved(6)
  end type ot2d_params

  ! one ihcbc.<n> file (ihc/src/ihc.f90:147-208)
  type, bind(C) :: ot2d_ihcbc
    integer(c_int32_t) :: nbc
    integer(c_int32_t) :: ibtyp(25), ibdir(25), jbcs(25), jbce(25), kbcs(25), kbce(25)
  end type ot2d_ihcbc

  interface
    integer(c_int32_t) function ot2d_default_params(p) bind(C, name='ot2d_default_params')
      import :: c_int32_t, ot2d_params
      type(ot2d_params), intent(inout) :: p
    end function ot2d_default_params

    integer(c_int32_t) function ot2d_sizes(p, m, jd, kd, nq, nv) bind(C, name='ot2d_sizes')
      import :: c_int32_t, ot2d_params, ot2d_mesh_desc
      type(ot2d_params), intent(in) :: p
      type(ot2d_mesh_desc), intent(in) :: m
      integer(c_int32_t), intent(out) :: jd, kd, nq, nv
    end function ot2d_sizes

    integer(c_int32_t) function ot2d_case_create(p, nmesh, meshes, case) bind(C, name='ot2d_case_create')
      import :: c_int32_t, c_ptr, ot2d_params, ot2d_mesh_desc
      type(ot2d_params), intent(in) :: p
      integer(c_int32_t), value :: nmesh
      type(ot2d_mesh_desc), intent(in) :: meshes(*)
      type(c_ptr), intent(out) :: case
    end function ot2d_case_create

    integer(c_int32_t) function ot2d_case_destroy(case) bind(C, name='ot2d_case_destroy')
      import :: c_int32_t, c_ptr
      type(c_ptr), value :: case
    end function ot2d_case_destroy

    integer(c_int32_t) function ot2d_set_grid(case, im, xv, yv) bind(C, name='ot2d_set_grid')
      import :: c_int32_t, c_ptr, c_double
      type(c_ptr), value :: case
      integer(c_int32_t), value :: im
      real(c_double), intent(in) :: xv(*), yv(*)
    end function ot2d_set_grid

    integer(c_int32_t) function ot2d_set_state(case, im, q) bind(C, name='ot2d_set_state')
      import :: c_int32_t, c_ptr, c_double
      type(c_ptr), value :: case
      integer(c_int32_t), value :: im
      real(c_double), intent(in) :: q(*)
    end function ot2d_set_state

    integer(c_int32_t) function ot2d_get_state(case, im, q) bind(C, name='ot2d_get_state')
      import :: c_int32_t, c_ptr, c_double
      type(c_ptr), value :: case
      integer(c_int32_t), value :: im
      real(c_double), intent(out) :: q(*)
    end function ot2d_get_state

    integer(c_int32_t) function ot2d_get_field(case, im, field, out) bind(C, name='ot2d_get_field')
      import :: c_int32_t, c_ptr, c_double
      type(c_ptr), value :: case
      integer(c_int32_t), value :: im, field
      real(c_double), intent(out) :: out(*)
    end function ot2d_get_field

    ! time-accurate runs: dual time per Newton sub-iteration, and the norms of
    ! every sub-iteration of the last time step
    integer(c_int32_t) function ot2d_set_dtpseudo(case, n, dtpseudo) bind(C, name='ot2d_set_dtpseudo')
      import :: c_int32_t, c_ptr, c_double
      type(c_ptr), value :: case
      integer(c_int32_t), value :: n
      real(c_double), intent(in) :: dtpseudo(*)
    end function ot2d_set_dtpseudo

    integer(c_int32_t) function ot2d_subiter_norms(case, norms, argmax) bind(C, name='ot2d_subiter_norms')
      import :: c_int32_t, c_ptr, c_double
      type(c_ptr), value :: case
      real(c_double), intent(out) :: norms(*)
      integer(c_int32_t), intent(out) :: argmax(*)
    end function ot2d_subiter_norms

    integer(c_int32_t) function ot2d_get_iblank(case, im, iblank) bind(C, name='ot2d_get_iblank')
      import :: c_int32_t, c_ptr
      type(c_ptr), value :: case
      integer(c_int32_t), value :: im
      integer(c_int32_t), intent(out) :: iblank(*)
    end function ot2d_get_iblank

    integer(c_int32_t) function ot2d_set_istep0(case, istep0) bind(C, name='ot2d_set_istep0')
      import :: c_int32_t, c_ptr
      type(c_ptr), value :: case
      integer(c_int32_t), value :: istep0
    end function ot2d_set_istep0

    integer(c_int32_t) function ot2d_get_istep0(case, istep0) bind(C, name='ot2d_get_istep0')
      import :: c_int32_t, c_ptr
      type(c_ptr), value :: case
      integer(c_int32_t), intent(out) :: istep0
    end function ot2d_get_istep0

    integer(c_int32_t) function ot2d_step(case, norms, argmax) bind(C, name='ot2d_step')
      import :: c_int32_t, c_ptr, c_double
      type(c_ptr), value :: case
      real(c_double), intent(out) :: norms(*)
      integer(c_int32_t), intent(out) :: argmax(*)
    end function ot2d_step

    integer(c_int32_t) function ot2d_iterate(case, niter, hist, done) bind(C, name='ot2d_iterate')
      import :: c_int32_t, c_ptr
      type(c_ptr), value :: case
      integer(c_int32_t), value :: niter
      type(c_ptr), value :: hist
      integer(c_int32_t), intent(out) :: done
    end function ot2d_iterate

    integer(c_int32_t) function ot2d_forces(case, im, nsp, out) bind(C, name='ot2d_forces')
      import :: c_int32_t, c_ptr, c_double
      type(c_ptr), value :: case
      integer(c_int32_t), value :: im, nsp
      real(c_double), intent(out) :: out(6)
    end function ot2d_forces

    type(c_ptr) function ot2d_last_error_c() bind(C, name='ot2d_last_error')
      import :: c_ptr
    end function ot2d_last_error_c

    integer(c_size_t) function c_strlen(s) bind(C, name='strlen')
      import :: c_size_t, c_ptr
      type(c_ptr), value :: s
    end function c_strlen
    ! overset meshes: connect2d on the host, do_interpolations in every step
    integer(c_int32_t) function ot2d_connect(case, ihcbc, nfringe) bind(C, name='ot2d_connect')
      import :: c_int32_t, c_ptr, ot2d_ihcbc
      type(c_ptr), value :: case
      type(ot2d_ihcbc), intent(in) :: ihcbc(*)
      integer(c_int32_t), intent(out) :: nfringe
    end function ot2d_connect
    integer(c_int32_t) function ot2d_interpolate(case) bind(C, name='ot2d_interpolate')
      import :: c_int32_t, c_ptr
      type(c_ptr), value :: case
    end function ot2d_interpolate
    ! serving primitives
    integer(c_int32_t) function ot2d_set_convergence(case, icase, tol) bind(C, name='ot2d_set_convergence')
      import :: c_int32_t, c_ptr, c_double
      type(c_ptr), value :: case
      integer(c_int32_t), value :: icase
      real(c_double), value :: tol
    end function ot2d_set_convergence
    integer(c_int32_t) function ot2d_iterate_async(case, niter) bind(C, name='ot2d_iterate_async')
      import :: c_int32_t, c_ptr
      type(c_ptr), value :: case
      integer(c_int32_t), value :: niter
    end function ot2d_iterate_async
    integer(c_int32_t) function ot2d_sync(case) bind(C, name='ot2d_sync')
      import :: c_int32_t, c_ptr
      type(c_ptr), value :: case
    end function ot2d_sync
  end interface

  integer(c_int32_t), parameter :: OT2D_F_X = 0, OT2D_F_Y = 1

contains

  ! text of the last failing call
  function ot2d_last_error() result(msg)
    character(len=:), allocatable :: msg
    type(c_ptr) :: p
    character(kind=c_char), pointer :: chars(:)
    integer :: n, i
    p = ot2d_last_error_c()
    n = int(c_strlen(p))
    call c_f_pointer(p, chars, [n])
    allocate(character(len=n) :: msg)
    do i = 1, n
      msg(i:i) = chars(i)
    end do
  end function ot2d_last_error

  ! stop with the library's message on a negative status
  subroutine ot2d_check(rc, where)
    integer(c_int32_t), intent(in) :: rc
    character(len=*), intent(in) :: where
    if (rc < 0) then
      write(6, '(a,a,a,i0,a,a)') ' libot2d: ', where, ' failed (', rc, '): ', ot2d_last_error()
      stop 'libot2d error'
    end if
  end subroutine ot2d_check
end module ot2d_capi
