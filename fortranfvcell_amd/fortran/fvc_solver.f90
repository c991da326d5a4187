! fvc_solver.f90 -- drop-in replacement for the reference's Solver.o.
!
! Exports the two external subroutines of src/Solver.f with their exact
! Fortran interfaces, so that linking libfvcell_fortran.so instead of
! Solver.o sends every caller (src/Simple.f:488,499,590,621,693,
! src/Transporte.f:111) to the gfx950 kernels through the C ABI of
! include/fvcell.h:
!
!   subroutine solver(ad,au,al,x,b,iax,ja,neq,nad,unsym,tol,maxit,code,matriz,nlit)
!   subroutine getRes(u,x,num,neq,numel,flag)
!
! The shim converts LOGICAL arguments to 0/1 (the value of .true. is
! compiler-dependent), reads openmpSolver from COMMON /omp_control/
! (include/openmp.fi:16-18) to pick the reference's initial-guess rule per
! path, and reproduces the reference's side effects: the unit-13 solver log
! line (src/Iterativos.f:91-93,216-218,429-431 and src/Iterativos_omp.f:
! 109-111,253-255), the console block when nlit == 1, and on failure the
! warning followed by STOP (src/Iterativos.f:76-77,198-199,405-406).
module fvc_capi
  use iso_c_binding
  implicit none

  type, bind(C) :: fvc_stats_t
    integer(c_int32_t) :: status, iterations, inner_iterations, engine
    real(c_double) :: energy, conv, last, true_relres, seconds
    integer(c_int32_t) :: grid_blocks, block_size, flags, reserved_i
    real(c_double) :: reserved(3)
  end type fvc_stats_t

  interface
    integer(c_int) function fvc_solver(ad, au, al, x, b, iax, ja, neq, nad, unsym, tol, &
                                       maxit, code, matriz, nlit, omp, st) bind(C, name='fvc_solver')
      import :: c_int, c_int32_t, c_double, fvc_stats_t
      real(c_double) :: ad(*), au(*), al(*), x(*), b(*)
      integer(c_int32_t) :: iax(*), ja(*)
      integer(c_int32_t), value :: neq, nad, unsym, maxit, code, matriz, nlit, omp
      real(c_double), value :: tol
      type(fvc_stats_t) :: st
    end function fvc_solver

    integer(c_int) function fvc_getres(u, x, num, neq, numel, flag) bind(C, name='fvc_getres')
      import :: c_int, c_int32_t, c_double
      real(c_double) :: u(*), x(*)
      integer(c_int32_t) :: num(*)
      integer(c_int32_t), value :: neq, numel, flag
    end function fvc_getres

    type(c_ptr) function fvc_last_error() bind(C, name='fvc_last_error')
      import :: c_ptr
    end function fvc_last_error

    integer(c_int) function fvc_keep_matrix(keep) bind(C, name='fvc_keep_matrix')
      import :: c_int, c_int32_t
      integer(c_int32_t), value :: keep
    end function fvc_keep_matrix
  end interface

  ! .true. mirrors the reference (print the warning and STOP on failure);
  ! callers that want a status instead call fvc_set_stop_on_failure(.false.)
  ! and read fvc_last_status after each solve.
  logical, save :: stop_on_failure = .true.
  integer, save :: fvc_last_status = 0
  type(fvc_stats_t), save :: fvc_last_stats
end module fvc_capi

subroutine fvc_set_stop_on_failure(flag)
  use fvc_capi
  implicit none
  logical, intent(in) :: flag
  stop_on_failure = flag
end subroutine fvc_set_stop_on_failure

! The next solver call reuses the device copy of A uploaded by the previous
! one (same arrays, contents unchanged by the caller): the second
! pressure-correction solve of a SIMPLE step (src/Simple.f:621), after which
! pform rebuilt only the right-hand side.  One-shot.
subroutine solverKeepMatrix()
  use fvc_capi
  implicit none
  integer(c_int) :: rc
  rc = fvc_keep_matrix(1_c_int32_t)
end subroutine solverKeepMatrix

subroutine fvc_get_last(status, iterations, energy, true_relres, seconds)
  use fvc_capi
  implicit none
  integer, intent(out) :: status, iterations
  real*8, intent(out) :: energy, true_relres, seconds
  status = fvc_last_status
  iterations = fvc_last_stats%iterations
  energy = fvc_last_stats%energy
  true_relres = fvc_last_stats%true_relres
  seconds = fvc_last_stats%seconds
end subroutine fvc_get_last

subroutine solver(ad, au, al, x, b, iax, ja, neq, nad, unsym, tol, maxit, code, matriz, nlit)
  use fvc_capi
  implicit none
  real*8 ad(*), au(*), al(*), x(*), b(*), tol
  integer iax(*), ja(*), neq, nad, maxit, code, matriz, nlit
  logical unsym
  ! COMMON /omp_control/ as laid out in include/openmp.fi:16-18
  integer :: num_threads, nThreadsCell, nThreadsSolver, nThreadsDot, nThreadsMatVec
  logical :: openmpCell, openmpSolver
  common /omp_control/ num_threads, openmpCell, openmpSolver, nThreadsCell, &
                       nThreadsSolver, nThreadsDot, nThreadsMatVec
  ! COMMON /timectrl/ as laid out in include/time.fi:10-12.  The reference's
  ! solver leaves its start stamp and elapsed time in time0/time
  ! (src/Iterativos.f:42,85-86,153,208) and its kernels accumulate matvectime,
  ! dottime and precontime (src/Matvec.f:27-37,394-399; src/Solver.f:83-85).
  ! The GPU fuses SpMV, dot products and the Jacobi step into the same
  ! kernels, so the whole solve time is added to matvectime.
  real*8 :: solvTime, elmTime, vecTime, dsTime, numeqTime, reordtime, totaltime, matvectime, &
            dottime, time, time0, readtime, viztime, grtime, precontime
  common /timectrl/ solvTime, elmTime, vecTime, dsTime, numeqTime, reordtime, totaltime, &
                    matvectime, dottime, time, time0, readtime, viztime, grtime, precontime
  type(fvc_stats_t) :: st
  integer(c_int) :: rc
  integer(c_int32_t) :: iuns, iomp
  logical :: omp_path, log_open
  character(len=15) :: label
  integer :: nlab
  integer(8) :: clk, clk_rate

  call system_clock(clk, clk_rate)
  time0 = dble(clk) / dble(max(clk_rate, 1_8))
  iuns = 0
  if (unsym) iuns = 1
  iomp = 0
  if (openmpSolver) iomp = 1
  rc = fvc_solver(ad, au, al, x, b, iax, ja, int(neq, c_int32_t), int(nad, c_int32_t), iuns, &
                  tol, int(maxit, c_int32_t), int(code, c_int32_t), int(matriz, c_int32_t), &
                  int(nlit, c_int32_t), iomp, st)
  fvc_last_status = rc
  fvc_last_stats = st
  time = st%seconds
  matvectime = matvectime + st%seconds

  ! which reference routine this call replaces, for the log label
  omp_path = openmpSolver .and. (matriz == 4 .or. (matriz == 1 .and. (code == 1 .or. unsym)))

  if (rc /= 0) then
    if (.not. stop_on_failure) return
    if (rc == 1) then
      if (code == 1) then
        write(*, 1200) maxit, st%last, st%conv
      else if (code == 2) then
        write(*, 1210) maxit, st%last, st%conv
      else
        write(*, 2100) maxit
      end if
    else
      write(*, '(a,i3,a,a)') ' *** fvcell solver error ', rc, ': ', last_error_text()
    end if
    stop
  end if

  if (code == 1) then
    label = 'PCG:       '
    nlab = 11
    if (omp_path) then
      label = 'PCG OMP      : '
      nlab = 15
    end if
    if (nlit == 1) write(*, 1100) 'PCG', tol, neq, nad, st%iterations, st%energy, st%seconds
  else if (code == 2) then
    label = 'PBICGSTAB: '
    nlab = 11
    if (omp_path) then
      label = 'PBICGSTAB OMP: '
      nlab = 15
    end if
    if (nlit == 1) write(*, 1100) 'PBICGSTAB', tol, neq, nad, st%iterations, st%energy, st%seconds
  else
    label = 'GMRES: '
    nlab = 7
    if (nlit == 1) write(*, 2000) tol, neq, nad
    write(*, 2050) st%iterations, st%inner_iterations, st%last, st%energy, st%seconds
  end if
  inquire(unit=13, opened=log_open)
  if (log_open) write(13, '(a,a,i9,a,d20.10,a,d20.10)') label(1:nlab), 'it', st%iterations, &
                      ' energy norm ', st%energy, ' tol ', tol
  return
1100 format(' (', a, ') solver:'/ &
            5x, 'Solver tol           = ', d20.6/ &
            5x, 'Number of equations  = ', i20/ &
            5x, 'Nad                  = ', i20/ &
            5x, 'Number of iterations = ', i20/ &
            5x, 'Energy norm          = ', d20.10/ &
            5x, 'CPU time (s)         = ', f20.2/)
1200 format(' *** WARNING:  PCG no convergence reached after ', i9, ' iterations !', d20.6, &
            ' Res ', d20.6, ' Conv'/)
1210 format(' *** WARNING: BiPCGSTAB no convergence reached after ', i9, ' iterations !', &
            d20.6, ' Res ', d20.6, ' Conv'/)
2000 format(' (GMRES) solver:'/ &
            5x, 'Solver tol           = ', d20.6/ &
            5x, 'Number of equations  = ', i20/ &
            5x, 'Nad                  = ', i20)
2050 format(5x, 'Number of cycles     = ', i20/ &
            5x, 'Number of iterations = ', i20/ &
            5x, 'Norm                 = ', d20.6/ &
            5x, 'Energy norm          = ', d20.10/ &
            5x, 'CPU time (s)         = ', f20.2/)
2100 format(' *** WARNING: no convergence reached for ', i9, ' cycles !', /)

contains

  function last_error_text() result(txt)
    character(len=256) :: txt
    character(kind=c_char), pointer :: p(:)
    integer :: i
    txt = ' '
    call c_f_pointer(fvc_last_error(), p, [256])
    do i = 1, 256
      if (p(i) == c_null_char) exit
      txt(i:i) = p(i)
    end do
  end function last_error_text
end subroutine solver

subroutine getRes(u, x, num, neq, numel, flag)
  use fvc_capi
  implicit none
  integer neq, numel
  real*8 u(*), x(*)
  integer num(*)
  logical flag
  integer(c_int) :: rc
  integer(c_int32_t) :: iflag
  iflag = 0
  if (flag) iflag = 1
  rc = fvc_getres(u, x, num, int(neq, c_int32_t), int(numel, c_int32_t), iflag)
  if (rc /= 0) then
    write(*, '(a)') ' *** getRes: equation number out of range'
    stop
  end if
end subroutine getRes
