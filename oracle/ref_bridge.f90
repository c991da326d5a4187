! ref_bridge.f90 -- C-callable bridge into the reference Fortran solver,
! compiled from the reference's own sources by oracle/Makefile into
! oracle/_ref/libfvcref.so.
!
! TEST INFRASTRUCTURE ONLY: the tests use it to pin the C restatement
! (oracle/fvc_oracle.c) and bench.py times it as the CPU baseline.  The
! product never links it.
!
! Entry points (all bind(C), arrays passed as plain pointers, 1-based index
! values as in the reference):
!   ref_init(pool_words, omp, nthreads)   init_malloc (src/Malloc.f:33) and the
!                                         /omp_control/ flags (include/openmp.fi)
!   ref_open_log(path, n)                 unit 13, the solver log
!   ref_solver(...)                       solver (src/Solver.f:28)
!   ref_matvec(kind, ...)                 the matvec_* routines (src/Matvec.f,
!                                         src/Matvec_omp.f)
!   ref_dot / ref_dot_ompl4               src/Matvec.f:383, src/Matvec_omp.f:1247
!   ref_gmres(kind, k, ...)               gmres (src/Iterativos.f:271) with any k
!   ref_pcg / ref_pbicgstab(kind, omp...) the Krylov routines with explicit matvec
module ref_bridge_state
  implicit none
  logical :: pool_ready = .false.
end module ref_bridge_state

subroutine ref_init(pool_words, omp, nthreads) bind(C, name='ref_init')
  use iso_c_binding
  use Malloc
  use ref_bridge_state
  implicit none
  ! the layout of COMMON /omp_control/ of include/openmp.fi:16-18
  integer :: num_threads, nThreadsCell, nThreadsSolver, nThreadsDot, nThreadsMatVec
  logical :: openmpCell, openmpSolver
  common /omp_control/ num_threads, openmpCell, openmpSolver, nThreadsCell, &
                       nThreadsSolver, nThreadsDot, nThreadsMatVec
  integer(c_int64_t), value :: pool_words
  integer(c_int), value :: omp, nthreads
  if (.not. pool_ready) then
    maxmem = pool_words
    call init_malloc()
    flush(6)           ! its banner leaves now, while the caller may have silenced stdout
    pool_ready = .true.
  end if
  openmpSolver = (omp /= 0)
  nThreadsSolver = nthreads
end subroutine ref_init

subroutine ref_open_log(path, n) bind(C, name='ref_open_log')
  use iso_c_binding
  implicit none
  integer(c_int), value :: n
  character(kind=c_char), intent(in) :: path(n)
  character(len=n) :: fname
  integer :: i
  do i = 1, n
    fname(i:i) = path(i)
  end do
  open(13, file=fname, status='replace', action='write')
end subroutine ref_open_log

subroutine ref_close_log() bind(C, name='ref_close_log')
  implicit none
  close(13)
end subroutine ref_close_log

subroutine ref_solver(ad, au, al, x, b, iax, ja, neq, nad, unsym, tol, maxit, code, &
                      matriz, nlit) bind(C, name='ref_solver')
  use iso_c_binding
  implicit none
  real(c_double) :: ad(*), au(*), al(*), x(*), b(*)
  integer(c_int) :: iax(*), ja(*)
  integer(c_int), value :: neq, nad, unsym, maxit, code, matriz, nlit
  real(c_double), value :: tol
  logical :: lunsym
  integer :: ineq, inad, imaxit, icode, imatriz, inlit
  lunsym = (unsym /= 0)
  ineq = neq; inad = nad; imaxit = maxit; icode = code; imatriz = matriz; inlit = nlit
  call solver(ad, au, al, x, b, iax, ja, ineq, inad, lunsym, tol, imaxit, icode, &
              imatriz, inlit)
end subroutine ref_solver

! kind: 1 csrd, 2 csr, 3 csrc, 4 csrc_sym, 5 csrd_sym, 6 ellpack,
!       7 csrd_ilu2_omp, 8 ellpack_omp, 9 csrd_omp
subroutine ref_matvec(kind, neq, ad, au, al, ia, ja, x, y) bind(C, name='ref_matvec')
  use iso_c_binding
  implicit none
  integer(c_int), value :: kind, neq
  real(c_double) :: ad(*), au(*), al(*), x(*), y(*)
  integer(c_int) :: ia(*), ja(*)
  real(c_double) :: dum
  integer :: n
  n = neq
  dum = 0.0d0
  select case (kind)
  case (1)
    call matvec_csrd(ad, au, al, x, y, ia, ja, n)
  case (2)
    call matvec_csr(ad, au, al, x, y, ia, ja, n)
  case (3)
    call matvec_csrc(ad, au, al, x, y, ia, ja, n)
  case (4)
    call matvec_csrc_sym(ad, au, al, x, y, ia, ja, n)
  case (5)
    call matvec_csrd_sym(ad, au, al, x, y, ia, ja, n)
  case (6)
    call matvec_ellpack(ad, au, al, x, y, ia, ja, n)
  case (7)
    call matvec_csrd_ilu2_omp(ad, au, al, x, y, ia, ja, n, dum)
  case (8)
    call matvec_ellpack_omp(ad, au, al, x, y, ia, ja, n)
  case (9)
    call matvec_csrd_omp(ad, au, al, x, y, ia, ja, n, dum)
  end select
end subroutine ref_matvec

function ref_dot(a, b, n) result(d) bind(C, name='ref_dot')
  use iso_c_binding
  implicit none
  real(c_double) :: a(*), b(*)
  integer(c_int), value :: n
  real(c_double) :: d
  real*8, external :: dot
  integer :: nn
  nn = n
  d = dot(a, b, nn)
end function ref_dot

function ref_dot_ompl4(a, b, n) result(d) bind(C, name='ref_dot_ompl4')
  use iso_c_binding
  implicit none
  real(c_double) :: a(*), b(*)
  integer(c_int), value :: n
  real(c_double) :: d
  real*8, external :: dot_ompL4
  integer :: nn
  nn = n
  d = dot_ompL4(a, b, nn)
end function ref_dot_ompl4

! GMRES(k) with an explicit matvec: kind 1 csrd, 5 csrd_sym, 3 csrc, 4 csrc_sym,
! 6 ellpack.  Work arrays are allocated here (sizes of src/Iterativos.f:278).
subroutine ref_gmres(kind, neq, nad, ad, au, al, ia, ja, m, b, x, k, tol, maxit, nlit) &
    bind(C, name='ref_gmres')
  use iso_c_binding
  implicit none
  integer(c_int), value :: kind, neq, nad, k, maxit, nlit
  real(c_double) :: ad(*), au(*), al(*), m(*), b(*), x(*)
  integer(c_int) :: ia(*), ja(*)
  real(c_double), value :: tol
  real*8, allocatable :: g(:), h(:), y(:), c(:), s(:), e(:)
  external matvec_csrd, matvec_csrd_sym, matvec_csrc, matvec_csrc_sym, matvec_ellpack
  integer :: n, nd, kk, mx, nl
  n = neq; nd = nad; kk = k; mx = maxit; nl = nlit
  allocate(g(int(n, 8) * (kk + 1)), h((kk + 1) * kk), y(kk), c(kk), s(kk), e(kk + 1))
  select case (kind)
  case (1)
    call gmres(n, nd, ia, ja, ad, au, al, m, b, x, kk, g, h, y, c, s, e, tol, mx, matvec_csrd, nl)
  case (5)
    call gmres(n, nd, ia, ja, ad, au, al, m, b, x, kk, g, h, y, c, s, e, tol, mx, matvec_csrd_sym, nl)
  case (3)
    call gmres(n, nd, ia, ja, ad, au, al, m, b, x, kk, g, h, y, c, s, e, tol, mx, matvec_csrc, nl)
  case (4)
    call gmres(n, nd, ia, ja, ad, au, al, m, b, x, kk, g, h, y, c, s, e, tol, mx, matvec_csrc_sym, nl)
  case (6)
    call gmres(n, nd, ia, ja, ad, au, al, m, b, x, kk, g, h, y, c, s, e, tol, mx, matvec_ellpack, nl)
  end select
  deallocate(g, h, y, c, s, e)
end subroutine ref_gmres

! PCG / BiCGSTAB with explicit matvec kind (6 ellpack, 1 csrd, 7 csrd_ilu2) and
! either the sequential routines (omp = 0: pcg/pbicgstab with dot) or the
! OpenMP ones (omp /= 0: pcg_omp/pbicgstab_omp with the *_omp matvec and
! dot_ompL4, inside their own parallel region of nThreadsSolver threads).
subroutine ref_pcg(kind, omp, neq, nad, ad, au, al, ia, ja, m, b, x, tol, maxit, init, nlit) &
    bind(C, name='ref_pcg')
  use iso_c_binding
  implicit none
  integer(c_int), value :: kind, omp, neq, nad, maxit, init, nlit
  real(c_double) :: ad(*), au(*), al(*), m(*), b(*), x(*)
  integer(c_int) :: ia(*), ja(*)
  real(c_double), value :: tol
  real*8, allocatable :: z(:), r(:), ty(:)
  real*8, external :: dot, dot_ompL4
  external matvec_csrd, matvec_ellpack, matvec_csrd_omp, matvec_ellpack_omp
  logical :: linit
  integer :: n, nd, mx, nl
  n = neq; nd = nad; mx = maxit; nl = nlit
  linit = (init /= 0)
  allocate(z(n), r(n), ty(1))
  if (omp == 0) then
    if (kind == 6) then
      call pcg(n, nd, ia, ja, ad, au, al, m, b, x, z, r, tol, mx, matvec_ellpack, dot, nl, linit)
    else
      call pcg(n, nd, ia, ja, ad, au, al, m, b, x, z, r, tol, mx, matvec_csrd, dot, nl, linit)
    end if
  else
    if (kind == 6) then
      call pcg_omp(n, nd, ia, ja, ad, au, al, m, b, x, z, r, tol, mx, matvec_ellpack_omp, &
                   dot_ompL4, ty, nl, linit)
    else
      call pcg_omp(n, nd, ia, ja, ad, au, al, m, b, x, z, r, tol, mx, matvec_csrd_omp, &
                   dot_ompL4, ty, nl, linit)
    end if
  end if
  deallocate(z, r, ty)
end subroutine ref_pcg

subroutine ref_pbicgstab(kind, omp, neq, nad, ad, au, al, ia, ja, m, b, x, tol, maxit, init, &
                         nlit) bind(C, name='ref_pbicgstab')
  use iso_c_binding
  implicit none
  integer(c_int), value :: kind, omp, neq, nad, maxit, init, nlit
  real(c_double) :: ad(*), au(*), al(*), m(*), b(*), x(*)
  integer(c_int) :: ia(*), ja(*)
  real(c_double), value :: tol
  real*8, allocatable :: t(:), v(:), r(:), p(:), z(:)
  real*8, external :: dot, dot_ompL4
  external matvec_csrd, matvec_ellpack, matvec_csrd_ilu2_omp, matvec_ellpack_omp
  logical :: linit
  integer :: n, nd, mx, nl
  n = neq; nd = nad; mx = maxit; nl = nlit
  linit = (init /= 0)
  allocate(t(n), v(n), r(n), p(n), z(n))
  if (omp == 0) then
    if (kind == 6) then
      call pbicgstab(n, nd, ia, ja, ad, au, al, m, b, x, t, v, r, p, z, tol, mx, &
                     matvec_ellpack, dot, nl, linit)
    else
      call pbicgstab(n, nd, ia, ja, ad, au, al, m, b, x, t, v, r, p, z, tol, mx, &
                     matvec_csrd, dot, nl, linit)
    end if
  else
    if (kind == 6) then
      call pbicgstab_omp(n, nd, ia, ja, ad, au, al, m, b, x, t, v, r, p, z, tol, mx, &
                         matvec_ellpack_omp, dot_ompL4, nl, linit)
    else
      call pbicgstab_omp(n, nd, ia, ja, ad, au, al, m, b, x, t, v, r, p, z, tol, mx, &
                         matvec_csrd_ilu2_omp, dot_ompL4, nl, linit)
    end if
  end if
  deallocate(t, v, r, p, z)
end subroutine ref_pbicgstab

! get_time (src/Time.f:2) -- omp_get_wtime, for timing the baseline the way
! the reference times its solves.
function ref_wtime() result(t) bind(C, name='ref_wtime')
  use iso_c_binding
  implicit none
  real(c_double) :: t
  real*8, external :: get_time
  t = get_time()
end function ref_wtime
