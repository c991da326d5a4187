! drive_solver.f90 -- a caller of the reference's solver(...) interface, as
! src/Simple.f calls it, linked against libfvcell_fortran.a (the drop-in for
! Solver.o).  Reads one system from a raw stream file, calls
!   solver(ad,au,al,x,b,iax,ja,neq,nad,unsym,tol,maxit,code,matriz,nlit)
! and getRes, and writes x plus the status of the last solve.
!
! usage: drive_solver.exe <input.bin> <output.bin> <unit13.log>
! input : int32 code, matriz, unsym, omp, neq, nad, maxit, nlit, n_iax, n_ja,
!         n_ad, n_al, n_au; float64 tol; int32 iax(n_iax), ja(n_ja);
!         float64 ad(n_ad), al(n_al), au(n_au), b(neq), x0(neq)
! output: float64 x(neq); int32 status, iterations; float64 energy;
!         float64 u(neq) = x permuted back by getRes(num = reversed ids);
!         float64 time, matvectime, dottime, precontime of COMMON /timectrl/
program drive_solver
  implicit none
  integer :: num_threads, nThreadsCell, nThreadsSolver, nThreadsDot, nThreadsMatVec
  logical :: openmpCell, openmpSolver
  common /omp_control/ num_threads, openmpCell, openmpSolver, nThreadsCell, &
                       nThreadsSolver, nThreadsDot, nThreadsMatVec
  integer :: hdr(13), code, matriz, unsym_i, omp, neq, nad, maxit, nlit, i
  integer :: status, iterations
  real*8 :: tol, energy, true_relres, seconds
  integer, allocatable :: iax(:), ja(:), num(:)
  real*8, allocatable :: ad(:), al(:), au(:), b(:), x(:), u(:)
  character(len=512) :: fin, fout, flog
  logical :: unsym
  real*8 :: solvTime, elmTime, vecTime, dsTime, numeqTime, reordtime, totaltime, matvectime, &
            dottime, time, time0, readtime, viztime, grtime, precontime
  common /timectrl/ solvTime, elmTime, vecTime, dsTime, numeqTime, reordtime, totaltime, &
                    matvectime, dottime, time, time0, readtime, viztime, grtime, precontime

  call get_command_argument(1, fin)
  call get_command_argument(2, fout)
  call get_command_argument(3, flog)
  open(10, file=trim(fin), access='stream', form='unformatted', status='old')
  read(10) hdr
  read(10) tol
  code = hdr(1); matriz = hdr(2); unsym_i = hdr(3); omp = hdr(4); neq = hdr(5)
  nad = hdr(6); maxit = hdr(7); nlit = hdr(8)
  allocate(iax(max(hdr(9), 1)), ja(max(hdr(10), 1)), ad(max(hdr(11), 1)), al(max(hdr(12), 1)), &
           au(max(hdr(13), 1)), b(neq), x(neq), u(neq), num(neq))
  read(10) iax(1:hdr(9))
  read(10) ja(1:hdr(10))
  read(10) ad(1:hdr(11))
  if (hdr(12) > 0) read(10) al(1:hdr(12))
  if (hdr(13) > 0) read(10) au(1:hdr(13))
  read(10) b
  read(10) x
  close(10)
  if (hdr(13) == 0) au = al
  openmpSolver = (omp /= 0)
  nThreadsSolver = 1
  unsym = (unsym_i /= 0)
  matvectime = 0.0d0
  dottime = 0.0d0
  precontime = 0.0d0
  open(13, file=trim(flog), status='replace', action='write')
  call solver(ad, au, al, x, b, iax, ja, neq, nad, unsym, tol, maxit, code, matriz, nlit)
  close(13)
  call fvc_get_last(status, iterations, energy, true_relres, seconds)
  do i = 1, neq
    num(i) = neq + 1 - i
  end do
  call getRes(u, x, num, neq, neq, .false.)
  open(11, file=trim(fout), access='stream', form='unformatted', status='replace')
  write(11) x
  write(11) status, iterations
  write(11) energy
  write(11) u
  write(11) time, matvectime, dottime, precontime
  ! the second pressure-correction solve of a SIMPLE step: same A, kept on
  ! the device, right-hand side 2 b; every iterate scales by the power of
  ! two exactly, so the solution is exactly 2 x
  b = 2.0d0 * b
  x = 0.0d0
  open(13, file=trim(flog)//'.kept', status='replace', action='write')
  call solverKeepMatrix()
  call solver(ad, au, al, x, b, iax, ja, neq, nad, unsym, tol, maxit, code, matriz, 2)
  close(13)
  write(11) x
  close(11)
end program drive_solver
