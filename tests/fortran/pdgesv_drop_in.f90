! Drop-in check: an unmodified ScaLAPACK-style Fortran caller of PDGESV on a
! 1 x 1 BLACS grid, linked against libscalapack_amd (+ its 1 x 1 BLACS shim)
! instead of libscalapack.  The calling sequence is the reference's
! (SRC/pdgesv.f:1-2, TOOLS/descinit.f); A is a deterministic well-conditioned
! matrix, B = A * ones, and the program checks the scaled residual of the
! LU testers (TESTING/traditional/LIN/pdlaschk.f:296) and prints it.
program pdgesv_drop_in
  implicit none
  integer, parameter :: n = 300, nrhs = 2, nb = 64
  integer :: iam, nprocs, ictxt, nprow, npcol, myrow, mycol, info, i, j
  integer :: desca(9), descb(9), ipiv(n + nb)
  double precision, allocatable :: a(:,:), a0(:,:), b(:,:), b0(:,:), r(:,:)
  double precision :: anorm, xnorm, rnorm, sresid, eps

  call blacs_pinfo(iam, nprocs)
  call blacs_get(-1, 0, ictxt)
  call blacs_gridinit(ictxt, 'Row-major', 1, 1)
  call blacs_gridinfo(ictxt, nprow, npcol, myrow, mycol)
  call descinit(desca, n, n, nb, nb, 0, 0, ictxt, n, info)
  call descinit(descb, n, nrhs, nb, nb, 0, 0, ictxt, n, info)

  allocate(a(n, n), a0(n, n), b(n, nrhs), b0(n, nrhs), r(n, nrhs))
  do j = 1, n
    do i = 1, n
      a(i, j) = sin(dble(i * 7 + j * 13)) + merge(dble(n) / 8d0, 0d0, i == j)
    end do
  end do
  do j = 1, nrhs
    do i = 1, n
      b(i, j) = sum(a(i, :)) * dble(j)
    end do
  end do
  a0 = a
  b0 = b

  call pdgesv(n, nrhs, a, 1, 1, desca, ipiv, b, 1, 1, descb, info)
  if (info /= 0) then
    print '(a, i0)', 'FAIL info=', info
    stop 1
  end if

  r = b0 - matmul(a0, b)
  eps = epsilon(1d0) * 0.5d0
  anorm = maxval(sum(abs(a0), dim=2))
  sresid = 0d0
  do j = 1, nrhs
    xnorm = maxval(abs(b(:, j)))
    rnorm = maxval(abs(r(:, j)))
    sresid = max(sresid, rnorm / (anorm * xnorm * dble(n) * eps))
  end do
  if (sresid < 16d0 .and. abs(b(1, 2) - 2d0) < 1d-10) then
    print '(a, es10.3)', 'OK sresid=', sresid
  else
    print '(a, es10.3)', 'FAIL sresid=', sresid
    stop 1
  end if
  call blacs_gridexit(ictxt)
  call blacs_exit(0)
end program pdgesv_drop_in
