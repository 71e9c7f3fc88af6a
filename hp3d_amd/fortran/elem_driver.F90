! elem_driver.F90 -- a Fortran host calling libhp3d_gpu through hp3d_gpu_mod,
! the way an hp3D `elem` would: element codes and geometry dofs in, ALOC /
! BLOC written in place with the caller's leading dimension (MAXDOFS in
! mumps_sc.F90:143-145).
!
! stdin (list-directed):
!   mode            1 = counts only (no device), 2 = elem, 3 = celem
!   space ld        energy space and leading dimension of ALOC
!   norder(19) nori_e(12) nori_f(6)
!   ngeo            then xnod(3, ngeo)
!   coef load
! stdout: "counts N nb", then for mode 2 "rc", K(1:N,1:N) column by column and
! b(1:N); for mode 3 "rc", Aii(1:ni,1:ni), Bi(1:ni), ASchur(nb,ni), BSchur(nb).
program elem_driver
  use iso_c_binding, only: c_int, c_double
  use hp3d_gpu_mod
  implicit none
  integer(c_int) :: mode, space, ld, norder(19), nori_e(12), nori_f(6), ngeo, load
  integer(c_int) :: n, nb, ni, rc, i, j
  real(c_double) :: coef
  real(c_double), allocatable :: xnod(:,:), aloc(:,:), bloc(:), aschur(:,:), bschur(:)

  read(*, *) mode
  read(*, *) space, ld
  read(*, *) norder, nori_e, nori_f
  read(*, *) ngeo
  allocate(xnod(3, ngeo))
  read(*, *) xnod
  read(*, *) coef, load

  rc = hp3d_gpu_ndof(space, norder, n, nb)
  call hp3d_gpu_check(rc, 'hp3d_gpu_ndof')
  write(*, '(a,1x,i0,1x,i0)') 'counts', n, nb
  if (mode == 1) stop
  ni = n - nb

  call hp3d_gpu_check(hp3d_gpu_init(0_c_int, 0_c_int), 'hp3d_gpu_init')
  allocate(aloc(ld, ld), bloc(ld), aschur(max(nb, 1), ni), bschur(max(nb, 1)))
  aloc = 0
  bloc = 0
  if (mode == 2) then
    rc = hp3d_gpu_elem(space, norder, nori_e, nori_f, xnod, ngeo, coef, load, aloc, ld, bloc)
    call hp3d_gpu_check(rc, 'hp3d_gpu_elem')
    write(*, '(a,1x,i0)') 'rc', rc
    do j = 1, n
      do i = 1, n
        write(*, '(es25.17e3)') aloc(i, j)
      end do
    end do
    do i = 1, n
      write(*, '(es25.17e3)') bloc(i)
    end do
  else
    rc = hp3d_gpu_celem(space, norder, nori_e, nori_f, xnod, ngeo, coef, load, aloc, ld, bloc, &
                        aschur, bschur)
    call hp3d_gpu_check(rc, 'hp3d_gpu_celem')
    write(*, '(a,1x,i0)') 'rc', rc
    do j = 1, ni
      do i = 1, ni
        write(*, '(es25.17e3)') aloc(i, j)
      end do
    end do
    do i = 1, ni
      write(*, '(es25.17e3)') bloc(i)
    end do
    do j = 1, ni
      do i = 1, nb
        write(*, '(es25.17e3)') aschur(i, j)
      end do
    end do
    do i = 1, nb
      write(*, '(es25.17e3)') bschur(i)
    end do
  end if
  call hp3d_gpu_check(hp3d_gpu_finalize(), 'hp3d_gpu_finalize')
end program elem_driver
