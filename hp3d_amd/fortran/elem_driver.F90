! elem_driver.F90 -- a Fortran host calling libhp3d_gpu through hp3d_gpu_mod,
! the way an hp3D `elem` would: element codes and geometry dofs in, ALOC /
! BLOC written in place with the caller's leading dimension (MAXDOFS in
! mumps_sc.F90:143-145).
!
! Modes 4 and 5 drive the batched entry the way a Fortran element loop over
! the active elements would (mumps_sc.F90:224-342): the host packs the codes,
! geometry and offsets of nel elements, stages them in device memory with the
! HIP runtime, calls hp3d_gpu_celem_batch (mode 4, with its workspace) or
! hp3d_gpu_celem_batch_noalloc (mode 5) and copies the condensed systems back.
!
! stdin (list-directed):
!   mode            1 = counts only (no device), 2 = elem, 3 = celem,
!                   4 = celem_batch, 5 = celem_batch_noalloc
!   modes 1-3:
!     space ld        energy space and leading dimension of ALOC
!     norder(19) nori_e(12) nori_f(6)
!     ngeo            then xnod(3, ngeo)
!     coef load
!   modes 4-5:
!     space precision nel ngeo load
!     nel times: norder(19) nori_e(12) nori_f(6), xnod(3, ngeo), coef
! stdout: "counts N nb", then for mode 2 "rc", K(1:N,1:N) column by column and
! b(1:N); for mode 3 "rc", Aii(1:ni,1:ni), Bi(1:ni), ASchur(nb,ni), BSchur(nb).
! Modes 4-5: "batch nel", status(1:nel), then the packed arrays Aii (ni_e x
! ni_e per element), Bi and BSchur (N_e slots per element) and ASchur (nb_e x
! ni_e per element), in that order, elements one after the other.
module hip_host
  use iso_c_binding, only: c_int, c_size_t, c_ptr, c_null_ptr, c_associated
  implicit none
  integer(c_int), parameter :: H2D = 1, D2H = 2
  interface
    integer(c_int) function hipMalloc(ptr, bytes) bind(C, name="hipMalloc")
      import :: c_int, c_size_t, c_ptr
      type(c_ptr) :: ptr
      integer(c_size_t), value :: bytes
    end function
    integer(c_int) function hipFree(ptr) bind(C, name="hipFree")
      import :: c_int, c_ptr
      type(c_ptr), value :: ptr
    end function
    integer(c_int) function hipMemcpy(dst, src, bytes, kind) bind(C, name="hipMemcpy")
      import :: c_int, c_size_t, c_ptr
      type(c_ptr), value :: dst, src
      integer(c_size_t), value :: bytes
      integer(c_int), value :: kind
    end function
    integer(c_int) function hipDeviceSynchronize() bind(C, name="hipDeviceSynchronize")
      import :: c_int
    end function
  end interface
contains
  subroutine hip_ok(rc, where)
    integer(c_int), intent(in) :: rc
    character(len=*), intent(in) :: where
    if (rc /= 0) then
      write(*, '(a,a,a,i0)') 'hip: ', where, ' failed ', rc
      stop 1
    end if
  end subroutine

  ! device copy of `bytes` bytes at host address src (c_null_ptr: left unset)
  function to_dev(src, bytes) result(d)
    type(c_ptr), intent(in) :: src
    integer(c_size_t), intent(in) :: bytes
    type(c_ptr) :: d
    d = c_null_ptr
    call hip_ok(hipMalloc(d, max(bytes, 256_c_size_t)), 'hipMalloc')
    if (c_associated(src) .and. bytes > 0) call hip_ok(hipMemcpy(d, src, bytes, H2D), 'hipMemcpy')
  end function
end module hip_host

program elem_driver
  use iso_c_binding, only: c_int, c_double
  use hp3d_gpu_mod
  implicit none
  integer(c_int) :: mode, space, ld, norder(19), nori_e(12), nori_f(6), ngeo, load
  integer(c_int) :: n, nb, ni, rc, i, j
  real(c_double) :: coef
  real(c_double), allocatable :: xnod(:,:), aloc(:,:), bloc(:), aschur(:,:), bschur(:)

  read(*, *) mode
  if (mode >= 4) then
    call run_batch(mode == 5)
    stop
  end if
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

contains

  subroutine run_batch(noalloc)
    use iso_c_binding, only: c_long_long, c_size_t, c_ptr, c_null_ptr, c_loc, c_sizeof
    use hip_host
    logical, intent(in) :: noalloc
    integer(c_int) :: prec, nel, e, maxn, maxb, esz
    integer(c_int), allocatable, target :: no(:,:), ne(:,:), nf(:,:), status(:)
    real(c_double), allocatable, target :: xn(:,:,:), cf(:), aii(:), bi(:), as(:), bs(:)
    integer(c_long_long), allocatable, target :: off_k(:), off_v(:), off_c(:), off_s(:)
    integer(c_long_long) :: sk, sv, sc, ss
    integer(c_size_t) :: i8, i4, wsb
    type(c_ptr) :: d_no, d_ne, d_nf, d_x, d_cf, d_ok, d_ov, d_oc, d_os, d_st
    type(c_ptr) :: d_k, d_b, d_aii, d_bi, d_as, d_bs, d_ws

    read(*, *) space, prec, nel, ngeo, load
    allocate(no(19, nel), ne(12, nel), nf(6, nel), xn(3, ngeo, nel), cf(nel), status(nel))
    allocate(off_k(nel), off_v(nel), off_c(nel), off_s(nel))
    do e = 1, nel
      read(*, *) no(:, e), ne(:, e), nf(:, e)
      read(*, *) xn(:, :, e)
      read(*, *) cf(e)
    end do
    ! packed offsets: K (N x N), vectors (N), Aii (ni x ni), ASchur (nb x ni)
    sk = 0; sv = 0; sc = 0; ss = 0; maxn = 0; maxb = 0
    do e = 1, nel
      call hp3d_gpu_check(hp3d_gpu_ndof(space, no(:, e), n, nb), 'hp3d_gpu_ndof')
      ni = n - nb
      off_k(e) = sk; off_v(e) = sv; off_c(e) = sc; off_s(e) = ss
      sk = sk + int(n, c_long_long) * n
      sv = sv + n
      sc = sc + int(ni, c_long_long) * ni
      ss = ss + int(nb, c_long_long) * ni
      maxn = max(maxn, n); maxb = max(maxb, nb)
    end do
    allocate(aii(max(sc, 1_c_long_long)), bi(max(sv, 1_c_long_long)))
    allocate(as(max(ss, 1_c_long_long)), bs(max(sv, 1_c_long_long)))
    status = 0
    i8 = 8; i4 = 4
    esz = merge(8, 4, prec == HP3D_FP64)

    call hp3d_gpu_check(hp3d_gpu_init(0_c_int, 0_c_int), 'hp3d_gpu_init')
    d_no = to_dev(c_loc(no), i4 * 19 * nel)
    d_ne = to_dev(c_loc(ne), i4 * 12 * nel)
    d_nf = to_dev(c_loc(nf), i4 * 6 * nel)
    d_x = to_dev(c_loc(xn), i8 * 3 * ngeo * nel)
    d_cf = to_dev(c_loc(cf), i8 * nel)
    d_ok = to_dev(c_loc(off_k), i8 * nel)
    d_ov = to_dev(c_loc(off_v), i8 * nel)
    d_oc = to_dev(c_loc(off_c), i8 * nel)
    d_os = to_dev(c_loc(off_s), i8 * nel)
    d_st = to_dev(c_loc(status), i4 * nel)
    d_k = to_dev(c_null_ptr, int(esz, c_size_t) * sk)
    d_b = to_dev(c_null_ptr, i8 * sv)
    d_aii = to_dev(c_null_ptr, i8 * sc)
    d_bi = to_dev(c_null_ptr, i8 * sv)
    d_as = to_dev(c_null_ptr, i8 * ss)
    d_bs = to_dev(c_null_ptr, i8 * sv)
    if (noalloc) then
      rc = hp3d_gpu_celem_batch_noalloc(space, prec, nel, d_no, d_ne, d_nf, d_x, ngeo, d_cf, &
                                        load, d_ok, 0_c_int, d_k, d_ov, d_b, d_oc, d_aii, d_bi, &
                                        d_os, d_as, d_bs, d_st, maxn, maxb, c_null_ptr)
      call hp3d_gpu_check(rc, 'hp3d_gpu_celem_batch_noalloc')
    else
      wsb = hp3d_gpu_celem_workspace(space, prec, nel, maxn, maxb)
      d_ws = c_null_ptr
      if (wsb > 0) d_ws = to_dev(c_null_ptr, wsb)
      rc = hp3d_gpu_celem_batch(space, prec, nel, d_no, d_ne, d_nf, d_x, ngeo, d_cf, load, d_ok, &
                                0_c_int, d_k, d_ov, d_b, d_oc, d_aii, d_bi, d_os, d_as, d_bs, &
                                d_ws, d_st, maxn, maxb, c_null_ptr)
      call hp3d_gpu_check(rc, 'hp3d_gpu_celem_batch')
    end if
    call hip_ok(hipDeviceSynchronize(), 'hipDeviceSynchronize')
    call hip_ok(hipMemcpy(c_loc(status), d_st, i4 * nel, D2H), 'hipMemcpy')
    if (sc > 0) call hip_ok(hipMemcpy(c_loc(aii), d_aii, i8 * sc, D2H), 'hipMemcpy')
    call hip_ok(hipMemcpy(c_loc(bi), d_bi, i8 * sv, D2H), 'hipMemcpy')
    if (ss > 0) call hip_ok(hipMemcpy(c_loc(as), d_as, i8 * ss, D2H), 'hipMemcpy')
    call hip_ok(hipMemcpy(c_loc(bs), d_bs, i8 * sv, D2H), 'hipMemcpy')
    write(*, '(a,1x,i0)') 'batch', nel
    do e = 1, nel
      write(*, '(i0)') status(e)
    end do
    write(*, '(es25.17e3)') aii(1:sc), bi(1:sv), bs(1:sv), as(1:ss)
    call hp3d_gpu_check(hp3d_gpu_finalize(), 'hp3d_gpu_finalize')
  end subroutine run_batch
end program elem_driver
