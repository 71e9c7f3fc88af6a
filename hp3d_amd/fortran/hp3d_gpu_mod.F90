! hp3d_gpu_mod.F90 -- ISO_C_BINDING interface of libhp3d_gpu for Fortran hosts.
!
! An hp3D user `elem` routine (problems/POISSON/GALERKIN/elem.F90:20-55) that
! uses this module replaces its call to elem_opt (elem.F90:49-50) by
! hp3d_gpu_elem on the element's ALOC(1,1)%array / BLOC(1)%array, and
! celem_systemI (src/constrs/celem_systemI.F90:523-534) may replace the pair
! elem + stc_fwd_wrapper by hp3d_gpu_celem.  The batched entries take device
! pointers (type(c_ptr)) and a HIP stream.
!
! Every interface mirrors a declaration of include/hp3d_gpu.h.  Arrays are
! passed by address (assumed size), scalars by value, exactly as the C
! prototypes expect.
module hp3d_gpu_mod
  use iso_c_binding, only: c_int, c_long_long, c_double, c_size_t, c_ptr, c_char, &
                           c_null_char, c_f_pointer, c_associated
  implicit none

  integer(c_int), parameter :: HP3D_MAXP = 6
  integer(c_int), parameter :: HP3D_OK = 0, HP3D_NEG_JACOBIAN = 1
  integer(c_int), parameter :: HP3D_E_ORDER = -1, HP3D_E_NOIFACE = -3, HP3D_E_FACTOR = -4
  integer(c_int), parameter :: HP3D_E_ARG = -5, HP3D_E_HIP = -6, HP3D_E_INIT = -7
  integer(c_int), parameter :: HP3D_H1 = 0, HP3D_HCURL = 1
  integer(c_int), parameter :: HP3D_FP64 = 0, HP3D_FP32 = 1, HP3D_BF16 = 2
  integer(c_int), parameter :: HP3D_LOAD_ZERO = 0, HP3D_LOAD_SINSINSIN = 1, HP3D_LOAD_CONV = 2
  integer(c_int), parameter :: HP3D_LOAD_ONE = 3, HP3D_LOAD_POLY2 = 4, HP3D_LOAD_XYZ = 5
  integer(c_int), parameter :: HP3D_LOAD_CURL_SIN = 10
  integer(c_int), parameter :: HP3D_FIELD_ERROR = 0, HP3D_FIELD_ADJOINT = 1, HP3D_FIELD_L2 = 2
  integer(c_int), parameter :: HP3D_OPT_PERSISTENT = 1
  integer(c_size_t), parameter :: HP3D_SCRATCH_BYTES = 134217728_c_size_t

  interface
    integer(c_int) function hp3d_gpu_init(device, max_ndof) bind(C, name="hp3d_gpu_init")
      import :: c_int
      integer(c_int), value :: device, max_ndof
    end function

    integer(c_int) function hp3d_gpu_finalize() bind(C, name="hp3d_gpu_finalize")
      import :: c_int
    end function

    type(c_ptr) function hp3d_gpu_last_error() bind(C, name="hp3d_gpu_last_error")
      import :: c_ptr
    end function

    integer(c_int) function hp3d_gpu_set_option(option, value) bind(C, name="hp3d_gpu_set_option")
      import :: c_int
      integer(c_int), value :: option, value
    end function

    integer(c_int) function hp3d_gpu_ndof(space, norder, nrdof, nrdofb) &
        bind(C, name="hp3d_gpu_ndof")
      import :: c_int
      integer(c_int), value :: space
      integer(c_int), intent(in) :: norder(19)
      integer(c_int), intent(out) :: nrdof, nrdofb
    end function

    integer(c_int) function hp3d_gpu_elem(space, norder, nori_e, nori_f, xnod, ngeo, coef, &
        load, aloc, ld_aloc, bloc) bind(C, name="hp3d_gpu_elem")
      import :: c_int, c_double
      integer(c_int), value :: space, ngeo, load, ld_aloc
      integer(c_int), intent(in) :: norder(19), nori_e(12), nori_f(6)
      real(c_double), intent(in) :: xnod(*)
      real(c_double), value :: coef
      real(c_double), intent(inout) :: aloc(*), bloc(*)
    end function

    integer(c_int) function hp3d_gpu_stc_fwd(ni, nb, aloc, ld_aloc, bloc, aschur, bschur) &
        bind(C, name="hp3d_gpu_stc_fwd")
      import :: c_int, c_double
      integer(c_int), value :: ni, nb, ld_aloc
      real(c_double), intent(inout) :: aloc(*), bloc(*)
      real(c_double), intent(out) :: aschur(*), bschur(*)
    end function

    integer(c_int) function hp3d_gpu_celem(space, norder, nori_e, nori_f, xnod, ngeo, coef, &
        load, aloc, ld_aloc, bloc, aschur, bschur) bind(C, name="hp3d_gpu_celem")
      import :: c_int, c_double
      integer(c_int), value :: space, ngeo, load, ld_aloc
      integer(c_int), intent(in) :: norder(19), nori_e(12), nori_f(6)
      real(c_double), intent(in) :: xnod(*)
      real(c_double), value :: coef
      real(c_double), intent(inout) :: aloc(*), bloc(*)
      real(c_double), intent(out) :: aschur(*), bschur(*)
    end function

    integer(c_int) function hp3d_gpu_elem_batch(space, precision, nel, norder, nori_e, nori_f, &
        xnod, ngeo, coef, load, off_k, ld_k, kout, off_v, bout, workspace, status, max_ndof, &
        stream) bind(C, name="hp3d_gpu_elem_batch")
      import :: c_int, c_ptr
      integer(c_int), value :: space, precision, nel, ngeo, load, ld_k, max_ndof
      type(c_ptr), value :: norder, nori_e, nori_f, xnod, coef, off_k, kout, off_v, bout
      type(c_ptr), value :: workspace, status, stream
    end function

    integer(c_size_t) function hp3d_gpu_elem_workspace(space, precision, nel, max_ndof) &
        bind(C, name="hp3d_gpu_elem_workspace")
      import :: c_int, c_size_t
      integer(c_int), value :: space, precision, nel, max_ndof
    end function

    integer(c_size_t) function hp3d_gpu_celem_workspace(space, precision, nel, max_ndof, &
        max_nbubble) bind(C, name="hp3d_gpu_celem_workspace")
      import :: c_int, c_size_t
      integer(c_int), value :: space, precision, nel, max_ndof, max_nbubble
    end function

    integer(c_int) function hp3d_gpu_stc_batch(space, precision, nel, norder, off_k, ld_k, kin, &
        off_v, bin, off_c, ld_c, aii, bi, off_s, aschur, bschur, workspace, status, &
        max_ndof, max_nbubble, stream) bind(C, name="hp3d_gpu_stc_batch")
      import :: c_int, c_ptr
      integer(c_int), value :: space, precision, nel, ld_k, ld_c, max_ndof, max_nbubble
      type(c_ptr), value :: norder, off_k, kin, off_v, bin, off_c, aii, bi, off_s
      type(c_ptr), value :: aschur, bschur, workspace, status, stream
    end function

    integer(c_size_t) function hp3d_gpu_stc_workspace(nel, max_ndof, max_nbubble) &
        bind(C, name="hp3d_gpu_stc_workspace")
      import :: c_int, c_size_t
      integer(c_int), value :: nel, max_ndof, max_nbubble
    end function

    integer(c_int) function hp3d_gpu_celem_batch(space, precision, nel, norder, nori_e, nori_f, &
        xnod, ngeo, coef, load, off_k, ld_k, kout, off_v, bout, off_c, aii, bi, off_s, &
        aschur, bschur, workspace, status, max_ndof, max_nbubble, stream) &
        bind(C, name="hp3d_gpu_celem_batch")
      import :: c_int, c_ptr
      integer(c_int), value :: space, precision, nel, ngeo, load, ld_k, max_ndof, max_nbubble
      type(c_ptr), value :: norder, nori_e, nori_f, xnod, coef, off_k, kout, off_v, bout
      type(c_ptr), value :: off_c, aii, bi, off_s, aschur, bschur, workspace, status, stream
    end function

    integer(c_int) function hp3d_gpu_celem_batch_noalloc(space, precision, nel, norder, nori_e, &
        nori_f, xnod, ngeo, coef, load, off_k, ld_k, kout, off_v, bout, off_c, aii, bi, off_s, &
        aschur, bschur, status, max_ndof, max_nbubble, stream) &
        bind(C, name="hp3d_gpu_celem_batch_noalloc")
      import :: c_int, c_ptr
      integer(c_int), value :: space, precision, nel, ngeo, load, ld_k, max_ndof, max_nbubble
      type(c_ptr), value :: norder, nori_e, nori_f, xnod, coef, off_k, kout, off_v, bout
      type(c_ptr), value :: off_c, aii, bi, off_s, aschur, bschur, status, stream
    end function

    integer(c_int) function hp3d_gpu_dirichlet_lift(space, nel, norder, off_c, aii, off_v, ud, &
        bi, stream) bind(C, name="hp3d_gpu_dirichlet_lift")
      import :: c_int, c_ptr
      integer(c_int), value :: space, nel
      type(c_ptr), value :: norder, off_c, aii, off_v, ud, bi, stream
    end function

    integer(c_int) function hp3d_gpu_shape(space, norder, nori_e, nori_f, npts, xi, val, grad) &
        bind(C, name="hp3d_gpu_shape")
      import :: c_int, c_double
      integer(c_int), value :: space, npts
      integer(c_int), intent(in) :: norder(19), nori_e(12), nori_f(6)
      real(c_double), intent(in) :: xi(*)
      real(c_double), intent(out) :: val(*), grad(*)
    end function

    integer(c_int) function hp3d_gpu_stc_bwd_batch(space, nel, norder, off_s, aschur, off_v, &
        bschur, x, stream) bind(C, name="hp3d_gpu_stc_bwd_batch")
      import :: c_int, c_ptr
      integer(c_int), value :: space, nel
      type(c_ptr), value :: norder, off_s, aschur, off_v, bschur, x, stream
    end function

    integer(c_int) function hp3d_gpu_field_batch(space, mode, nel, norder, nori_e, nori_f, xnod, &
        ngeo, coef, off_v, u, v, exact, out, r, status, stream) bind(C, name="hp3d_gpu_field_batch")
      import :: c_int, c_ptr
      integer(c_int), value :: space, mode, nel, ngeo, exact
      type(c_ptr), value :: norder, nori_e, nori_f, xnod, coef, off_v, u, v, out, r, status, stream
    end function

    integer(c_int) function hp3d_gpu_gather_sum(nout, ptr, src, in, out, stream) &
        bind(C, name="hp3d_gpu_gather_sum")
      import :: c_int, c_long_long, c_ptr
      integer(c_long_long), value :: nout
      type(c_ptr), value :: ptr, src, in, out, stream
    end function

    integer(c_int) function hp3d_gpu_csr_spmv(nrows, indptr, indices, vals, x, y, stream) &
        bind(C, name="hp3d_gpu_csr_spmv")
      import :: c_int, c_long_long, c_ptr
      integer(c_long_long), value :: nrows
      type(c_ptr), value :: indptr, indices, vals, x, y, stream
    end function
  end interface

contains

  ! The reference stops on fatal conditions (order above MAXP, failed
  ! factorisation) and only warns on a negative Jacobian (geom3D.F90:93-110):
  ! map the return codes the same way.
  subroutine hp3d_gpu_check(rc, where)
    integer(c_int), intent(in) :: rc
    character(len=*), intent(in) :: where
    if (rc == HP3D_NEG_JACOBIAN) then
      write(*, '(a,a)') 'hp3d_gpu: negative Jacobian in ', where
    else if (rc < 0) then
      write(*, '(a,a,a,i0,a,a)') 'hp3d_gpu: ', where, ' failed (', rc, ') ', &
        trim(hp3d_gpu_error_string())
      stop 1
    end if
  end subroutine

  function hp3d_gpu_error_string() result(s)
    character(len=256) :: s
    character(kind=c_char), pointer :: cs(:)
    type(c_ptr) :: p
    integer :: i
    s = ' '
    p = hp3d_gpu_last_error()
    if (.not. c_associated(p)) return
    call c_f_pointer(p, cs, [256])
    do i = 1, 256
      if (cs(i) == c_null_char) exit
      s(i:i) = cs(i)
    end do
  end function

end module hp3d_gpu_mod
