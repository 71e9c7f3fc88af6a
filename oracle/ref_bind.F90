!----------------------------------------------------------------------
! ref_bind.F90 -- C-callable entry points onto the hp3D reference routines
! compiled from the reference sources by oracle/Makefile into oracle/_ref.
!
! TEST INFRASTRUCTURE ONLY: used by tests/ to pin the C restatement in
! oracle/hp3d_oracle.c against the reference's own shape functions,
! Legendre recurrences, Gauss tables and Jacobian inverse.  Each wrapper
! only forwards its arguments; nothing here restates reference arithmetic.
!----------------------------------------------------------------------
module ref_bind
   use iso_c_binding
   implicit none
   integer, parameter :: RB_MAXP = 6
contains

!..the Nsize(1) that checkorder (CheckOrder.F90:60-152) computes for a brick:
!  the largest order over edges, face directions and interior directions
   integer function max_order(Nord)
      integer, intent(in) :: Nord(19)
      integer :: f, hv(2), b(3)
      max_order = maxval(Nord(1:12))
      do f = 1, 6
         call decod(Nord(12+f), 10, 2, hv)
         max_order = max(max_order, hv(1), hv(2))
      enddo
      call decod(Nord(19), 10, 3, b)
      max_order = max(max_order, b(1), b(2), b(3))
   end function max_order

!..shape3DHHexa: trunk/src/element/shape_1/Hexahedron.F90:33
   subroutine ref_shape3dh_hexa(Xi, Norder, NoriE, NoriF, Nrdof, ShapH, GradH) &
         bind(C, name="ref_shape3dh_hexa")
      real(c_double), intent(in) :: Xi(3)
      integer(c_int), intent(in) :: Norder(19), NoriE(12), NoriF(6)
      integer(c_int), intent(out) :: Nrdof
      real(c_double), intent(out) :: ShapH((RB_MAXP+1)**3), GradH(3,(RB_MAXP+1)**3)
      integer :: nord(19), noe(12), nof(6), nsize(2), nrdofh
      nord = Norder; noe = NoriE; nof = NoriF
      nsize(1) = max_order(nord); nsize(2) = (nsize(1)+1)**3
      ShapH = 0.d0; GradH = 0.d0
      call shape3DHHexa(Xi, nord, noe, nof, nsize, nrdofh, ShapH, GradH)
      Nrdof = nrdofh
   end subroutine ref_shape3dh_hexa

!..shape3DEHexa: trunk/src/element/shape_1/Hexahedron.F90:240
   subroutine ref_shape3de_hexa(Xi, Norder, NoriE, NoriF, Nrdof, ShapE, CurlE) &
         bind(C, name="ref_shape3de_hexa")
      real(c_double), intent(in) :: Xi(3)
      integer(c_int), intent(in) :: Norder(19), NoriE(12), NoriF(6)
      integer(c_int), intent(out) :: Nrdof
      real(c_double), intent(out) :: ShapE(3,3*RB_MAXP*(RB_MAXP+1)**2), &
                                     CurlE(3,3*RB_MAXP*(RB_MAXP+1)**2)
      integer :: nord(19), noe(12), nof(6), nsize(2), nrdofe
      nord = Norder; noe = NoriE; nof = NoriF
      nsize(1) = max_order(nord); nsize(2) = 3*nsize(1)*(nsize(1)+1)**2
      ShapE = 0.d0; CurlE = 0.d0
      call shape3DEHexa(Xi, nord, noe, nof, nsize, nrdofe, ShapE, CurlE)
      Nrdof = nrdofe
   end subroutine ref_shape3de_hexa

!..PolyLegendre: trunk/src/element/shape_1/Polynomials.F90:34
   subroutine ref_poly_legendre(X, T, Nord, P) bind(C, name="ref_poly_legendre")
      real(c_double), value :: X, T
      integer(c_int), value :: Nord
      real(c_double), intent(out) :: P(0:Nord)
      real(8) :: xx, tt
      integer :: n
      xx = X; tt = T; n = Nord
      call PolyLegendre(xx, tt, n, P)
   end subroutine ref_poly_legendre

!..PolyILegendre (Idec=.true.): trunk/src/element/shape_1/Polynomials.F90:109
   subroutine ref_poly_ilegendre(X, T, Nord, L, P) bind(C, name="ref_poly_ilegendre")
      real(c_double), value :: X, T
      integer(c_int), value :: Nord
      real(c_double), intent(out) :: L(2:Nord), P(1:Nord-1)
      real(8) :: xx, tt, r(Nord-1)
      integer :: n
      logical :: idec
      xx = X; tt = T; n = Nord; idec = .true.
      call PolyILegendre(xx, tt, n, idec, L, P, r)
   end subroutine ref_poly_ilegendre

!..XIGAUS1/WAGAUS1 after init_gauss_quadrature:
!  trunk/src/element/quadrature/gauss_quadrature.F90:24-26,518-752
   subroutine ref_gauss1d(N, X, W) bind(C, name="ref_gauss1d")
      use gauss_quadrature, only : INITIALIZED, XIGAUS1, WAGAUS1
      integer(c_int), value :: N
      real(c_double), intent(out) :: X(N), W(N)
      if (.not. INITIALIZED) call init_gauss_quadrature
      X(1:N) = XIGAUS1(1:N, N)
      W(1:N) = WAGAUS1(1:N, N)
   end subroutine ref_gauss1d

!..geom: trunk/src/element/util/geom.F90:30
   subroutine ref_geom(Dxdxi, Dxidx, Rjac, Iflag) bind(C, name="ref_geom")
      real(c_double), intent(in) :: Dxdxi(3,3)
      real(c_double), intent(out) :: Dxidx(3,3), Rjac
      integer(c_int), intent(out) :: Iflag
      integer :: ifl
      real(8) :: r
      call geom(Dxdxi, Dxidx, r, ifl)
      Rjac = r; Iflag = ifl
   end subroutine ref_geom

end module ref_bind
