/*
 * hp3d_oracle.h -- CPU restatement of hp3D's hexahedral element path.
 *
 * TEST INFRASTRUCTURE ONLY.  This library is the correctness checker for the
 * HIP path in hp3d_amd/ and the timed CPU baseline of bench.py.  Only tests/,
 * __graft_entry__.smoke() and bench.py's cpu_baseline leg may load it; the
 * product path never calls it.
 *
 * Conventions follow the reference (Fortran, column-major, 1-based codes):
 *   norder[19]  : 12 edge orders, 6 face orders coded h*10+v, interior x*100+y*10+z
 *   nori_e[12]  : edge orientations in {0,1}
 *   nori_f[6]   : face orientations in {0..7}
 *   xnod        : geometry dofs, 3 x ngeo column-major (first 8 = vertices)
 *   matrices    : column-major with explicit leading dimension
 *
 * Pinning status of each function is recorded in the header of hp3d_oracle.c.
 */
#ifndef HP3D_ORACLE_H
#define HP3D_ORACLE_H

#ifdef __cplusplus
extern "C" {
#endif

#define ORC_MAXP 6
#define ORC_MAXBRICKH 343
#define ORC_MAXBRICKE 882
#define ORC_MAX_NINT3 392

/* load selectors (the user getf routine) */
#define ORC_LOAD_ZERO 0      /* f = 0 */
#define ORC_LOAD_SINSINSIN 1 /* f = 3 pi^2 sin(pi x) sin(pi y) sin(pi z) */
#define ORC_LOAD_CONV 2      /* f = -2 sin(pi x) sinh(pi y)  (u = sin(pi x) sinh(pi y) z^2) */
#define ORC_LOAD_ONE 3       /* f = 1 */
#define ORC_LOAD_POLY2 4     /* f = -Laplace(u), u = x(1-x) y(1-y) z(1-z) */
#define ORC_LOAD_XYZ 5       /* f = 0, u = x y z */
#define ORC_LOAD_CURL_SIN 10 /* J = (1+2 pi^2) sin(pi y) sin(pi z) e_x */

void orc_decod(int nick, int mod, int n, int *narr);

/* Legendre family (Polynomials.F90) */
void orc_poly_legendre(double x, double t, int nord, double *p);
void orc_poly_ilegendre(double x, double t, int nord, int idec, double *l, double *p, double *r);

/* 1D Gauss-Legendre rule on [0,1] with n points (gauss_quadrature.F90) */
void orc_gauss1d(int n, double *x, double *w);
/* brick quadrature (set_3D_int.F90) */
void orc_set_3d_int(const int *norder, const int *nori_f, int *nint, double *xiloc, double *waloc);
/* set_3Dint (no orientation swap) with a given INTEGRATION */
void orc_set_3dint(const int *norder, int integration, int *nint, double *xiloc, double *waloc);
void orc_find_order_loc(const int *norder, const int *nori_f, int *norder_loc);

/* element dof counts (celndof.F90 / element_data.F90 ndof_nod, hexahedron) */
int orc_ndof_h1(const int *norder);
int orc_ndof_hcurl(const int *norder);
int orc_nbubble_h1(int nord_mdl);
int orc_nbubble_hcurl(int nord_mdl);
int orc_checkorder(const int *norder);

/* shape functions (Hexahedron.F90) */
int orc_shape3dh_hexa(const double *xi, const int *norder, const int *nori_e, const int *nori_f,
                      int *nrdof, double *shap, double *grad);
int orc_shape3de_hexa(const double *xi, const int *norder, const int *nori_e, const int *nori_f,
                      int *nrdof, double *shap, double *curl);

/* geometry (geom.F90, geom3D.F90) */
int orc_geom(const double *dxdxi, double *dxidx, double *rjac);
int orc_geom3d(const double *xnod, int ngeo, const double *shap, const double *grad,
               double *x, double *dxdxi, double *dxidx, double *rjac);

/* load (getf) */
void orc_getf_h1(int load, const double *x, double *f);
void orc_getf_hcurl(int load, const double *x, double *j);

/* element routines (POISSON/GALERKIN/elem_opt.F90, MAXWELL/GALERKIN/elem_opt.F90 real form) */
int orc_elem_h1(const int *norder, const int *nori_e, const int *nori_f, const double *xnod,
                int ngeo, double kappa, int load, double *zaloc, int ld, double *zbloc);
int orc_elem_hcurl(const int *norder, const int *nori_e, const int *nori_f, const double *xnod,
                   int ngeo, double mu, double eps, int load, double *zaloc, int ld,
                   double *zbloc);

/* static condensation (stc.F90): in place on aloc/bloc, Schur factors out */
int orc_stc_fwd_gen(int ni, int nb, double *aloc, int ld, double *bloc, double *aschur,
                    double *bschur);
int orc_stc_fwd_herm(int ni, int nb, double *aloc, int ld, double *bloc, double *aschur,
                     double *bschur);

/* back substitution (stc.F90 stc_bwd): xb = bschur - aschur * xi */
void orc_stc_bwd(int ni, int nb, const double *aschur, const double *bschur, const double *xi,
                 double *xb);

/* manufactured solutions (selector = the load selector of the same problem) */
void orc_exact_h1(int sel, const double *x, double *u, double *du);
void orc_exact_hcurl(int sel, const double *x, double *e, double *curl);
/* element_error (compute_error.F90:226-400): squared error and exact norm */
int orc_element_error(int space, const int *norder, const int *nori_e, const int *nori_f,
                      const double *xnod, int ngeo, const double *dof, int sel, double *err,
                      double *rnorm);
/* element mass matrix (H1 or H(curl)), element quadrature */
int orc_elem_mass(int space, const int *norder, const int *nori_e, const int *nori_f,
                  const double *xnod, int ngeo, double *m, int ld);

/*
 * Batched element loop, the mumps_sc Step 2 OpenMP loop (omp_mumps/mumps_sc.F90:224-342)
 * restated for a batch of hexahedra that share no state:
 *   space     0 = H1 Poisson, 1 = H(curl) curl-curl + mass
 *   norder    [nel][19], nori_e [nel][12], nori_f [nel][6]
 *   xnod      [nel][3*ngeo]
 *   coef      [nel] kappa_e (H1) or eps_e (H(curl), mu = 1)
 *   kout      [nel][ldk*ldk] element matrices (may be NULL), ldk >= max N
 *   bout      [nel][ldk]     element loads    (may be NULL)
 *   aii       [nel][ldk*ldk] condensed matrices (may be NULL), leading dim ldk
 *   bi        [nel][ldk]
 *   aschur    [nel][ldk*ldk] nb x ni, leading dim nb
 *   bschur    [nel][ldk]
 *   herm      0 = LU (stc_fwd_gen, the mumps_sc('G') default), 1 = Cholesky
 * Returns 0, or the first nonzero element status.
 */
int orc_element_loop(int space, int nel, const int *norder, const int *nori_e, const int *nori_f,
                     const double *xnod, int ngeo, const double *coef, int load, int ldk,
                     double *kout, double *bout, double *aii, double *bi, double *aschur,
                     double *bschur, int herm, int nthreads);
/* host BLAS/LAPACK (Fortran ABI, 32-bit integers) for orc_element_loop_blas */
int orc_set_blas(void *dsyrk, void *dgemm, void *dgetrf, void *dgetrs);
int orc_element_loop_blas(int space, int nel, const int *norder, const int *nori_e,
                          const int *nori_f, const double *xnod, int ngeo, const double *coef,
                          int load, int ldk, double *kout, double *bout, double *aii, double *bi,
                          double *aschur, double *bschur, int nthreads);

#ifdef __cplusplus
}
#endif
#endif
