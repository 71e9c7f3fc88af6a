/*
 * hp3d_gpu.h -- C ABI of the MI355X (gfx950) element-computation library.
 *
 * Replaces, for hexahedral H1 and H(curl) elements with trilinear or
 * isoparametric geometry, the per-element path of hp3D:
 *   user elem / elem_opt   problems/POISSON/GALERKIN/elem.F90:20-55,
 *                          problems/POISSON/GALERKIN/elem_opt.F90:22-141,
 *                          problems/MAXWELL/GALERKIN/elem_opt.F90:22-158
 *   static condensation    src/modules/stc.F90:182-311 (stc_fwd_wrapper),
 *                          :338-416 (stc_fwd_herm), :443-509 (stc_fwd_gen)
 *   back substitution      src/modules/stc.F90:664-679 (stc_bwd)
 *   the OpenMP element loop src/solver/omp_mumps/mumps_sc.F90:224-342
 *
 * Conventions are hp3D's: integer codes as in find_order / find_orient
 * (norder[19]: 12 edge orders, 6 face orders h*10+v in face-node axes,
 * interior x*100+y*10+z; nori_e[12] in {0,1}; nori_f[6] in {0..7}),
 * geometry dofs xnod as 3 x ngeo column-major (ngeo = 8: vertices only,
 * i.e. a trilinear 'Linear' GMP block; ngeo = N: full isoparametric),
 * matrices column-major with an explicit leading dimension, element-local
 * dof order vertices, edges, faces, bubbles (interface first, bubbles last).
 *
 * Return codes (all entry points): 0 ok; 1 a negative Jacobian was seen
 * (geom.F90:70-71 sets Iflag and the reference continues); negative values
 * are fatal conditions on which the reference stops:
 *   HP3D_E_ORDER   order outside [1, MAXP]        (CheckOrder.F90:139-152)
 *   HP3D_E_FACTOR  bubble block not SPD           (stc.F90:360-396, 464-489)
 *   HP3D_E_NOIFACE no interface dof               (stc.F90:214-217)
 *   HP3D_E_ARG     bad argument (sizes, pointers, leading dimensions)
 *   HP3D_E_HIP     HIP runtime error
 *   HP3D_E_INIT    library not initialised / no device
 */
#ifndef HP3D_GPU_H
#define HP3D_GPU_H

#include <stddef.h>

#ifdef __cplusplus
extern "C" {
#endif

#define HP3D_MAXP 6

#define HP3D_OK 0
#define HP3D_NEG_JACOBIAN 1
#define HP3D_E_ORDER (-1)
#define HP3D_E_FACTOR (-4)
#define HP3D_E_NOIFACE (-3)
#define HP3D_E_ARG (-5)
#define HP3D_E_HIP (-6)
#define HP3D_E_INIT (-7)

/* energy spaces */
#define HP3D_H1 0    /* Poisson: (kappa grad u, grad v) */
#define HP3D_HCURL 1 /* (mu^-1 curl E, curl F) + (eps E, F), real form */

/* arithmetic of the B^T W B contraction; geometry, Jacobian inverse and the
 * bubble factorisation are always FP64 */
#define HP3D_FP64 0
#define HP3D_FP32 1
/* trial: operands rounded to bfloat16, FP32 accumulation (no carry), K stored
 * as FP32; the same kernels as FP32 (single-block and block-tiled), and the
 * condensation as for FP32 */
#define HP3D_BF16 2

/* load selectors (the user getf routine) */
#define HP3D_LOAD_ZERO 0
#define HP3D_LOAD_SINSINSIN 1 /* f = 3 pi^2 sin(pi x) sin(pi y) sin(pi z)     */
#define HP3D_LOAD_CONV 2      /* f = -2 sin(pi x) sinh(pi y)                 */
#define HP3D_LOAD_ONE 3       /* f = 1                                        */
#define HP3D_LOAD_POLY2 4     /* f = -Laplace(u), u = x(1-x) y(1-y) z(1-z)    */
#define HP3D_LOAD_XYZ 5       /* f = 0, u = x y z (test/poly_pois.F90 exact)  */
#define HP3D_LOAD_CURL_SIN 10 /* J = (1 + 2 pi^2) sin(pi y) sin(pi z) e_x     */

/* ------------------------------------------------------------ lifecycle */

/* Select the device and fix the size of the per-element entries' staging:
 * elements of up to max_ndof dofs (0 = the MAXP maximum, 882); a larger
 * element is rejected with HP3D_E_ARG.  Every host thread that calls a
 * per-element entry gets its own context (a stream, device staging and pinned
 * host staging of about 5 max_ndof^2 doubles), made on its first call after
 * init (the calling thread's now), with the library's device made current on
 * that thread; later calls on the thread allocate nothing.  Calls from
 * different threads therefore run concurrently on the GPU.  This replaces the
 * THREADPRIVATE workspace setup of mumps_sc Step 2 (omp_mumps/mumps_sc.F90:
 * 228-251).  hp3d_gpu_finalize frees every context; call it when no other
 * call is in flight. */
int hp3d_gpu_init(int device, int max_ndof);
int hp3d_gpu_finalize(void);
/* the calling thread's last error message (a HIP error string, or which
 * argument or state was refused) */
const char *hp3d_gpu_last_error(void);

/* Process-wide launch options, read when a batched launch is made (a
 * captured graph keeps the form it was captured with); not synchronised with
 * launches in flight on other threads.  No counterpart in the reference.
 *   HP3D_OPT_PERSISTENT  1: the single-block element kernel runs as a
 *                        persistent grid (2 workgroups per CU walking the
 *                        elements), 0: one workgroup per element (default) */
#define HP3D_OPT_PERSISTENT 1
int hp3d_gpu_set_option(int option, int value);

/* ------------------------------------------------------------- counts */

/* celndof / ndof_nod for a hexahedron (element/util/celndof.F90:24,
 * element_data.F90:808-883): total dofs, and bubble dofs of the middle node
 * (stc_get_nrdof, stc.F90:94-160). */
int hp3d_gpu_ndof(int space, const int norder[19], int *nrdof, int *nrdofb);

/* ------------------------------------------------ per-element drop-in */

/* elem_opt for one element, host pointers.  Writes the N x N stiffness into
 * aloc (leading dimension ld_aloc >= N, the caller's ALOC(1,1)%array) and the
 * load into bloc.  coef is kappa_e (H1) or eps_e (H(curl), mu = 1).
 * Replaces POISSON/GALERKIN/elem_opt.F90:22-141 and the real form of
 * MAXWELL/GALERKIN/elem_opt.F90:22-158. */
int hp3d_gpu_elem(int space, const int norder[19], const int nori_e[12], const int nori_f[6],
                  const double *xnod, int ngeo, double coef, int load, double *aloc,
                  int ld_aloc, double *bloc);

/* stc_fwd_wrapper for one element, host pointers: condenses the last nb of
 * ni+nb dofs of aloc/bloc in place (aloc(1:ni,1:ni) <- Aii - Aib Abb^-1 Abi,
 * bloc(1:ni) <- Bi - Aib Abb^-1 Bb) and returns ASchur = Abb^-1 Abi (nb x ni,
 * leading dimension nb) and BSchur = Abb^-1 Bb (either may be NULL).
 * Replaces stc.F90:182-311 with stc_fwd_herm semantics (Cholesky; Abb is SPD
 * for these forms, where stc_fwd_gen's LU gives the same factors' solution). */
int hp3d_gpu_stc_fwd(int ni, int nb, double *aloc, int ld_aloc, double *bloc, double *aschur,
                     double *bschur);

/* elem + stc_fwd_wrapper fused for one element (celem_systemI.F90:523-534). */
int hp3d_gpu_celem(int space, const int norder[19], const int nori_e[12], const int nori_f[6],
                   const double *xnod, int ngeo, double coef, int load, double *aloc,
                   int ld_aloc, double *bloc, double *aschur, double *bschur);

/* ----------------------------------------------- batched device entries */

/*
 * Element matrices and loads for nel elements, all pointers DEVICE pointers,
 * launched on `stream` (a hipStream_t, NULL = default stream).  Per element e:
 *   norder[19e..], nori_e[12e..], nori_f[6e..], xnod[3*ngeo*e..], coef[e]
 *   K at kout + off_k[e] (FP64 or FP32 by `precision`), leading dimension
 *     ld_k, or N_e when ld_k == 0 (packed)
 *   load at bout + off_v[e] (FP64; bout may be NULL)
 *   status[e] (may be NULL): 0, 1 (negative Jacobian) or HP3D_E_ORDER
 * max_ndof bounds N_e over the batch (it sizes the launch grid).
 * workspace: device memory of hp3d_gpu_elem_workspace(space, precision, nel,
 * max_ndof) bytes (the per-element quadrature-point geometry of the
 * block-tiled kernel, N_e > 128); may be NULL when that size is 0.
 * Replaces the element loop of mumps_sc Step 2 up to elem
 * (mumps_sc.F90:253-259 -> celem_systemI.F90:523).
 */
int hp3d_gpu_elem_batch(int space, int precision, int nel, const int *norder, const int *nori_e,
                        const int *nori_f, const double *xnod, int ngeo, const double *coef,
                        int load, const long long *off_k, int ld_k, void *kout,
                        const long long *off_v, double *bout, void *workspace, int *status,
                        int max_ndof, void *stream);
/* bytes of device workspace hp3d_gpu_elem_batch needs for nel elements of at
 * most max_ndof dofs (0 when the single-block kernel serves the batch).  No
 * counterpart in the reference: its elem computes the geometry point by point
 * inside the element loop (elem_opt.F90:101-122). */
size_t hp3d_gpu_elem_workspace(int space, int precision, int nel, int max_ndof);

/*
 * Static condensation of nel element systems already in device memory
 * (stc_fwd_wrapper + stc_fwd_herm, stc.F90:182-416), FP64 arithmetic:
 *   kin  at off_k[e], ld_k (0 = N_e), FP64 or FP32 by `precision`
 *   bin  at off_v[e]
 *   aii  at off_c[e], ld_c (0 = ni_e): condensed matrix (may alias kin when
 *        precision is FP64, off_c == off_k and ld_c == ld_k: in place as ALOC)
 *   bi   at off_v[e]: condensed load (may alias bin)
 *   aschur at off_s[e] (nb_e x ni_e, leading dimension nb_e), bschur at off_v[e]
 *   workspace: device memory of hp3d_gpu_stc_workspace(nel, max_ndof,
 *   max_nbubble) bytes (one slab per element: the bubble block's Cholesky
 *   factor in 16 x 16 tiles, its transformed copy and the inverted diagonal
 *   blocks; independent of max_ndof)
 *   max_ndof, max_nbubble bound N_e and nb_e over the batch (an element above
 *   them reports HP3D_E_ARG in status[e])
 *   status[e]: 0, a negative code, or 1 when status[e] already held 1 (the
 *   negative-Jacobian warning of the element pass that produced kin)
 */
int hp3d_gpu_stc_batch(int space, int precision, int nel, const int *norder,
                       const long long *off_k, int ld_k, const void *kin,
                       const long long *off_v, const double *bin, const long long *off_c,
                       int ld_c, double *aii, double *bi, const long long *off_s,
                       double *aschur, double *bschur, void *workspace, int *status,
                       int max_ndof, int max_nbubble, void *stream);
/* bytes of device workspace the condensation of nel elements with at most
 * max_ndof dofs and max_nbubble bubble dofs needs (0 when max_nbubble = 0) */
size_t hp3d_gpu_stc_workspace(int nel, int max_ndof, int max_nbubble);

/*
 * Element matrices, loads and static condensation in one call (the work of
 * celem_systemI.F90:523-534 for nel elements; DEVICE pointers).  Arguments as
 * in hp3d_gpu_elem_batch and hp3d_gpu_stc_batch; max_nbubble bounds nb_e.
 * When max_ndof <= 128 and max_nbubble <= 32 the condensation runs inside the
 * element kernel on the register-resident K (no second pass over HBM); kout
 * and bout may then be NULL (K and the full load are not written).  Otherwise
 * K and the load are written and condensed by a second kernel; kout, bout and
 * the hp3d_gpu_celem_workspace(space, precision, nel, max_ndof, max_nbubble)
 * workspace are then required.  The condensed matrix is packed (leading
 * dimension ni_e).
 */
int hp3d_gpu_celem_batch(int space, int precision, int nel, const int *norder, const int *nori_e,
                         const int *nori_f, const double *xnod, int ngeo, const double *coef,
                         int load, const long long *off_k, int ld_k, void *kout,
                         const long long *off_v, double *bout, const long long *off_c,
                         double *aii, double *bi, const long long *off_s, double *aschur,
                         double *bschur, void *workspace, int *status, int max_ndof,
                         int max_nbubble, void *stream);
/* bytes of device workspace hp3d_gpu_celem_batch needs: 0 for the fused path,
 * else the larger of the element pass's and the condensation's (they use the
 * same memory one after the other).  No counterpart in the reference (its
 * THREADPRIVATE workspaces are per thread, mumps_sc.F90:224-342). */
size_t hp3d_gpu_celem_workspace(int space, int precision, int nel, int max_ndof, int max_nbubble);

/*
 * hp3d_gpu_celem_batch without a workspace argument: allocates nothing (no
 * hipMalloc, no hipFree, no allocating library call), for callers that cannot
 * allocate device memory (code under HIP graph capture, serving loops,
 * harnesses that forbid it).  No counterpart in the reference.  The two-pass
 * bins use the library's own device scratch of HP3D_SCRATCH_BYTES (a static
 * device array, one per device, held until the process ends), filled chunk by
 * chunk of elements on `stream`; an element whose workspace exceeds it
 * (hp3d_gpu_celem_workspace(space, precision, 1, max_ndof, max_nbubble) >
 * HP3D_SCRATCH_BYTES) is rejected with HP3D_E_ARG.  The scratch is shared by
 * every call on a device and keeps the last call's contents: calls that
 * overlap on one device (other streams, other threads) race.  Everything else
 * as in hp3d_gpu_celem_batch.
 */
#define HP3D_SCRATCH_BYTES ((size_t)128 << 20)
int hp3d_gpu_celem_batch_noalloc(int space, int precision, int nel, const int *norder,
                                 const int *nori_e, const int *nori_f, const double *xnod, int ngeo,
                                 const double *coef, int load, const long long *off_k, int ld_k,
                                 void *kout, const long long *off_v, double *bout,
                                 const long long *off_c, double *aii, double *bi,
                                 const long long *off_s, double *aschur, double *bschur,
                                 int *status, int max_ndof, int max_nbubble, void *stream);

/*
 * Back substitution for nel elements (stc_bwd_wrapper + stc_bwd + stc_solout,
 * stc.F90:529-679): element e's dof vector sits at x + off_v[e] in element
 * order; its interface part x[0:ni_e] is read and its bubble part
 * x[ni_e:N_e] = BSchur - ASchur x[0:ni_e] is written (device pointers;
 * bschur at off_v[e] as hp3d_gpu_stc_batch writes it).
 */
int hp3d_gpu_stc_bwd_batch(int space, int nel, const int *norder, const long long *off_s,
                           const double *aschur, const long long *off_v, const double *bschur,
                           double *x, void *stream);

/*
 * Dirichlet lift of condensed element loads (celem_systemI.F90:794-808:
 * ZBMOD(:) -= ZAMOD(:,k2) ZDOFD(k2) for the Dirichlet dofs k2), device
 * pointers: bi[off_v[e] + k] -= sum_k2 Aii_e(k, k2) ud[off_v[e] + k2] over the
 * interface dofs (Aii_e at off_c[e], leading dimension ni_e), where ud holds
 * the element's Dirichlet dof values and 0 at its free dofs.
 */
int hp3d_gpu_dirichlet_lift(int space, int nel, const int *norder, const long long *off_c,
                            const double *aii, const long long *off_v, const double *ud, double *bi,
                            void *stream);

/*
 * Shape functions of one hexahedron at npts master-element points xi
 * (3 x npts, in [0,1]^3), HOST memory and host arithmetic, from the same dof
 * descriptors and 1D factor formulas as the kernels: shape3DH
 * (ContExactSequence.F90:397-437) gives val[p*N + k] and the reference
 * gradient grad[3*(p*N + k) + c]; shape3DE (:468-508) gives the vector
 * val[3*(p*N + k) + c] and the reference curl in grad.  For set-up work on
 * the host (the Dirichlet data of update_Ddof).
 */
int hp3d_gpu_shape(int space, const int norder[19], const int nori_e[12], const int nori_f[6],
                   int npts, const double *xi, double *val, double *grad);

/* ------------------------------------------------ solution-level entries */

#define HP3D_FIELD_ERROR 0   /* out[2e] = error^2, out[2e+1] = exact norm^2       */
#define HP3D_FIELD_ADJOINT 1 /* out[e] = -coef_e int grad v . grad u  (H(curl): v.u) */
#define HP3D_FIELD_L2 2      /* out[e] = int u . u;  r[off_v[e]+k] = int phi_k . u  */

/*
 * Integrals of finite-element fields given by element dof vectors (device
 * pointers; u, v at off_v[e], element order, full N_e including bubbles):
 *   HP3D_FIELD_ERROR    element_error (element/util/compute_error.F90:226-400)
 *                       against the manufactured solution `exact` (a load
 *                       selector: SINSINSIN, CONV, POLY2 for H1, CURL_SIN for
 *                       H(curl)): H1 norm (H(curl) norm) squared of u - u_h and
 *                       of u, quadrature of set_3Dint with INTEGRATION = 2
 *   HP3D_FIELD_ADJOINT  -lambda_e^T K_e u_e for K_e linear in coef_e, with
 *                       lambda = v: dJ/dtheta_e when coef_e = exp(theta_e)
 *                       (no counterpart in the reference)
 *   HP3D_FIELD_L2       int |u_h|^2 and r_e = M_e u_e (element mass matrix)
 * status[e] (may be NULL): 0, 1 (negative Jacobian) or HP3D_E_ORDER.
 */
int hp3d_gpu_field_batch(int space, int mode, int nel, const int *norder, const int *nori_e,
                         const int *nori_f, const double *xnod, int ngeo, const double *coef,
                         const long long *off_v, const double *u, const double *v, int exact,
                         double *out, double *r, int *status, void *stream);

/*
 * Deterministic segmented sum out[t] = sum_{ptr[t] <= j < ptr[t+1]} in[src[j]]
 * (device pointers).  With src the positions of the element contributions to
 * each global nonzero (or right-hand-side entry) in the packed condensed
 * outputs, this assembles the element systems into CSR values: the triplet
 * accumulation of mumps_sc.F90:266-323 (duplicates summed), done on device.
 */
int hp3d_gpu_gather_sum(long long nout, const long long *ptr, const long long *src,
                        const double *in, double *out, void *stream);

/*
 * y = A x for the assembled global matrix in CSR form (indptr int64,
 * indices int32, FP64 values; device pointers): the operator of the
 * iterative solve that stands in for the reference's MUMPS factorisation
 * (the dmumps analysis, factorisation and solve of mumps_sc.F90:368-426) on meshes too large for a host
 * direct solve.  No counterpart in the reference.
 */
int hp3d_gpu_csr_spmv(long long nrows, const long long *indptr, const int *indices,
                      const double *vals, const double *x, double *y, void *stream);

#ifdef __cplusplus
}
#endif
#endif
