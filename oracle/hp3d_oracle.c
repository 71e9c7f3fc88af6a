/*
 * hp3d_oracle.c -- CPU restatement of hp3D's hexahedral element path.
 *
 * TEST INFRASTRUCTURE ONLY: the correctness checker for the HIP product in
 * hp3d_amd/ and the timed CPU baseline ("port") of bench.py.  Nothing in the
 * product links or calls this file.
 *
 * Each function cites the reference routine it restates (paths relative to
 * the hp3d repository root).  How each one is pinned (tests/test_oracle.py, tests/test_mesh.py):
 *
 *   orc_poly_legendre / orc_poly_ilegendre
 *       bitwise against the reference Polynomials.F90 compiled by
 *       oracle/Makefile (oracle/_ref), and against numpy.polynomial.legendre
 *       closed forms (P_i((2x-t)/t) t^i, L_i = (P_i - P_{i-2})/(2(2i-1))).
 *   orc_gauss1d
 *       bitwise against the reference tables (oracle/_ref) and to 1e-14
 *       against numpy leggauss; exactness to degree 2n-1.
 *   orc_shape3dh_hexa / orc_shape3de_hexa
 *       bitwise (<=2 ulp) against the reference Hexahedron.F90 (oracle/_ref)
 *       for every orientation code and anisotropic orders; plus partition of
 *       unity, polynomial reproduction of Q^{px,py,pz}, exact sequence
 *       (grad H1 in span H(curl)) and trace conformity across shared faces.
 *   orc_geom
 *       bitwise against the reference geom.F90 (oracle/_ref); J * J^-1 = I.
 *   orc_set_3d_int / orc_find_order_loc
 *       point counts and ordering against set_3D_int.F90:203-258 semantics,
 *       exactness of the tensor rule, sum of weights.
 *   orc_elem_h1
 *       8x8 closed-form trilinear Laplacian on the unit cube (p=1),
 *       K 1 = 0, symmetry / PSD, h-scaling K(hx) = h K(x), and the
 *       rank-one (elem_poisson) form equals the SYRK form.
 *   orc_elem_hcurl
 *       symmetry / PSD, curl-curl kernel contains gradients of H1
 *       (via the exact sequence), mass block against an independent
 *       quadrature in numpy.
 *   orc_stc_fwd_gen / orc_stc_fwd_herm / orc_stc_bwd
 *       against scipy.linalg LAPACK (getrf/getrs, potrf) and the Schur
 *       identity (full solve == condensed solve + back substitution).
 *   orc_element_loop
 *       composition of the above; checked element by element.
 *   orc_element_loop_blas (the CPU baseline's reference-shaped loop)
 *       the same elements with DSYRK and DGETRF/DGETRS/DGEMM from a host
 *       BLAS/LAPACK: equal to orc_element_loop to 1e-12 (tests/test_oracle.py).
 *
 * Build: oracle/Makefile (gcc -O2 -ffp-contract=off, no FMA on x86-64
 * baseline, so rounding follows the source order like the reference build).
 */
#include "hp3d_oracle.h"

#include <math.h>
#include <stdlib.h>
#include <string.h>
#ifdef _OPENMP
#include <omp.h>
#endif

#define MODORDER 10

/* ---------------------------------------------------------------- utility */

/* decod: utility/decod.F90:21-39 (digits of Nick in base Mod, most significant first) */
void orc_decod(int nick, int mod, int n, int *narr)
{
    int nick1 = nick;
    for (int i = 1; i <= n - 1; i++) {
        int nick2 = nick1 / mod;
        narr[n - i] = nick1 - nick2 * mod;
        nick1 = nick2;
    }
    narr[0] = nick1;
}

/* cross: utility/cross_product.F90:130-144 (N = 3) */
static void cross3(const double *a, const double *b, double *c)
{
    c[0] = a[1] * b[2] - a[2] * b[1];
    c[1] = -(a[0] * b[2]) + a[2] * b[0];
    c[2] = a[0] * b[1] - a[1] * b[0];
}

/* ------------------------------------------------------------- Legendre */

/* PolyLegendre: element/shape_1/Polynomials.F90:34-78 */
void orc_poly_legendre(double x, double t, int nord, double *p)
{
    double y = 0.0;
    p[0] = 1.0;
    if (nord >= 1) {
        y = 2.0 * x - t;
        p[1] = y;
    }
    if (nord >= 2) {
        double tt = t * t;
        for (int i = 2; i <= nord; i++) {
            p[i] = (double)(2 * i - 1) * y * p[i - 1] - (double)(i - 1) * tt * p[i - 2];
            p[i] = p[i] / (double)i;
        }
    }
}

/* PolyILegendre: Polynomials.F90:109-167.  l[2..nord], p[1..nord-1], r[1..nord-1] */
void orc_poly_ilegendre(double x, double t, int nord, int idec, double *l, double *p, double *r)
{
    double ptemp[ORC_MAXP + 2];
    orc_poly_legendre(x, t, nord, ptemp);
    for (int i = 1; i <= nord - 1; i++) p[i] = ptemp[i];
    double tt = t * t;
    for (int i = 2; i <= nord; i++) {
        int ifact = 4 * i - 2;
        l[i] = (ptemp[i] - tt * ptemp[i - 2]) / (double)ifact;
        if (!idec) r[i - 1] = -(ptemp[i - 1] + t * ptemp[i - 2]) / 2.0;
    }
}

/* HomILegendre, Idec = .true. branch (Polynomials.F90:497-513): hexahedra only use it */
static void hom_ilegendre(const double *s, const double ds[2][3], int nord, double *homl,
                          double dhoml[][3])
{
    double homp[ORC_MAXP + 2], homr[ORC_MAXP + 2];
    orc_poly_ilegendre(s[1], 1.0, nord, 1, homl, homp, homr);
    for (int i = 2; i <= nord; i++)
        for (int c = 0; c < 3; c++) dhoml[i][c] = homp[i - 1] * ds[1][c];
}

/* AncPhiE: element/shape_1/Ancillary.F90:37-59 */
static void anc_phie(const double *s, const double ds[2][3], int nord, double *phie,
                     double dphie[][3])
{
    hom_ilegendre(s, ds, nord, phie, dphie);
}

/* AncEE, Idec = .true. branch: Ancillary.F90:88-131 (EE_i = P_i(s1; s0+s1) grad s1, curl = 0) */
static void anc_ee(const double *s, const double ds[2][3], int nord, double ee[][3],
                   double curlee[][3])
{
    double homp[ORC_MAXP + 2];
    orc_poly_legendre(s[1], s[0] + s[1], nord - 1, homp);
    for (int i = 0; i <= nord - 1; i++)
        for (int c = 0; c < 3; c++) {
            ee[i][c] = homp[i] * ds[1][c];
            curlee[i][c] = 0.0;
        }
}

/* AncPhiQuad: Ancillary.F90:171-210 */
static void anc_phiquad(const double st[2][2], const double dst[2][2][3], const int *nord,
                        double phiq[ORC_MAXP + 1][ORC_MAXP + 1],
                        double dphiq[ORC_MAXP + 1][ORC_MAXP + 1][3])
{
    double phis[ORC_MAXP + 1], dphis[ORC_MAXP + 1][3];
    double phit[ORC_MAXP + 1], dphit[ORC_MAXP + 1][3];
    anc_phie(st[0], dst[0], nord[0], phis, dphis);
    anc_phie(st[1], dst[1], nord[1], phit, dphit);
    for (int j = 2; j <= nord[1]; j++)
        for (int i = 2; i <= nord[0]; i++) {
            phiq[i][j] = phis[i] * phit[j];
            for (int c = 0; c < 3; c++) dphiq[i][j][c] = phis[i] * dphit[j][c] + phit[j] * dphis[i][c];
        }
}

/* AncEQuad: Ancillary.F90:237-303.  equad[i][j], i = 0..nord[0]-1, j = 2..nord[1] */
static void anc_equad(const double st[2][2], const double dst[2][2][3], const int *nord,
                      double equad[ORC_MAXP][ORC_MAXP + 1][3],
                      double cequad[ORC_MAXP][ORC_MAXP + 1][3])
{
    double ees[ORC_MAXP][3], cees[ORC_MAXP][3];
    double phit[ORC_MAXP + 1], dphit[ORC_MAXP + 1][3];
    anc_ee(st[0], dst[0], nord[0], ees, cees);
    anc_phie(st[1], dst[1], nord[1], phit, dphit);
    for (int j = 2; j <= nord[1]; j++)
        for (int i = 0; i <= nord[0] - 1; i++) {
            double dxe[3];
            for (int c = 0; c < 3; c++) equad[i][j][c] = ees[i][c] * phit[j];
            cross3(dphit[j], ees[i], dxe);
            for (int c = 0; c < 3; c++) cequad[i][j][c] = cees[i][c] * phit[j] + dxe[c];
        }
}

/* ---------------------------------------------------- hexahedron tables */

/* AffineHexahedron: AffineCoordinates.F90:56-72.  mu[d][a], dmu[d][a][:] */
static void affine_hexahedron(const double *xi, double mu[3][2], double dmu[3][2][3])
{
    for (int d = 0; d < 3; d++) {
        mu[d][0] = 1.0 - xi[d];
        mu[d][1] = xi[d];
        for (int c = 0; c < 3; c++) {
            dmu[d][0][c] = 0.0;
            dmu[d][1][c] = 0.0;
        }
        dmu[d][0][d] = -1.0;
        dmu[d][1][d] = 1.0;
    }
}

/* BlendHexaV vertex -> affine index per direction: BlendProject.F90:197-253 */
static const int VTX_MU[8][3] = {{0, 0, 0}, {1, 0, 0}, {1, 1, 0}, {0, 1, 0},
                                 {0, 0, 1}, {1, 0, 1}, {1, 1, 1}, {0, 1, 1}};

/* BlendProjectHexaE: BlendProject.F90:255-363.  {mu index, dir} of the two blends, and the
 * direction of the projected pair */
static const int EDGE_B1[12][2] = {{0, 1}, {1, 0}, {1, 1}, {0, 0}, {0, 1}, {1, 0},
                                   {1, 1}, {0, 0}, {0, 0}, {1, 0}, {1, 0}, {0, 0}};
static const int EDGE_B2[12][2] = {{0, 2}, {0, 2}, {0, 2}, {0, 2}, {1, 2}, {1, 2},
                                   {1, 2}, {1, 2}, {0, 1}, {0, 1}, {1, 1}, {1, 1}};
static const int EDGE_DIR[12] = {0, 1, 0, 1, 0, 1, 0, 1, 2, 2, 2, 2};

/* BlendProjectHexaF: BlendProject.F90:365-439.  {mu index, normal dir}, local axes */
static const int FACE_B[6][2] = {{0, 2}, {1, 2}, {0, 1}, {1, 0}, {1, 1}, {0, 0}};
static const int FACE_AX[6][2] = {{0, 1}, {0, 1}, {0, 2}, {1, 2}, {0, 2}, {1, 2}};

/* OrientE: element/shape_1/Orient.F90:9-37 */
static void orient_e(const double *s, const double ds[2][3], int nori, double *gs,
                     double gds[2][3])
{
    int o0 = nori == 0 ? 0 : 1, o1 = nori == 0 ? 1 : 0;
    gs[0] = s[o0];
    gs[1] = s[o1];
    for (int c = 0; c < 3; c++) {
        gds[0][c] = ds[o0][c];
        gds[1][c] = ds[o1][c];
    }
}

/* OrientQuad: Orient.F90:39-116.  OrPa (1-based pair), OrGS, OrGT per code */
static const int ORPA[8][2] = {{1, 2}, {2, 1}, {1, 2}, {2, 1}, {2, 1}, {1, 2}, {2, 1}, {1, 2}};
static const int ORGS[8][2] = {{0, 1}, {0, 1}, {1, 0}, {1, 0}, {0, 1}, {1, 0}, {1, 0}, {0, 1}};
static const int ORGT[8][2] = {{0, 1}, {1, 0}, {1, 0}, {0, 1}, {0, 1}, {0, 1}, {1, 0}, {1, 0}};

/* st[pair][a], dst[pair][a][:] -> gst, gdst */
static void orient_quad(const double st[2][2], const double dst[2][2][3], int nori,
                        double gst[2][2], double gdst[2][2][3])
{
    int p1 = ORPA[nori][0] - 1, p2 = ORPA[nori][1] - 1;
    for (int a = 0; a < 2; a++) {
        gst[0][a] = st[p1][ORGS[nori][a]];
        gst[1][a] = st[p2][ORGT[nori][a]];
        for (int c = 0; c < 3; c++) {
            gdst[0][a][c] = dst[p1][ORGS[nori][a]][c];
            gdst[1][a][c] = dst[p2][ORGT[nori][a]][c];
        }
    }
}

/* NFAXES(3,:): element_data.F90:246-249 (1 = face axes swapped) */
static const int NFAXES3[8] = {0, 1, 0, 1, 1, 0, 1, 0};

/* checkorder: element/shape_1/CheckOrder.F90:31-152 (hexahedron).  Returns max order, or
 * -1 if some order is outside [1, MAXP] (the reference stops there). */
int orc_checkorder(const int *norder)
{
    int nps[27], m = 0;
    for (int e = 0; e < 12; e++) nps[m++] = norder[e];
    for (int f = 0; f < 6; f++) {
        int hv[2];
        orc_decod(norder[12 + f], MODORDER, 2, hv);
        nps[m++] = hv[0];
        nps[m++] = hv[1];
    }
    int b[3];
    orc_decod(norder[18], MODORDER, 3, b);
    nps[m++] = b[0];
    nps[m++] = b[1];
    nps[m++] = b[2];
    int mn = nps[0], mx = nps[0];
    for (int i = 1; i < m; i++) {
        if (nps[i] < mn) mn = nps[i];
        if (nps[i] > mx) mx = nps[i];
    }
    if (mn < 1 || mx > ORC_MAXP) return -1;
    return mx;
}

/* ------------------------------------------------------------ dof counts */

/* celndof + ndof_nod (element/util/celndof.F90:24-149, element_data.F90:808-883), hexa H1 */
int orc_ndof_h1(const int *norder)
{
    int n = 8;
    for (int e = 0; e < 12; e++) n += norder[e] - 1;
    for (int f = 0; f < 6; f++) {
        int hv[2];
        orc_decod(norder[12 + f], MODORDER, 2, hv);
        n += (hv[0] - 1) * (hv[1] - 1);
    }
    return n + orc_nbubble_h1(norder[18]);
}

int orc_nbubble_h1(int nord_mdl)
{
    int b[3];
    orc_decod(nord_mdl, MODORDER, 3, b);
    return (b[0] - 1) * (b[1] - 1) * (b[2] - 1);
}

int orc_ndof_hcurl(const int *norder)
{
    int n = 0;
    for (int e = 0; e < 12; e++) n += norder[e];
    for (int f = 0; f < 6; f++) {
        int hv[2];
        orc_decod(norder[12 + f], MODORDER, 2, hv);
        n += hv[0] * (hv[1] - 1) + (hv[0] - 1) * hv[1];
    }
    return n + orc_nbubble_hcurl(norder[18]);
}

int orc_nbubble_hcurl(int nord_mdl)
{
    int b[3];
    orc_decod(nord_mdl, MODORDER, 3, b);
    return b[0] * (b[1] - 1) * (b[2] - 1) + (b[0] - 1) * b[1] * (b[2] - 1) +
           (b[0] - 1) * (b[1] - 1) * b[2];
}

/* ------------------------------------------------------------- quadrature */

/* 1D Gauss-Legendre nodes/weights on [-1,1] as tabulated in
 * element/quadrature/gauss_quadrature.F90:518-660 (15 digits for n <= 9). */
static const double GX[10][10] = {
    {0.0},
    {-.577350269189626, .577350269189626},
    {-.774596669241483, .000000000000000, .774596669241483},
    {-.861136311594053, -.339981043584856, .339981043584856, .861136311594053},
    {-.906179845938664, -.538469310105683, .000000000000000, .538469310105683,
     .906179845938664},
    {-.932469514203152, -.661209386466265, -.238619186083197, .238619186083197,
     .661209386466265, .932469514203152},
    {-.949107912342759, -.741531185599394, -.405845151377397, .000000000000000,
     .405845151377397, .741531185599394, .949107912342759},
    {-.960289856497536, -.796666477413627, -.525532409916329, -.183434642495650,
     .183434642495650, .525532409916329, .796666477413627, .960289856497536},
    {-.968160239507626, -.836031107326636, -.613371432700590, -.324253423403809,
     .000000000000000, .324253423403809, .613371432700590, .836031107326636,
     .968160239507626},
    {-.9739065285171717200779640, -.8650633666889845107320967, -.6794095682990244062343274,
     -.4333953941292471907992659, -.1488743389816312108848260, .1488743389816312108848260,
     .4333953941292471907992659, .6794095682990244062343274, .8650633666889845107320967,
     .9739065285171717200779640}};
static const double GW[10][10] = {
    {2.000000000000000},
    {1.000000000000000, 1.000000000000000},
    {.555555555555556, .888888888888889, .555555555555556},
    {.347854845137454, .652145154862546, .652145154862546, .347854845137454},
    {.236926885056189, .478628670499366, .568888888888889, .478628670499366,
     .236926885056189},
    {.171324492379170, .360761573048139, .467913934572691, .467913934572691,
     .360761573048139, .171324492379170},
    {.129484966168870, .279705391489277, .381830050505119, .417959183673469,
     .381830050505119, .279705391489277, .129484966168870},
    {.101228536290376, .222381034453374, .313706645877887, .362683783378362,
     .362683783378362, .313706645877887, .222381034453374, .101228536290376},
    {.081274388361574, .180648160694857, .260610696402935, .312347077040003,
     .330239355001260, .312347077040003, .260610696402935, .180648160694857,
     .081274388361574},
    {.0666713443086881375935688, .1494513491505805931457763, .2190863625159820439955349,
     .2692667193099963550912269, .2955242247147528701738930, .2955242247147528701738930,
     .2692667193099963550912269, .2190863625159820439955349, .1494513491505805931457763,
     .0666713443086881375935688}};

/* init_gauss_quadrature rescaling to [0,1]: gauss_quadrature.F90:746-752 */
void orc_gauss1d(int n, double *x, double *w)
{
    for (int j = 0; j < n; j++) {
        x[j] = 0.5 * (1.0 + GX[n - 1][j]);
        w[j] = 0.5 * GW[n - 1][j];
    }
}

/* find_order_loc: datstrs/find_order.F90:68-100 (brick) */
void orc_find_order_loc(const int *norder, const int *nori_f, int *norder_loc)
{
    for (int i = 0; i < 19; i++) norder_loc[i] = norder[i];
    for (int j = 0; j < 6; j++) {
        if (NFAXES3[nori_f[j]] == 1) {
            int h = norder[12 + j] / 10, v = norder[12 + j] % 10;
            norder_loc[12 + j] = v * 10 + h;
        }
    }
}

/* per-direction point counts of set_3Dint_aux, brick branch (set_3D_int.F90:203-244);
 * INTEGRATION = 0 (files/control/control_0) for the element routines, 2 in
 * element_error (compute_error.F90:308-310) */
static void brick_nint_int(const int *nl, int integration, int *nint3);
static void brick_nint(const int *nl, int *nint3) { brick_nint_int(nl, 0, nint3); }
static void brick_nint_int(const int *nl, int integration, int *nint3)
{
    int nx = 0, ny = 0, nz = 0, hv[2], b[3];
    for (int i = 1; i <= 19; i++) {
        int v = nl[i - 1];
        switch (i) {
        case 1: case 3: case 5: case 7: nx = v > nx ? v : nx; break;
        case 2: case 4: case 6: case 8: ny = v > ny ? v : ny; break;
        case 9: case 10: case 11: case 12: nz = v > nz ? v : nz; break;
        case 13: case 14:
            orc_decod(v, MODORDER, 2, hv);
            nx = hv[0] > nx ? hv[0] : nx;
            ny = hv[1] > ny ? hv[1] : ny;
            break;
        case 15: case 17:
            orc_decod(v, MODORDER, 2, hv);
            nx = hv[0] > nx ? hv[0] : nx;
            nz = hv[1] > nz ? hv[1] : nz;
            break;
        case 16: case 18:
            orc_decod(v, MODORDER, 2, hv);
            ny = hv[0] > ny ? hv[0] : ny;
            nz = hv[1] > nz ? hv[1] : nz;
            break;
        case 19:
            orc_decod(v, MODORDER, 3, b);
            nx = b[0] > nx ? b[0] : nx;
            ny = b[1] > ny ? b[1] : ny;
            nz = b[2] > nz ? b[2] : nz;
            break;
        }
    }
    nx += integration;
    ny += integration;
    nz += integration;
    nx = nx < ORC_MAXP ? nx : ORC_MAXP;
    ny = ny < ORC_MAXP ? ny : ORC_MAXP;
    nz = nz < ORC_MAXP ? nz : ORC_MAXP;
    nint3[0] = nx + 1;
    nint3[1] = ny + 1;
    nint3[2] = nz + 1;
}

/* set_3D_int + set_3Dint_aux, brick branch: set_3D_int.F90:47-63,203-258 */
void orc_set_3d_int(const int *norder, const int *nori_f, int *nint, double *xiloc, double *waloc)
{
    int nl[19], n3[3];
    double gx[3][10], gw[3][10];
    orc_find_order_loc(norder, nori_f, nl);
    brick_nint(nl, n3);
    for (int d = 0; d < 3; d++) orc_gauss1d(n3[d], gx[d], gw[d]);
    int l = 0;
    for (int l3 = 0; l3 < n3[2]; l3++)
        for (int l2 = 0; l2 < n3[1]; l2++)
            for (int l1 = 0; l1 < n3[0]; l1++) {
                xiloc[3 * l + 0] = gx[0][l1];
                xiloc[3 * l + 1] = gx[1][l2];
                xiloc[3 * l + 2] = gx[2][l3];
                waloc[l] = gw[0][l1] * gw[1][l2] * gw[2][l3];
                l++;
            }
    *nint = l;
}

/* set_3Dint (set_3D_int.F90:8-21): the same brick rule from the face orders
 * as given (no find_order_loc swap), with INTEGRATION = `integration` */
void orc_set_3dint(const int *norder, int integration, int *nint, double *xiloc, double *waloc)
{
    int n3[3];
    double gx[3][10], gw[3][10];
    brick_nint_int(norder, integration, n3);
    for (int d = 0; d < 3; d++) orc_gauss1d(n3[d], gx[d], gw[d]);
    int l = 0;
    for (int l3 = 0; l3 < n3[2]; l3++)
        for (int l2 = 0; l2 < n3[1]; l2++)
            for (int l1 = 0; l1 < n3[0]; l1++) {
                xiloc[3 * l + 0] = gx[0][l1];
                xiloc[3 * l + 1] = gx[1][l2];
                xiloc[3 * l + 2] = gx[2][l3];
                waloc[l] = gw[0][l1] * gw[1][l2] * gw[2][l3];
                l++;
            }
    *nint = l;
}

/* ------------------------------------------------------- shape functions */

/* shape3DHHexa: element/shape_1/Hexahedron.F90:33-210.  grad is 3 x N column-major. */
int orc_shape3dh_hexa(const double *xi, const int *norder, const int *nori_e, const int *nori_f,
                      int *nrdof, double *shap, double *grad)
{
    if (orc_checkorder(norder) < 0) return -1;
    double mu[3][2], dmu[3][2][3];
    affine_hexahedron(xi, mu, dmu);
    int m = 0;

    /* vertices: Hexahedron.F90:73-82 */
    for (int v = 0; v < 8; v++) {
        double mb[3];
        const double *dmb[3];
        for (int d = 0; d < 3; d++) {
            mb[d] = mu[d][VTX_MU[v][d]];
            dmb[d] = dmu[d][VTX_MU[v][d]];
        }
        shap[m] = mb[0] * mb[1] * mb[2];
        for (int c = 0; c < 3; c++)
            grad[3 * m + c] =
                mb[0] * mb[1] * dmb[2][c] + mb[0] * dmb[1][c] * mb[2] + dmb[0][c] * mb[1] * mb[2];
        m++;
    }

    /* edges: Hexahedron.F90:86-104 */
    for (int e = 0; e < 12; e++) {
        int nord = norder[e];
        if (nord - 1 <= 0) continue;
        double mb1 = mu[EDGE_B1[e][1]][EDGE_B1[e][0]], mb2 = mu[EDGE_B2[e][1]][EDGE_B2[e][0]];
        const double *dmb1 = dmu[EDGE_B1[e][1]][EDGE_B1[e][0]];
        const double *dmb2 = dmu[EDGE_B2[e][1]][EDGE_B2[e][0]];
        int p = EDGE_DIR[e];
        double gs[2], gds[2][3], phie[ORC_MAXP + 1], dphie[ORC_MAXP + 1][3];
        orient_e(mu[p], dmu[p], nori_e[e], gs, gds);
        anc_phie(gs, gds, nord, phie, dphie);
        for (int i = 2; i <= nord; i++) {
            shap[m] = mb1 * mb2 * phie[i];
            for (int c = 0; c < 3; c++)
                grad[3 * m + c] = mb1 * mb2 * dphie[i][c] + mb1 * dmb2[c] * phie[i] +
                                  dmb1[c] * mb2 * phie[i];
            m++;
        }
    }

    /* faces: Hexahedron.F90:108-131 */
    for (int f = 0; f < 6; f++) {
        int nf[2];
        orc_decod(norder[12 + f], MODORDER, 2, nf);
        if ((nf[0] - 1) * (nf[1] - 1) <= 0) continue;
        double mbf = mu[FACE_B[f][1]][FACE_B[f][0]];
        const double *dmbf = dmu[FACE_B[f][1]][FACE_B[f][0]];
        double st[2][2], dst[2][2][3], gst[2][2], gdst[2][2][3];
        for (int q = 0; q < 2; q++) {
            int ax = FACE_AX[f][q];
            for (int a = 0; a < 2; a++) {
                st[q][a] = mu[ax][a];
                for (int c = 0; c < 3; c++) dst[q][a][c] = dmu[ax][a][c];
            }
        }
        orient_quad(st, dst, nori_f[f], gst, gdst);
        double phiq[ORC_MAXP + 1][ORC_MAXP + 1], dphiq[ORC_MAXP + 1][ORC_MAXP + 1][3];
        anc_phiquad(gst, gdst, nf, phiq, dphiq);
        for (int j = 2; j <= nf[1]; j++)
            for (int i = 2; i <= nf[0]; i++) {
                shap[m] = mbf * phiq[i][j];
                for (int c = 0; c < 3; c++)
                    grad[3 * m + c] = mbf * dphiq[i][j][c] + dmbf[c] * phiq[i][j];
                m++;
            }
    }

    /* bubbles: Hexahedron.F90:135-156 */
    int nb[3];
    orc_decod(norder[18], MODORDER, 3, nb);
    if ((nb[0] - 1) * (nb[1] - 1) * (nb[2] - 1) > 0) {
        double st[2][2], dst[2][2][3];
        for (int q = 0; q < 2; q++)
            for (int a = 0; a < 2; a++) {
                st[q][a] = mu[q][a];
                for (int c = 0; c < 3; c++) dst[q][a][c] = dmu[q][a][c];
            }
        double phiq[ORC_MAXP + 1][ORC_MAXP + 1], dphiq[ORC_MAXP + 1][ORC_MAXP + 1][3];
        double phie[ORC_MAXP + 1], dphie[ORC_MAXP + 1][3];
        anc_phiquad(st, dst, nb, phiq, dphiq);
        anc_phie(mu[2], dmu[2], nb[2], phie, dphie);
        for (int k = 2; k <= nb[2]; k++)
            for (int j = 2; j <= nb[1]; j++)
                for (int i = 2; i <= nb[0]; i++) {
                    shap[m] = phiq[i][j] * phie[k];
                    for (int c = 0; c < 3; c++)
                        grad[3 * m + c] = phiq[i][j] * dphie[k][c] + dphiq[i][j][c] * phie[k];
                    m++;
                }
    }
    *nrdof = m;
    return 0;
}

/* shape3DEHexa: element/shape_1/Hexahedron.F90:240-439.  shap, curl are 3 x N col-major */
int orc_shape3de_hexa(const double *xi, const int *norder, const int *nori_e, const int *nori_f,
                      int *nrdof, double *shap, double *curl)
{
    if (orc_checkorder(norder) < 0) return -1;
    double mu[3][2], dmu[3][2][3];
    affine_hexahedron(xi, mu, dmu);
    int m = 0;

    /* edges: Hexahedron.F90:283-302 */
    for (int e = 0; e < 12; e++) {
        int nord = norder[e];
        if (nord <= 0) continue;
        double mb1 = mu[EDGE_B1[e][1]][EDGE_B1[e][0]], mb2 = mu[EDGE_B2[e][1]][EDGE_B2[e][0]];
        const double *dmb1 = dmu[EDGE_B1[e][1]][EDGE_B1[e][0]];
        const double *dmb2 = dmu[EDGE_B2[e][1]][EDGE_B2[e][0]];
        int p = EDGE_DIR[e];
        double gs[2], gds[2][3], ee[ORC_MAXP][3], cee[ORC_MAXP][3];
        orient_e(mu[p], dmu[p], nori_e[e], gs, gds);
        anc_ee(gs, gds, nord, ee, cee);
        for (int i = 0; i <= nord - 1; i++) {
            double dt[3], ct[3];
            for (int c = 0; c < 3; c++) dt[c] = mb1 * dmb2[c] + dmb1[c] * mb2;
            cross3(dt, ee[i], ct);
            for (int c = 0; c < 3; c++) {
                shap[3 * m + c] = mb1 * mb2 * ee[i][c];
                curl[3 * m + c] = ct[c];
            }
            m++;
        }
    }

    /* faces: Hexahedron.F90:306-346 */
    for (int f = 0; f < 6; f++) {
        int nf[2];
        orc_decod(norder[12 + f], MODORDER, 2, nf);
        double mbf = mu[FACE_B[f][1]][FACE_B[f][0]];
        const double *dmbf = dmu[FACE_B[f][1]][FACE_B[f][0]];
        double st[2][2], dst[2][2][3], gst[2][2], gdst[2][2][3];
        for (int q = 0; q < 2; q++) {
            int ax = FACE_AX[f][q];
            for (int a = 0; a < 2; a++) {
                st[q][a] = mu[ax][a];
                for (int c = 0; c < 3; c++) dst[q][a][c] = dmu[ax][a][c];
            }
        }
        orient_quad(st, dst, nori_f[f], gst, gdst);
        for (int fam = 0; fam <= 1; fam++) {
            int a = fam == 0 ? 0 : 1, b = 1 - a;
            if (nf[a] * (nf[b] - 1) <= 0) continue;
            double sab[2][2], dsab[2][2][3];
            for (int q = 0; q < 2; q++) {
                int src = q == 0 ? a : b;
                for (int s = 0; s < 2; s++) {
                    sab[q][s] = gst[src][s];
                    for (int c = 0; c < 3; c++) dsab[q][s][c] = gdst[src][s][c];
                }
            }
            int nab[2] = {nf[a], nf[b]};
            double eq[ORC_MAXP][ORC_MAXP + 1][3], ceq[ORC_MAXP][ORC_MAXP + 1][3];
            anc_equad(sab, dsab, nab, eq, ceq);
            /* outer loop over the second global face axis, inner over the first */
            int ig0 = fam == 0 ? 0 : 2, ig1 = fam == 0 ? nf[0] - 1 : nf[0];
            int jg0 = fam == 0 ? 2 : 0, jg1 = fam == 0 ? nf[1] : nf[1] - 1;
            for (int jg = jg0; jg <= jg1; jg++)
                for (int ig = ig0; ig <= ig1; ig++) {
                    int i = fam == 0 ? ig : jg, j = fam == 0 ? jg : ig;
                    double ct[3];
                    cross3(dmbf, eq[i][j], ct);
                    for (int c = 0; c < 3; c++) {
                        shap[3 * m + c] = mbf * eq[i][j][c];
                        curl[3 * m + c] = mbf * ceq[i][j][c] + ct[c];
                    }
                    m++;
                }
        }
    }

    /* bubbles: Hexahedron.F90:350-388 */
    int nb[3];
    orc_decod(norder[18], MODORDER, 3, nb);
    for (int fam = 0; fam <= 2; fam++) {
        int a = fam, b = (fam + 1) % 3, c3 = (fam + 2) % 3;
        if (nb[a] * (nb[b] - 1) * (nb[c3] - 1) <= 0) continue;
        double st[2][2], dst[2][2][3];
        for (int q = 0; q < 2; q++) {
            int ax = q == 0 ? a : b;
            for (int s = 0; s < 2; s++) {
                st[q][s] = mu[ax][s];
                for (int c = 0; c < 3; c++) dst[q][s][c] = dmu[ax][s][c];
            }
        }
        int nab[2] = {nb[a], nb[b]};
        double eq[ORC_MAXP][ORC_MAXP + 1][3], ceq[ORC_MAXP][ORC_MAXP + 1][3];
        double phie[ORC_MAXP + 1], dphie[ORC_MAXP + 1][3];
        anc_equad(st, dst, nab, eq, ceq);
        anc_phie(mu[c3], dmu[c3], nb[c3], phie, dphie);
        /* global loop order z (outer), y, x (inner); the family axis runs from 0 */
        int lo[3] = {2, 2, 2}, hi[3] = {nb[0], nb[1], nb[2]};
        lo[a] = 0;
        hi[a] = nb[a] - 1;
        for (int kg = lo[2]; kg <= hi[2]; kg++)
            for (int jg = lo[1]; jg <= hi[1]; jg++)
                for (int ig = lo[0]; ig <= hi[0]; ig++) {
                    int g[3] = {ig, jg, kg};
                    int i = g[a], j = g[b], k = g[c3];
                    double ct[3];
                    cross3(dphie[k], eq[i][j], ct);
                    for (int c = 0; c < 3; c++) {
                        shap[3 * m + c] = eq[i][j][c] * phie[k];
                        curl[3 * m + c] = phie[k] * ceq[i][j][c] + ct[c];
                    }
                    m++;
                }
    }
    *nrdof = m;
    return 0;
}

/* --------------------------------------------------------------- geometry */

/* geom: element/util/geom.F90:30-134.  Column-major 3x3; returns Iflag. */
int orc_geom(const double *a, double *b, double *rjac)
{
#define A(i, j) a[((i)-1) + 3 * ((j)-1)]
#define B(i, j) b[((i)-1) + 3 * ((j)-1)]
    int iflag = 0;
    double r = A(1, 1) * A(2, 2) * A(3, 3) + A(2, 1) * A(3, 2) * A(1, 3) +
               A(3, 1) * A(1, 2) * A(2, 3) - A(3, 1) * A(2, 2) * A(1, 3) -
               A(1, 1) * A(3, 2) * A(2, 3) - A(2, 1) * A(1, 2) * A(3, 3);
    if (r < 0.0) iflag = 1;
    double d1, d2, d3;
    d1 = A(2, 2) * A(3, 3) - A(3, 2) * A(2, 3);
    d2 = -(A(2, 1) * A(3, 3)) + A(3, 1) * A(2, 3);
    d3 = A(2, 1) * A(3, 2) - A(3, 1) * A(2, 2);
    B(1, 1) = d1 / r;
    B(2, 1) = d2 / r;
    B(3, 1) = d3 / r;
    d1 = A(3, 2) * A(1, 3) - A(1, 2) * A(3, 3);
    d2 = A(1, 1) * A(3, 3) - A(3, 1) * A(1, 3);
    d3 = -(A(1, 1) * A(3, 2)) + A(3, 1) * A(1, 2);
    B(1, 2) = d1 / r;
    B(2, 2) = d2 / r;
    B(3, 2) = d3 / r;
    d1 = A(1, 2) * A(2, 3) - A(2, 2) * A(1, 3);
    d2 = -(A(1, 1) * A(2, 3)) + A(2, 1) * A(1, 3);
    d3 = A(1, 1) * A(2, 2) - A(2, 1) * A(1, 2);
    B(1, 3) = d1 / r;
    B(2, 3) = d2 / r;
    B(3, 3) = d3 / r;
    *rjac = r;
    return iflag;
#undef A
#undef B
}

/* geom3D, isoparametric branch (EXGEOM = 0): element/util/geom3D.F90:77-91 */
int orc_geom3d(const double *xnod, int ngeo, const double *shap, const double *grad, double *x,
               double *dxdxi, double *dxidx, double *rjac)
{
    for (int i = 0; i < 3; i++) x[i] = 0.0;
    for (int i = 0; i < 9; i++) dxdxi[i] = 0.0;
    for (int k = 0; k < ngeo; k++) {
        for (int c = 0; c < 3; c++) x[c] = x[c] + xnod[3 * k + c] * shap[k];
        for (int i = 0; i < 3; i++)
            for (int c = 0; c < 3; c++)
                dxdxi[c + 3 * i] = dxdxi[c + 3 * i] + xnod[3 * k + c] * grad[3 * k + i];
    }
    return orc_geom(dxdxi, dxidx, rjac);
}

/* ------------------------------------------------------------------ loads */

/* getf restated per load selector (POISSON/GALERKIN/common/getf.F90:20-65 computes
 * f = -Laplace(u) of the manufactured solution) */
void orc_getf_h1(int load, const double *x, double *f)
{
    const double pi = 3.14159265358979323846;
    switch (load) {
    case ORC_LOAD_POLY2: {
        double a = x[0] * (1.0 - x[0]), b = x[1] * (1.0 - x[1]), c = x[2] * (1.0 - x[2]);
        *f = 2.0 * (b * c + a * c + a * b);
        break;
    }
    case ORC_LOAD_SINSINSIN:
        *f = 3.0 * pi * pi * sin(pi * x[0]) * sin(pi * x[1]) * sin(pi * x[2]);
        break;
    case ORC_LOAD_CONV:
        *f = -2.0 * sin(pi * x[0]) * sinh(pi * x[1]);
        break;
    case ORC_LOAD_ONE:
        *f = 1.0;
        break;
    default:
        *f = 0.0;
    }
}

/* J = curl curl E* + E* for E* = (sin(pi y) sin(pi z), 0, 0) */
void orc_getf_hcurl(int load, const double *x, double *j)
{
    const double pi = 3.14159265358979323846;
    j[0] = j[1] = j[2] = 0.0;
    if (load == ORC_LOAD_CURL_SIN) j[0] = (1.0 + 2.0 * pi * pi) * sin(pi * x[1]) * sin(pi * x[2]);
}

/* --------------------------------------------------------- element routines */

/* SYRK upper, C = A A^T (A is n x k col-major, lda = n), then mirror to lower:
 * POISSON/GALERKIN/elem_opt.F90:131-137 (BLAS DSYRK, summation order unpinned) */
static void syrk_upper_add(int n, int k, const double *a, double alpha, double *c, int ldc,
                           int accumulate)
{
    for (int j = 0; j < n; j++)
        for (int i = 0; i <= j; i++) {
            double s = 0.0;
            for (int l = 0; l < k; l++) s += a[i + (size_t)l * n] * a[j + (size_t)l * n];
            c[i + (size_t)j * ldc] = (accumulate ? c[i + (size_t)j * ldc] : 0.0) + alpha * s;
        }
}

static void mirror_lower(int n, double *c, int ldc)
{
    for (int k = 0; k < n; k++)
        for (int i = k + 1; i < n; i++) c[i + (size_t)k * ldc] = c[k + (size_t)i * ldc];
}

/* elem_opt (Poisson): problems/POISSON/GALERKIN/elem_opt.F90:22-141, with a per-element
 * conductivity kappa folded into the weight: A_TEST = grad(phi) sqrt(kappa w detJ) */
static void syrk_blas(int n, int k, const double *a, double beta, double *c, int ldc);

static int elem_h1_core(const int *norder, const int *nori_e, const int *nori_f,
                        const double *xnod, int ngeo, double kappa, int load, double *zaloc,
                        int ld, double *zbloc, int blas)
{
    if (orc_checkorder(norder) < 0) return -1;
    int n = orc_ndof_h1(norder);
    if (ngeo > n) return -2;
    int nint, status = 0;
    double xiloc[3 * ORC_MAX_NINT3], waloc[ORC_MAX_NINT3];
    orc_set_3d_int(norder, nori_f, &nint, xiloc, waloc);
    int nda = 3 * nint;
    double *atest = (double *)malloc(sizeof(double) * (size_t)n * nda);
    double shap[ORC_MAXBRICKH], grad[3 * ORC_MAXBRICKH];
    for (int k = 0; k < n; k++) zbloc[k] = 0.0;
    for (int l = 0; l < nint; l++) {
        int nrdof;
        double x[3], dxdxi[9], dxidx[9], rjac, fval;
        orc_shape3dh_hexa(&xiloc[3 * l], norder, nori_e, nori_f, &nrdof, shap, grad);
        if (orc_geom3d(xnod, ngeo, shap, grad, x, dxdxi, dxidx, &rjac)) status = 1;
        double weight = rjac * waloc[l];
        orc_getf_h1(load, x, &fval);
        double sw = sqrt(kappa * weight);
        for (int k = 0; k < nrdof; k++) {
            double q = shap[k];
            const double *g = &grad[3 * k];
            zbloc[k] = zbloc[k] + q * fval * weight;
            for (int d = 0; d < 3; d++) {
                double dq = g[0] * dxidx[0 + 3 * d] + g[1] * dxidx[1 + 3 * d] + g[2] * dxidx[2 + 3 * d];
                atest[k + (size_t)(3 * l + d) * n] = dq * sw;
            }
        }
    }
    if (blas)
        syrk_blas(n, nda, atest, 0.0, zaloc, ld);
    else
        syrk_upper_add(n, nda, atest, 1.0, zaloc, ld, 0);
    mirror_lower(n, zaloc, ld);
    free(atest);
    return status;
}

int orc_elem_h1(const int *norder, const int *nori_e, const int *nori_f, const double *xnod,
                int ngeo, double kappa, int load, double *zaloc, int ld, double *zbloc)
{
    return elem_h1_core(norder, nori_e, nori_f, xnod, ngeo, kappa, load, zaloc, ld, zbloc, 0);
}

/* elem_opt (Maxwell) real restatement: problems/MAXWELL/GALERKIN/elem_opt.F90:22-158 with
 * K = A A^T + M M^T, A = curl(F) sqrt(w/mu), M = F sqrt(eps w), b = (J, F) */
static int elem_hcurl_core(const int *norder, const int *nori_e, const int *nori_f,
                           const double *xnod, int ngeo, double mu, double eps, int load,
                           double *zaloc, int ld, double *zbloc, int blas)
{
    if (orc_checkorder(norder) < 0) return -1;
    int nh = orc_ndof_h1(norder), n = orc_ndof_hcurl(norder);
    if (ngeo > nh) return -2;
    int nint, status = 0;
    double xiloc[3 * ORC_MAX_NINT3], waloc[ORC_MAX_NINT3];
    orc_set_3d_int(norder, nori_f, &nint, xiloc, waloc);
    int nda = 3 * nint;
    double *at = (double *)malloc(sizeof(double) * (size_t)n * nda);
    double *mt = (double *)malloc(sizeof(double) * (size_t)n * nda);
    double shap[ORC_MAXBRICKH], grad[3 * ORC_MAXBRICKH];
    double *she = (double *)malloc(sizeof(double) * 3 * ORC_MAXBRICKE);
    double *cue = (double *)malloc(sizeof(double) * 3 * ORC_MAXBRICKE);
    for (int k = 0; k < n; k++) zbloc[k] = 0.0;
    for (int l = 0; l < nint; l++) {
        int nrh, nre;
        double x[3], dxdxi[9], dxidx[9], rjac, zj[3];
        orc_shape3dh_hexa(&xiloc[3 * l], norder, nori_e, nori_f, &nrh, shap, grad);
        orc_shape3de_hexa(&xiloc[3 * l], norder, nori_e, nori_f, &nre, she, cue);
        if (orc_geom3d(xnod, ngeo, shap, grad, x, dxdxi, dxidx, &rjac)) status = 1;
        double weight = rjac * waloc[l];
        orc_getf_hcurl(load, x, zj);
        double sa = sqrt(weight / mu), sm = sqrt(eps * weight);
        for (int k = 0; k < nre; k++) {
            const double *e = &she[3 * k], *c = &cue[3 * k];
            double f[3], cf[3];
            for (int d = 0; d < 3; d++) {
                f[d] = e[0] * dxidx[0 + 3 * d] + e[1] * dxidx[1 + 3 * d] + e[2] * dxidx[2 + 3 * d];
                cf[d] = dxdxi[d + 0] * c[0] + dxdxi[d + 3] * c[1] + dxdxi[d + 6] * c[2];
                cf[d] = cf[d] / rjac;
            }
            double za = f[0] * zj[0] + f[1] * zj[1] + f[2] * zj[2];
            zbloc[k] = zbloc[k] + za * weight;
            for (int d = 0; d < 3; d++) {
                mt[k + (size_t)(3 * l + d) * n] = f[d] * sm;
                at[k + (size_t)(3 * l + d) * n] = cf[d] * sa;
            }
        }
    }
    if (blas) {
        syrk_blas(n, nda, at, 0.0, zaloc, ld);
        syrk_blas(n, nda, mt, 1.0, zaloc, ld);
    } else {
        syrk_upper_add(n, nda, at, 1.0, zaloc, ld, 0);
        syrk_upper_add(n, nda, mt, 1.0, zaloc, ld, 1);
    }
    mirror_lower(n, zaloc, ld);
    free(at);
    free(mt);
    free(she);
    free(cue);
    return status;
}

int orc_elem_hcurl(const int *norder, const int *nori_e, const int *nori_f, const double *xnod,
                   int ngeo, double mu, double eps, int load, double *zaloc, int ld,
                   double *zbloc)
{
    return elem_hcurl_core(norder, nori_e, nori_f, xnod, ngeo, mu, eps, load, zaloc, ld, zbloc,
                           0);
}

/* ---------------------------------------------------- static condensation */

/* pack of stc_fwd_wrapper (modules/stc.F90:227-261): interface dofs first, bubbles last */
static void stc_pack(int ni, int nb, const double *aloc, int ld, const double *bloc, double *aii,
                     double *abi, double *aib, double *abb, double *bi, double *bb)
{
    for (int j = 0; j < ni; j++)
        for (int i = 0; i < ni; i++) aii[i + (size_t)j * ni] = aloc[i + (size_t)j * ld];
    for (int j = 0; j < ni; j++)
        for (int i = 0; i < nb; i++) abi[i + (size_t)j * nb] = aloc[ni + i + (size_t)j * ld];
    for (int j = 0; j < nb; j++)
        for (int i = 0; i < ni; i++) aib[i + (size_t)j * ni] = aloc[i + (size_t)(ni + j) * ld];
    for (int j = 0; j < nb; j++)
        for (int i = 0; i < nb; i++) abb[i + (size_t)j * nb] = aloc[ni + i + (size_t)(ni + j) * ld];
    for (int i = 0; i < ni; i++) bi[i] = bloc[i];
    for (int i = 0; i < nb; i++) bb[i] = bloc[ni + i];
}

/* condensed system back into ALOC/BLOC, Schur factors out: stc.F90:270-305, and the two
 * Schur-complement DGEMMs of stc_fwd_*: stc.F90:404-416 / 497-509 */
static void stc_finish(int ni, int nb, double *aii, const double *aib, const double *x,
                       double *bi, const double *xb, double *aloc, int ld, double *bloc,
                       double *aschur, double *bschur)
{
    for (int i = 0; i < ni; i++) {
        double s = 0.0;
        for (int b = 0; b < nb; b++) s += aib[i + (size_t)b * ni] * xb[b];
        bi[i] = -s + bi[i];
    }
    for (int j = 0; j < ni; j++)
        for (int i = 0; i < ni; i++) {
            double s = 0.0;
            for (int b = 0; b < nb; b++) s += aib[i + (size_t)b * ni] * x[b + (size_t)j * nb];
            aii[i + (size_t)j * ni] = -s + aii[i + (size_t)j * ni];
        }
    for (int j = 0; j < ni; j++)
        for (int i = 0; i < ni; i++) aloc[i + (size_t)j * ld] = aii[i + (size_t)j * ni];
    for (int i = 0; i < ni; i++) bloc[i] = bi[i];
    if (aschur)
        for (size_t i = 0; i < (size_t)nb * ni; i++) aschur[i] = x[i];
    if (bschur)
        for (int i = 0; i < nb; i++) bschur[i] = xb[i];
}

/* stc_fwd_gen: stc.F90:443-509 (DGETRF / DGETRS restated as unblocked partial-pivot LU) */
int orc_stc_fwd_gen(int ni, int nb, double *aloc, int ld, double *bloc, double *aschur,
                    double *bschur)
{
    if (nb == 0) return 0;
    if (ni == 0) return -3;
    double *aii = malloc(sizeof(double) * (size_t)ni * ni);
    double *abi = malloc(sizeof(double) * (size_t)nb * ni);
    double *aib = malloc(sizeof(double) * (size_t)ni * nb);
    double *abb = malloc(sizeof(double) * (size_t)nb * nb);
    double *bi = malloc(sizeof(double) * ni), *bb = malloc(sizeof(double) * nb);
    int *piv = malloc(sizeof(int) * nb);
    int info = 0;
    stc_pack(ni, nb, aloc, ld, bloc, aii, abi, aib, abb, bi, bb);
#define LU(i, j) abb[(i) + (size_t)(j)*nb]
    for (int j = 0; j < nb; j++) {
        int p = j;
        double amax = fabs(LU(j, j));
        for (int i = j + 1; i < nb; i++)
            if (fabs(LU(i, j)) > amax) {
                amax = fabs(LU(i, j));
                p = i;
            }
        piv[j] = p;
        if (LU(p, j) == 0.0) {
            if (!info) info = j + 1;
            continue;
        }
        if (p != j)
            for (int k = 0; k < nb; k++) {
                double t = LU(j, k);
                LU(j, k) = LU(p, k);
                LU(p, k) = t;
            }
        double r = 1.0 / LU(j, j);
        for (int i = j + 1; i < nb; i++) LU(i, j) *= r;
        for (int k = j + 1; k < nb; k++)
            for (int i = j + 1; i < nb; i++) LU(i, k) -= LU(i, j) * LU(j, k);
    }
    if (info) goto done;
    /* DGETRS('N'): apply interchanges, unit-lower solve, upper solve; columns bb, abi */
    for (int col = -1; col < ni; col++) {
        double *x = col < 0 ? bb : &abi[(size_t)col * nb];
        for (int j = 0; j < nb; j++)
            if (piv[j] != j) {
                double t = x[j];
                x[j] = x[piv[j]];
                x[piv[j]] = t;
            }
        for (int j = 0; j < nb; j++)
            for (int i = j + 1; i < nb; i++) x[i] -= x[j] * LU(i, j);
        for (int j = nb - 1; j >= 0; j--) {
            x[j] = x[j] / LU(j, j);
            for (int i = 0; i < j; i++) x[i] -= x[j] * LU(i, j);
        }
    }
#undef LU
    stc_finish(ni, nb, aii, aib, abi, bi, bb, aloc, ld, bloc, aschur, bschur);
done:
    free(aii);
    free(abi);
    free(aib);
    free(abb);
    free(bi);
    free(bb);
    free(piv);
    return info ? -4 : 0;
}

/* stc_fwd_herm: stc.F90:338-416 (DTRTTF + DPFTRF + DPFTRS restated as Abb = U^T U) */
int orc_stc_fwd_herm(int ni, int nb, double *aloc, int ld, double *bloc, double *aschur,
                     double *bschur)
{
    if (nb == 0) return 0;
    if (ni == 0) return -3;
    double *aii = malloc(sizeof(double) * (size_t)ni * ni);
    double *abi = malloc(sizeof(double) * (size_t)nb * ni);
    double *aib = malloc(sizeof(double) * (size_t)ni * nb);
    double *abb = malloc(sizeof(double) * (size_t)nb * nb);
    double *bi = malloc(sizeof(double) * ni), *bb = malloc(sizeof(double) * nb);
    int info = 0;
    stc_pack(ni, nb, aloc, ld, bloc, aii, abi, aib, abb, bi, bb);
#define U(i, j) abb[(i) + (size_t)(j)*nb]
    for (int j = 0; j < nb; j++) {
        double s = U(j, j);
        for (int k = 0; k < j; k++) s -= U(k, j) * U(k, j);
        if (!(s > 0.0)) {
            info = j + 1;
            break;
        }
        double ujj = sqrt(s);
        U(j, j) = ujj;
        for (int i = j + 1; i < nb; i++) {
            double t = U(j, i);
            for (int k = 0; k < j; k++) t -= U(k, j) * U(k, i);
            U(j, i) = t / ujj;
        }
    }
    if (!info) {
        for (int col = -1; col < ni; col++) {
            double *x = col < 0 ? bb : &abi[(size_t)col * nb];
            for (int i = 0; i < nb; i++) { /* U^T y = b */
                double t = x[i];
                for (int k = 0; k < i; k++) t -= U(k, i) * x[k];
                x[i] = t / U(i, i);
            }
            for (int i = nb - 1; i >= 0; i--) { /* U x = y */
                double t = x[i];
                for (int k = i + 1; k < nb; k++) t -= U(i, k) * x[k];
                x[i] = t / U(i, i);
            }
        }
        stc_finish(ni, nb, aii, aib, abi, bi, bb, aloc, ld, bloc, aschur, bschur);
    }
#undef U
    free(aii);
    free(abi);
    free(aib);
    free(abb);
    free(bi);
    free(bb);
    return info ? -4 : 0;
}

/* stc_bwd: stc.F90:664-679, Xb = BSchur - ASchur Xi */
void orc_stc_bwd(int ni, int nb, const double *aschur, const double *bschur, const double *xi,
                 double *xb)
{
    for (int b = 0; b < nb; b++) {
        double s = 0.0;
        for (int i = 0; i < ni; i++) s += aschur[b + (size_t)i * nb] * xi[i];
        xb[b] = -s + bschur[b];
    }
}

/* ------------------------------------------ BLAS / LAPACK form (CPU baseline) */

/* Host BLAS and LAPACK entry points (Fortran calling convention, 32-bit
 * integers), set by orc_set_blas from whatever library the caller found (the
 * reference links BLAS/LAPACK the same way: elem_opt.F90:131 DSYRK,
 * stc.F90:464-509 DGETRF/DGETRS/DGEMM).  Used only by orc_element_loop_blas. */
typedef void (*dsyrk_fn)(const char *, const char *, const int *, const int *, const double *,
                         const double *, const int *, const double *, double *, const int *,
                         size_t, size_t);
typedef void (*dgemm_fn)(const char *, const char *, const int *, const int *, const int *,
                         const double *, const double *, const int *, const double *,
                         const int *, const double *, double *, const int *, size_t, size_t);
typedef void (*dgetrf_fn)(const int *, const int *, double *, const int *, int *, int *);
typedef void (*dgetrs_fn)(const char *, const int *, const int *, const double *, const int *,
                          const int *, double *, const int *, int *, size_t);
static dsyrk_fn g_dsyrk;
static dgemm_fn g_dgemm;
static dgetrf_fn g_dgetrf;
static dgetrs_fn g_dgetrs;

int orc_set_blas(void *dsyrk, void *dgemm, void *dgetrf, void *dgetrs)
{
    g_dsyrk = (dsyrk_fn)dsyrk;
    g_dgemm = (dgemm_fn)dgemm;
    g_dgetrf = (dgetrf_fn)dgetrf;
    g_dgetrs = (dgetrs_fn)dgetrs;
    return (g_dsyrk && g_dgemm && g_dgetrf && g_dgetrs) ? 0 : -5;
}

/* DSYRK('U','N',n,k,1,A,n,beta,C,ldc): elem_opt.F90:131-132 */
static void syrk_blas(int n, int k, const double *a, double beta, double *c, int ldc)
{
    const double one = 1.0;
    g_dsyrk("U", "N", &n, &k, &one, a, &n, &beta, c, &ldc, 1, 1);
}

/* stc_fwd_gen through LAPACK and BLAS, in the reference's call order
 * (stc.F90:443-509): DGETRF on Abb, DGETRS for Bb then Abi, DGEMM for Bi and
 * Aii; packing and write-back as stc_fwd_wrapper (stc.F90:227-305) */
static int stc_fwd_gen_blas(int ni, int nb, double *aloc, int ld, double *bloc, double *aschur,
                            double *bschur)
{
    if (nb == 0) return 0;
    if (ni == 0) return -3;
    double *aii = malloc(sizeof(double) * (size_t)ni * ni);
    double *abi = malloc(sizeof(double) * (size_t)nb * ni);
    double *aib = malloc(sizeof(double) * (size_t)ni * nb);
    double *abb = malloc(sizeof(double) * (size_t)nb * nb);
    double *bi = malloc(sizeof(double) * ni), *bb = malloc(sizeof(double) * nb);
    int *piv = malloc(sizeof(int) * nb);
    int info = 0, one = 1;
    const double mone = -1.0, pone = 1.0;
    stc_pack(ni, nb, aloc, ld, bloc, aii, abi, aib, abb, bi, bb);
    g_dgetrf(&nb, &nb, abb, &nb, piv, &info);
    if (info == 0) g_dgetrs("N", &nb, &one, abb, &nb, piv, bb, &nb, &info, 1);
    if (info == 0) g_dgetrs("N", &nb, &ni, abb, &nb, piv, abi, &nb, &info, 1);
    if (info == 0) {
        g_dgemm("N", "N", &ni, &one, &nb, &mone, aib, &ni, bb, &nb, &pone, bi, &ni, 1, 1);
        g_dgemm("N", "N", &ni, &ni, &nb, &mone, aib, &ni, abi, &nb, &pone, aii, &ni, 1, 1);
        for (int j = 0; j < ni; j++)
            memcpy(&aloc[(size_t)j * ld], &aii[(size_t)j * ni], sizeof(double) * ni);
        memcpy(bloc, bi, sizeof(double) * ni);
        if (aschur) memcpy(aschur, abi, sizeof(double) * (size_t)nb * ni);
        if (bschur) memcpy(bschur, bb, sizeof(double) * nb);
    }
    free(aii);
    free(abi);
    free(aib);
    free(abb);
    free(bi);
    free(bb);
    free(piv);
    return info ? -4 : 0;
}

/* ------------------------------------------------------------ element loop */

/* mumps_sc Step 2 (omp_mumps/mumps_sc.F90:224-342): !$OMP DO SCHEDULE(DYNAMIC) over
 * elements, each running elem + stc_fwd on thread-private ALOC/BLOC workspaces */
static int element_loop_core(int space, int nel, const int *norder, const int *nori_e,
                             const int *nori_f, const double *xnod, int ngeo, const double *coef,
                             int load, int ldk, double *kout, double *bout, double *aii,
                             double *bi, double *aschur, double *bschur, int herm, int nthreads,
                             int blas)
{
    int status = 0;
#ifdef _OPENMP
    if (nthreads > 0) omp_set_num_threads(nthreads);
#endif
#pragma omp parallel
    {
        double *aloc = malloc(sizeof(double) * (size_t)ldk * ldk);
        double *bloc = malloc(sizeof(double) * (size_t)ldk);
#pragma omp for schedule(dynamic)
        for (int e = 0; e < nel; e++) {
            const int *no = &norder[19 * (size_t)e];
            int n, nb, st;
            if (space == 0) {
                n = orc_ndof_h1(no);
                nb = orc_nbubble_h1(no[18]);
                st = elem_h1_core(no, &nori_e[12 * (size_t)e], &nori_f[6 * (size_t)e],
                                  &xnod[3 * (size_t)ngeo * e], ngeo, coef[e], load, aloc, ldk,
                                  bloc, blas);
            } else {
                n = orc_ndof_hcurl(no);
                nb = orc_nbubble_hcurl(no[18]);
                st = elem_hcurl_core(no, &nori_e[12 * (size_t)e], &nori_f[6 * (size_t)e],
                                     &xnod[3 * (size_t)ngeo * e], ngeo, 1.0, coef[e], load, aloc,
                                     ldk, bloc, blas);
            }
            size_t o2 = (size_t)e * ldk * ldk, o1 = (size_t)e * ldk;
            if (kout)
                for (int j = 0; j < n; j++)
                    memcpy(&kout[o2 + (size_t)j * ldk], &aloc[(size_t)j * ldk], sizeof(double) * n);
            if (bout) memcpy(&bout[o1], bloc, sizeof(double) * n);
            int ni = n - nb;
            if (st >= 0 && aii) {
                double *as = aschur ? &aschur[o2] : NULL, *bs = bschur ? &bschur[o1] : NULL;
                int s2 = blas ? stc_fwd_gen_blas(ni, nb, aloc, ldk, bloc, as, bs)
                         : herm ? orc_stc_fwd_herm(ni, nb, aloc, ldk, bloc, as, bs)
                                : orc_stc_fwd_gen(ni, nb, aloc, ldk, bloc, as, bs);
                if (s2) st = s2;
                for (int j = 0; j < ni; j++)
                    memcpy(&aii[o2 + (size_t)j * ldk], &aloc[(size_t)j * ldk], sizeof(double) * ni);
                if (bi) memcpy(&bi[o1], bloc, sizeof(double) * ni);
            }
            if (st != 0) {
#pragma omp critical
                if (status == 0 || (st < 0 && status > 0)) status = st;
            }
        }
        free(aloc);
        free(bloc);
    }
    return status;
}

int orc_element_loop(int space, int nel, const int *norder, const int *nori_e, const int *nori_f,
                     const double *xnod, int ngeo, const double *coef, int load, int ldk,
                     double *kout, double *bout, double *aii, double *bi, double *aschur,
                     double *bschur, int herm, int nthreads)
{
    return element_loop_core(space, nel, norder, nori_e, nori_f, xnod, ngeo, coef, load, ldk, kout,
                             bout, aii, bi, aschur, bschur, herm, nthreads, 0);
}

/* the same loop with the element's SYRK and the condensation's LU, solves and
 * Schur GEMMs through the host BLAS/LAPACK of orc_set_blas, as the reference
 * builds it (elem_opt.F90:131-132, stc_fwd_gen stc.F90:443-509) */
int orc_element_loop_blas(int space, int nel, const int *norder, const int *nori_e,
                          const int *nori_f, const double *xnod, int ngeo, const double *coef,
                          int load, int ldk, double *kout, double *bout, double *aii, double *bi,
                          double *aschur, double *bschur, int nthreads)
{
    if (!g_dsyrk || !g_dgemm || !g_dgetrf || !g_dgetrs) return -5;
    return element_loop_core(space, nel, norder, nori_e, nori_f, xnod, ngeo, coef, load, ldk, kout,
                             bout, aii, bi, aschur, bschur, 0, nthreads, 1);
}

/* ------------------------------------------- solution-level element routines */

/* exact solutions of the manufactured problems, value and gradient
 * (problems' exact routine: POISSON/GALERKIN/common/exact.F90:30-77,
 * test/conv_pois.F90 exact; load selectors name the solution they belong to) */
void orc_exact_h1(int sel, const double *x, double *u, double *du)
{
    const double pi = 3.14159265358979323846;
    double sx = sin(pi * x[0]), sy = sin(pi * x[1]), sz = sin(pi * x[2]);
    double cx = cos(pi * x[0]), cy = cos(pi * x[1]), cz = cos(pi * x[2]);
    *u = du[0] = du[1] = du[2] = 0.0;
    switch (sel) {
    case ORC_LOAD_SINSINSIN:
        *u = sx * sy * sz;
        du[0] = pi * cx * sy * sz;
        du[1] = pi * sx * cy * sz;
        du[2] = pi * sx * sy * cz;
        break;
    case ORC_LOAD_CONV: {
        double sh = sinh(pi * x[1]), ch = cosh(pi * x[1]);
        *u = sx * sh * x[2] * x[2];
        du[0] = pi * cx * sh * x[2] * x[2];
        du[1] = pi * sx * ch * x[2] * x[2];
        du[2] = 2.0 * sx * sh * x[2];
        break;
    }
    case ORC_LOAD_POLY2: {
        double a = x[0] * (1.0 - x[0]), b = x[1] * (1.0 - x[1]), c = x[2] * (1.0 - x[2]);
        *u = a * b * c;
        du[0] = (1.0 - 2.0 * x[0]) * b * c;
        du[1] = a * (1.0 - 2.0 * x[1]) * c;
        du[2] = a * b * (1.0 - 2.0 * x[2]);
        break;
    }
    case ORC_LOAD_XYZ: /* test/poly_pois.F90 exact: u = x y z */
        *u = x[0] * x[1] * x[2];
        du[0] = x[1] * x[2];
        du[1] = x[0] * x[2];
        du[2] = x[0] * x[1];
        break;
    default:
        break;
    }
}

/* E* = (sin(pi y) sin(pi z), 0, 0) and its curl (0, pi sin(pi y) cos(pi z), -pi cos(pi y) sin(pi z)) */
void orc_exact_hcurl(int sel, const double *x, double *e, double *curl)
{
    const double pi = 3.14159265358979323846;
    for (int i = 0; i < 3; i++) e[i] = curl[i] = 0.0;
    if (sel == ORC_LOAD_CURL_SIN) {
        e[0] = sin(pi * x[1]) * sin(pi * x[2]);
        curl[1] = pi * sin(pi * x[1]) * cos(pi * x[2]);
        curl[2] = -pi * cos(pi * x[1]) * sin(pi * x[2]);
    }
}

/* element_error, H1 and H(curl) branches (element/util/compute_error.F90:226-400):
 * quadrature of set_3Dint with INTEGRATION = 2, soleval of the element dofs,
 * errorH = int |grad(u - u_h)|^2 + |u - u_h|^2 (rnormH the same of u), and for
 * H(curl) errorE = int |curl(E - E_h)|^2 + |E - E_h|^2 (rnormE of E). */
int orc_element_error(int space, const int *norder, const int *nori_e, const int *nori_f,
                      const double *xnod, int ngeo, const double *dof, int sel, double *err,
                      double *rnorm)
{
    if (orc_checkorder(norder) < 0) return -1;
    int nh = orc_ndof_h1(norder);
    if (ngeo > nh) return -2;
    int nint, status = 0;
    double xiloc[3 * ORC_MAX_NINT3], waloc[ORC_MAX_NINT3];
    orc_set_3dint(norder, 2, &nint, xiloc, waloc);
    double shap[ORC_MAXBRICKH], grad[3 * ORC_MAXBRICKH];
    double *she = (double *)malloc(sizeof(double) * 3 * ORC_MAXBRICKE);
    double *cue = (double *)malloc(sizeof(double) * 3 * ORC_MAXBRICKE);
    double e2 = 0.0, r2 = 0.0;
    for (int l = 0; l < nint; l++) {
        int nrh, nre;
        double x[3], dxdxi[9], dxidx[9], rjac;
        orc_shape3dh_hexa(&xiloc[3 * l], norder, nori_e, nori_f, &nrh, shap, grad);
        if (orc_geom3d(xnod, ngeo, shap, grad, x, dxdxi, dxidx, &rjac)) status = 1;
        double weight = waloc[l] * rjac;
        if (space == 0) {
            double uh = 0.0, gref[3] = {0.0, 0.0, 0.0}, duh[3], u, du[3];
            for (int k = 0; k < nrh; k++) {
                uh += dof[k] * shap[k];
                for (int i = 0; i < 3; i++) gref[i] += dof[k] * grad[3 * k + i];
            }
            for (int d = 0; d < 3; d++)
                duh[d] = gref[0] * dxidx[0 + 3 * d] + gref[1] * dxidx[1 + 3 * d] + gref[2] * dxidx[2 + 3 * d];
            orc_exact_h1(sel, x, &u, du);
            for (int j = 0; j < 3; j++) {
                e2 += (du[j] - duh[j]) * (du[j] - duh[j]) * weight;
                r2 += du[j] * du[j] * weight;
            }
            e2 += (u - uh) * (u - uh) * weight;
            r2 += u * u * weight;
        } else {
            orc_shape3de_hexa(&xiloc[3 * l], norder, nori_e, nori_f, &nre, she, cue);
            double eref[3] = {0.0, 0.0, 0.0}, cref[3] = {0.0, 0.0, 0.0}, eh[3], ch[3], ex[3], cx[3];
            for (int k = 0; k < nre; k++)
                for (int i = 0; i < 3; i++) {
                    eref[i] += dof[k] * she[3 * k + i];
                    cref[i] += dof[k] * cue[3 * k + i];
                }
            for (int d = 0; d < 3; d++) {
                eh[d] = eref[0] * dxidx[0 + 3 * d] + eref[1] * dxidx[1 + 3 * d] + eref[2] * dxidx[2 + 3 * d];
                ch[d] = (dxdxi[d + 0] * cref[0] + dxdxi[d + 3] * cref[1] + dxdxi[d + 6] * cref[2]) / rjac;
            }
            orc_exact_hcurl(sel, x, ex, cx);
            for (int j = 0; j < 3; j++) {
                e2 += ((cx[j] - ch[j]) * (cx[j] - ch[j]) + (ex[j] - eh[j]) * (ex[j] - eh[j])) * weight;
                r2 += (cx[j] * cx[j] + ex[j] * ex[j]) * weight;
            }
        }
    }
    free(she);
    free(cue);
    *err = e2;
    *rnorm = r2;
    return status;
}

/* Element mass matrix M_e (H1: int phi_k phi_l; H(curl): int E_k . E_l) with
 * the element quadrature of elem_opt (set_3D_int, INTEGRATION = 0), column-major
 * with leading dimension ld.  Used for the L2 objective of the adjoint
 * problem: J_e = 1/2 d^T M_e d and its gradient M_e d. */
int orc_elem_mass(int space, const int *norder, const int *nori_e, const int *nori_f,
                  const double *xnod, int ngeo, double *m, int ld)
{
    if (orc_checkorder(norder) < 0) return -1;
    int nh = orc_ndof_h1(norder), n = space == 0 ? nh : orc_ndof_hcurl(norder);
    if (ngeo > nh) return -2;
    int nint, status = 0;
    double xiloc[3 * ORC_MAX_NINT3], waloc[ORC_MAX_NINT3];
    orc_set_3d_int(norder, nori_f, &nint, xiloc, waloc);
    double shap[ORC_MAXBRICKH], grad[3 * ORC_MAXBRICKH];
    double *she = (double *)malloc(sizeof(double) * 3 * ORC_MAXBRICKE);
    double *cue = (double *)malloc(sizeof(double) * 3 * ORC_MAXBRICKE);
    double *f = (double *)malloc(sizeof(double) * 3 * ORC_MAXBRICKE);
    for (int j = 0; j < n; j++)
        for (int i = 0; i < n; i++) m[i + (size_t)j * ld] = 0.0;
    for (int l = 0; l < nint; l++) {
        int nrh, nre;
        double x[3], dxdxi[9], dxidx[9], rjac;
        orc_shape3dh_hexa(&xiloc[3 * l], norder, nori_e, nori_f, &nrh, shap, grad);
        if (orc_geom3d(xnod, ngeo, shap, grad, x, dxdxi, dxidx, &rjac)) status = 1;
        double weight = waloc[l] * rjac;
        if (space == 0) {
            for (int j = 0; j < n; j++)
                for (int i = 0; i < n; i++) m[i + (size_t)j * ld] += shap[i] * shap[j] * weight;
        } else {
            orc_shape3de_hexa(&xiloc[3 * l], norder, nori_e, nori_f, &nre, she, cue);
            for (int k = 0; k < n; k++)
                for (int d = 0; d < 3; d++)
                    f[3 * k + d] = she[3 * k + 0] * dxidx[0 + 3 * d] + she[3 * k + 1] * dxidx[1 + 3 * d] +
                                   she[3 * k + 2] * dxidx[2 + 3 * d];
            for (int j = 0; j < n; j++)
                for (int i = 0; i < n; i++)
                    m[i + (size_t)j * ld] += (f[3 * i] * f[3 * j] + f[3 * i + 1] * f[3 * j + 1] +
                                              f[3 * i + 2] * f[3 * j + 2]) * weight;
        }
    }
    free(she);
    free(cue);
    free(f);
    return status;
}
