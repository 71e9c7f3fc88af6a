// hp3d_gpu.hip -- MI355X (gfx950) element computation for hp3D hexahedra.
//
// Kernels
//   elem_kernel   one workgroup per (element, 128x128 block of K): 1D factor
//                 tables and per-point geometry in LDS, A = B^T(geometry)
//                 streamed through LDS in chunks of quadrature points, and
//                 K += A A^T on MFMA (v_mfma_f64_16x16x4f64 or
//                 v_mfma_f32_16x16x4f32), upper-triangle tiles only, written
//                 out in both triangles with coalesced column segments.
//   elem_single_kernel  one workgroup per element with N <= 128: K kept in
//                 registers over the quadrature, then (nb <= 32) the bubble
//                 block condensed in the same launch.
//   stc_strip_kernel  one workgroup of 8 waves per element: left-looking
//                 blocked Cholesky of the bubble block, then [Abi | Bb] in
//                 register-resident strips through the forward and backward
//                 sweeps with Aii - Aib X accumulated on the way, FP64 MFMA.
//   field_kernel  element integrals of finite-element fields (element_error,
//                 L2 norm and residual, adjoint coefficient gradient).
//   gather_sum_kernel, csr_spmv_kernel  device assembly and the global operator.
//   stc_bwd_kernel  xb = BSchur - ASchur xi.
//
// The reference computes the same quantities per element on one CPU thread:
// POISSON/GALERKIN/elem_opt.F90:88-137 (A_TEST fill + DSYRK), the real form of
// MAXWELL/GALERKIN/elem_opt.F90:88-149, modules/stc.F90:182-416.
#include <hip/hip_runtime.h>

#include <math.h>
#include <stdint.h>
#include <stdio.h>
#include <string.h>

#include <algorithm>
#include <mutex>
#include <type_traits>
#include <vector>

#include "../../include/hp3d_gpu.h"
#include "hp3d_tables.h"

typedef double d4_t __attribute__((ext_vector_type(4)));
typedef float f4_t __attribute__((ext_vector_type(4)));

namespace hp3d {

constexpr int WG = 256;        // threads per workgroup (4 waves)
constexpr int BR = 128;        // rows of K per block (8 MFMA tiles)
constexpr int KC = 24;         // A columns per chunk (8 points H1, 4 points H(curl))
constexpr int LDA = 2 * BR + 16; // LDS stride of one A column (+16: the four lane
                                 // groups of an MFMA operand read hit disjoint banks)
constexpr int NPT = 8;         // padded max Gauss points per direction
constexpr int MAXQ = (MAXP + 1) * (MAXP + 1) * (MAXP + 1);
constexpr int MAXN_H1 = (MAXP + 1) * (MAXP + 1) * (MAXP + 1);
constexpr int MAXN_E = 3 * MAXP * (MAXP + 1) * (MAXP + 1);

__constant__ double c_gx[10][10] = HP3D_GAUSS_X;
__constant__ double c_gw[10][10] = HP3D_GAUSS_W;

// ------------------------------------------------------------------ helpers

// 1D factor tables at the Gauss points of the three directions (values and
// d/dxi), one work item per (direction, point, flip) so that one wave fills
// all of them: 2 + 2 (MAXP - 1) + 2 MAXP entries per (direction, point)
// Gauss node q of the n-point rule by selects over compile-time constants:
// no memory load on the table fill's critical path
__device__ __forceinline__ double gauss_x_sel(int n, int q)
{
    constexpr double gx[10][10] = HP3D_GAUSS_X;
    const int key = 16 * n + q;
    double v = 0.0;
#pragma unroll
    for (int nn = 1; nn <= 10; ++nn)
#pragma unroll
        for (int qq = 0; qq < nn; ++qq) v = key == 16 * nn + qq ? gx[nn - 1][qq] : v;
    return v;
}

__device__ __forceinline__ void fill_table_item(int d, int q, int flip, int n, double (*tv)[NTAB * NPT],
                                                double (*td)[NTAB * NPT])
{
    double x = 0.5 * (1.0 + gauss_x_sel(n, q));
    if (flip == 0) {
        tv[d][q * NTAB + 0] = 1.0 - x;
        td[d][q * NTAB + 0] = -1.0;
        tv[d][q * NTAB + 1] = x;
        td[d][q * NTAB + 1] = 1.0;
    }
    // integrated Legendre phi_i(s) = (P_i(2s-1) - P_{i-2}(2s-1)) / (4i-2),
    // d phi_i / ds = P_{i-1}(2s-1)   (PolyILegendre, Idec = .true.)
    double s = flip ? 1.0 - x : x;
    double y = 2.0 * s - 1.0, p[MAXP + 1];
    // the divisions by the small integers as products with their
    // (compile-time) reciprocals: within an ulp of the quotients
    p[0] = 1.0;
    p[1] = y;
#pragma unroll
    for (int i = 2; i <= MAXP; ++i)
        p[i] = ((double)(2 * i - 1) * y * p[i - 1] - (double)(i - 1) * p[i - 2]) * (1.0 / (double)i);
#pragma unroll
    for (int i = 2; i <= MAXP; ++i) {
        tv[d][q * NTAB + t_phi(i, flip)] = (p[i] - p[i - 2]) * (1.0 / (double)(4 * i - 2));
        td[d][q * NTAB + t_phi(i, flip)] = flip ? -p[i - 1] : p[i - 1];
    }
    // scaled Legendre P_i(s; t) with t = s0 + s1 evaluated as the reference
    // does (HomLegendre of the oriented pair)
    double t = (1.0 - x) + x;
    double yy = 2.0 * s - t, tt = t * t, r[MAXP];
    r[0] = 1.0;
    r[1] = yy;
#pragma unroll
    for (int i = 2; i < MAXP; ++i)
        r[i] = ((double)(2 * i - 1) * yy * r[i - 1] - (double)(i - 1) * tt * r[i - 2]) * (1.0 / (double)i);
    for (int i = 0; i < MAXP; ++i) {
        tv[d][q * NTAB + t_leg(i, flip)] = r[i];
        td[d][q * NTAB + t_leg(i, flip)] = 0.0;
    }
}

// all three directions' tables; `lane` runs over one wave (64 >= 3 x NPT x 2)
__device__ __forceinline__ void fill_tables3(const int *nq, int lane, double (*tv)[NTAB * NPT],
                                             double (*td)[NTAB * NPT])
{
    int d = lane / (2 * NPT), q = (lane / 2) % NPT, flip = lane & 1;
    if (d < 3 && q < nq[d]) fill_table_item(d, q, flip, nq[d], tv, td);
}

// the same, one direction per wave (kernels that fill them with waves 0-2)
__device__ __forceinline__ void fill_tables(int n, double *tv, double *td)
{
    double (*v)[NTAB * NPT] = reinterpret_cast<double (*)[NTAB * NPT]>(tv);
    double (*g)[NTAB * NPT] = reinterpret_cast<double (*)[NTAB * NPT]>(td);
    int l = threadIdx.x & 63, q = l >> 1, flip = l & 1;
    if (q < n) fill_table_item(0, q, flip, n, v, g);
}

struct Pt { int l1, l2, l3; };
__device__ inline Pt point_ix(int l, const int *nq)
{
    Pt p;
    p.l1 = l % nq[0];
    p.l2 = (l / nq[0]) % nq[1];
    p.l3 = l / (nq[0] * nq[1]);
    return p;
}

// shared state of one element (built by the first wave)
struct ElemShared {
    ElemOrders o;
    int n, nb, nq[3], nint, nh, ok, negjac;
};

// ------------------------------------------------------------ element kernel

template <typename T> struct Mfma;
// 16x16x4 MFMA.  Operands: lane l supplies A[l%16][l/16] and B[l/16][l%16].
// Result register j of lane l holds D[row(l, j)][l%16], with row(l, j) =
// l/16 + 4j for the f64 instruction and 4(l/16) + j for the f32 one (checked
// on gfx950 with a single-wave layout probe).
template <> struct Mfma<double> {
    typedef d4_t acc_t;
    __device__ static inline acc_t mma(double a, double b, acc_t c)
    {
        return __builtin_amdgcn_mfma_f64_16x16x4f64(a, b, c, 0, 0, 0);
    }
    __device__ static inline int row(int lane, int j) { return (lane >> 4) + 4 * j; }
};
template <> struct Mfma<float> {
    typedef f4_t acc_t;
    __device__ static inline acc_t mma(float a, float b, acc_t c)
    {
        return __builtin_amdgcn_mfma_f32_16x16x4f32(a, b, c, 0, 0, 0);
    }
    __device__ static inline int row(int lane, int j) { return 4 * (lane >> 4) + j; }
};

struct ElemArgs {
    int space, nel, ngeo, load, ld_k, max_ndof; // max_ndof: the caller's bound on N_e
    const int *norder, *nori_e, *nori_f;
    const double *xnod, *coef;
    const long long *off_k, *off_v;
    void *kout;
    double *bout;
    int *status;
    double *gws; // per-element geometry slabs of the block-tiled path
};

// geometry doubles per quadrature point, and the per-element slab of the
// block-tiled path: MAXQ points, then the negative-Jacobian flag
__host__ __device__ constexpr int geo_stride(int space) { return space == 0 ? 10 : 21; }
__host__ __device__ constexpr size_t geo_slab(int space) { return (size_t)MAXQ * geo_stride(space) + 2; }
constexpr int XS_MAX = 27; // geometry nodes staged in LDS by the geometry code

__device__ inline void getf(int load, const double *x, double *f)
{
    const double pi = 3.14159265358979323846;
    f[0] = f[1] = f[2] = 0.0;
    switch (load) {
    // sin(pi x) as sinpi(x): no reduction of a large argument is needed, and
    // the product pi x is not rounded first (the values agree to an ulp or two)
    case HP3D_LOAD_SINSINSIN: f[0] = 3.0 * pi * pi * sinpi(x[0]) * sinpi(x[1]) * sinpi(x[2]); break;
    case HP3D_LOAD_CONV: f[0] = -2.0 * sinpi(x[0]) * sinh(pi * x[1]); break;
    case HP3D_LOAD_ONE: f[0] = 1.0; break;
    case HP3D_LOAD_POLY2: {
        double u0 = x[0] * (1.0 - x[0]), u1 = x[1] * (1.0 - x[1]), u2 = x[2] * (1.0 - x[2]);
        f[0] = 2.0 * (u1 * u2 + u0 * u2 + u0 * u1);
        break;
    }
    case HP3D_LOAD_CURL_SIN: f[0] = (1.0 + 2.0 * pi * pi) * sinpi(x[1]) * sinpi(x[2]); break;
    default: break;
    }
}

// (I, J), I <= J, of block b of the upper block triangle enumerated column by
// column: column J holds blocks I = 0..J
__device__ inline void block_ij(int b, int &I, int &J)
{
    J = (int)((sqrt(8.0 * b + 1.0) - 1.0) * 0.5);
    while ((J + 1) * (J + 2) / 2 <= b) ++J;
    while (J * (J + 1) / 2 > b) --J;
    I = b - J * (J + 1) / 2;
}

// the trilinear / isoparametric map at point p of the tensor rule with nq
// points per direction (geom3D.F90:77-91, geom.F90:57-113): x, J (jm[c][i] =
// d x_c / d xi_i), det J, inv[i][j] = d xi_i / d x_j and the Gauss weight
__device__ __forceinline__ void point_map(Pt p, const int *nq, const double (*tv)[NTAB * NPT],
                                          const double (*td)[NTAB * NPT], const unsigned *dgeo,
                                          const double *xn, int ngeo, double x[3], double jm[3][3],
                                          double inv[3][3], double &det, double &w)
{
    for (int c = 0; c < 3; ++c) {
        x[c] = 0.0;
        jm[c][0] = jm[c][1] = jm[c][2] = 0.0;
    }
    for (int k = 0; k < ngeo; ++k) {
        unsigned d = dgeo[k];
        int i0 = desc_ix(d, 0), i1 = desc_ix(d, 1), i2 = desc_ix(d, 2);
        double v0 = tv[0][p.l1 * NTAB + i0], v1 = tv[1][p.l2 * NTAB + i1], v2 = tv[2][p.l3 * NTAB + i2];
        double g0 = td[0][p.l1 * NTAB + i0] * v1 * v2;
        double g1 = v0 * td[1][p.l2 * NTAB + i1] * v2;
        double g2 = v0 * v1 * td[2][p.l3 * NTAB + i2];
        double ph = v0 * v1 * v2;
        for (int c = 0; c < 3; ++c) {
            double xc = xn[3 * k + c];
            x[c] += xc * ph;
            jm[c][0] += xc * g0;
            jm[c][1] += xc * g1;
            jm[c][2] += xc * g2;
        }
    }
    det = jm[0][0] * jm[1][1] * jm[2][2] + jm[1][0] * jm[2][1] * jm[0][2] +
          jm[2][0] * jm[0][1] * jm[1][2] - jm[2][0] * jm[1][1] * jm[0][2] -
          jm[0][0] * jm[2][1] * jm[1][2] - jm[1][0] * jm[0][1] * jm[2][2];
    const double rdet = 1.0 / det;
    inv[0][0] = (jm[1][1] * jm[2][2] - jm[2][1] * jm[1][2]) * rdet;
    inv[1][0] = (-jm[1][0] * jm[2][2] + jm[2][0] * jm[1][2]) * rdet;
    inv[2][0] = (jm[1][0] * jm[2][1] - jm[2][0] * jm[1][1]) * rdet;
    inv[0][1] = (jm[2][1] * jm[0][2] - jm[0][1] * jm[2][2]) * rdet;
    inv[1][1] = (jm[0][0] * jm[2][2] - jm[2][0] * jm[0][2]) * rdet;
    inv[2][1] = (-jm[0][0] * jm[2][1] + jm[2][0] * jm[0][1]) * rdet;
    inv[0][2] = (jm[0][1] * jm[1][2] - jm[1][1] * jm[0][2]) * rdet;
    inv[1][2] = (-jm[0][0] * jm[1][2] + jm[1][0] * jm[0][2]) * rdet;
    inv[2][2] = (jm[0][0] * jm[1][1] - jm[1][0] * jm[0][1]) * rdet;
    w = (0.5 * c_gw[nq[0] - 1][p.l1]) * (0.5 * c_gw[nq[1] - 1][p.l2]) * (0.5 * c_gw[nq[2] - 1][p.l3]);
}

// geometry of one quadrature point (FP64): x, J, det J, J^-1 folded with the
// weights into GS doubles at gp (geom3D.F90:77-91, geom.F90:57-113,
// elem_opt.F90:101-122).  Returns true when det J < 0.
template <int SPACE, bool SPLIT = false>
__device__ __forceinline__ bool point_geometry(const ElemShared &es, int l, const double (*tv)[NTAB * NPT],
                               const double (*td)[NTAB * NPT], const unsigned *dgeo,
                               const double *xn, int ngeo, double coef, int load, double *gp)
{
    double x[3], jm[3][3], inv[3][3], det, w;
    point_map(point_ix(l, es.nq), es.nq, tv, td, dgeo, xn, ngeo, x, jm, inv, det, w);
    double wd = det * w;
    if (SPACE == 0) {
        double s = sqrt(coef * wd);
        for (int i = 0; i < 3; ++i)
            for (int j = 0; j < 3; ++j) gp[3 * i + j] = inv[i][j] * s; // [ref][phys]
        if (SPLIT) {
            gp[9] = wd; // the load f(x) comes from point_load_h1 into gp[10]
        } else {
            double f[3];
            getf(load, x, f);
            gp[9] = f[0] * wd;
        }
        return det < 0.0;
    }
    double f[3];
    getf(load, x, f);
    {
        double sc = sqrt(wd) / det, sm = sqrt(coef * wd); // mu = 1
        for (int i = 0; i < 3; ++i)
            for (int j = 0; j < 3; ++j) {
                gp[3 * i + j] = jm[i][j] * sc;      // curl: [phys][ref]
                gp[9 + 3 * i + j] = inv[i][j] * sm; // mass: [ref][phys]
            }
        for (int c = 0; c < 3; ++c)
            gp[18 + c] = (inv[c][0] * f[0] + inv[c][1] * f[1] + inv[c][2] * f[2]) * wd;
    }
    return det < 0.0;
}

// The scalar load f(x) of the H1 form at point l: x summed over the geometry
// nodes in point_map's order (so it is the same x), then getf.  Lets the
// 64-row fused kernel evaluate the load on another wave than the Jacobian.
__device__ __forceinline__ double point_load_h1(const ElemShared &es, int l, const double (*tv)[NTAB * NPT],
                                                const unsigned *dgeo, const double *xn, int ngeo, int load)
{
    const Pt p = point_ix(l, es.nq);
    double x[3] = {0.0, 0.0, 0.0};
    for (int k = 0; k < ngeo; ++k) {
        unsigned d = dgeo[k];
        int i0 = desc_ix(d, 0), i1 = desc_ix(d, 1), i2 = desc_ix(d, 2);
        double v0 = tv[0][p.l1 * NTAB + i0], v1 = tv[1][p.l2 * NTAB + i1], v2 = tv[2][p.l3 * NTAB + i2];
        double ph = v0 * v1 * v2;
        for (int c = 0; c < 3; ++c) x[c] += xn[3 * k + c] * ph;
    }
    double f[3];
    getf(load, x, f);
    return f[0];
}

// for the H(curl) component cp (0..2) with sign sg and the reference gradient
// gr of the scalar factor: cu = curl of (factor e_cp) = gr x e_cp scaled by sg,
// mg = row cp of the mass factor gp[9..17] and fl = gp[18 + cp].  With LOCAL
// (gp a local copy) no array is indexed by the run-time component: 0/1
// weights instead of a selected address keep gp in registers (its entries are
// finite); with gp in LDS one indexed load is cheaper.
template <bool LOCAL>
__device__ __forceinline__ void curl_mass(int cp, double sg, const double gr[3], const double *gp,
                                          double cu[3], double mg[3], double &fl)
{
    cu[0] = cp == 0 ? 0.0 : (cp == 1 ? -sg * gr[2] : sg * gr[1]);
    cu[1] = cp == 0 ? sg * gr[2] : (cp == 1 ? 0.0 : -sg * gr[0]);
    cu[2] = cp == 0 ? -sg * gr[1] : (cp == 1 ? sg * gr[0] : 0.0);
    if (LOCAL) {
        const double w0 = cp == 0 ? 1.0 : 0.0, w1 = cp == 1 ? 1.0 : 0.0, w2 = cp == 2 ? 1.0 : 0.0;
#pragma unroll
        for (int c = 0; c < 3; ++c) mg[c] = w0 * gp[9 + c] + w1 * gp[12 + c] + w2 * gp[15 + c];
        fl = w0 * gp[18] + w1 * gp[19] + w2 * gp[20];
    } else {
#pragma unroll
        for (int c = 0; c < 3; ++c) mg[c] = gp[9 + 3 * cp + c];
        fl = gp[18 + cp];
    }
}

// the NC entries of row d (descriptor) at point p of the A matrix
// (A_TEST of elem_opt.F90:122, or the [curl | mass] pair of the H(curl) form)
template <int SPACE>
__device__ __forceinline__ void a_entries(unsigned d, Pt p, const double (*tv)[NTAB * NPT],
                                 const double (*td)[NTAB * NPT], const double *gp, double *out)
{
    int i0 = desc_ix(d, 0), i1 = desc_ix(d, 1), i2 = desc_ix(d, 2);
    double v0 = tv[0][p.l1 * NTAB + i0], v1 = tv[1][p.l2 * NTAB + i1], v2 = tv[2][p.l3 * NTAB + i2];
    double d0 = td[0][p.l1 * NTAB + i0], d1 = td[1][p.l2 * NTAB + i1], d2 = td[2][p.l3 * NTAB + i2];
    double gr[3] = {d0 * v1 * v2, v0 * d1 * v2, v0 * v1 * d2};
    if (SPACE == 0) {
#pragma unroll
        for (int c = 0; c < 3; ++c) out[c] = gr[0] * gp[c] + gr[1] * gp[3 + c] + gr[2] * gp[6 + c];
    } else {
        int cp = desc_comp(d);
        double sg = desc_neg(d) ? -1.0 : 1.0;
        double val = sg * v0 * v1 * v2;
        double cu[3], mg[3], fl;
        curl_mass<false>(cp, sg, gr, gp, cu, mg, fl);
#pragma unroll
        for (int c = 0; c < 3; ++c) {
            out[c] = gp[3 * c] * cu[0] + gp[3 * c + 1] * cu[1] + gp[3 * c + 2] * cu[2];
            out[3 + c] = val * mg[c];
        }
    }
}

// A entries and the load contribution of row d at one point (one pass over
// the 1D tables for both)
template <int SPACE, bool LOCAL = false>
__device__ __forceinline__ double a_entries_load(unsigned d, Pt p, const double (*tv)[NTAB * NPT],
                                                 const double (*td)[NTAB * NPT], const double *gp,
                                                 double *out)
{
    int i0 = desc_ix(d, 0), i1 = desc_ix(d, 1), i2 = desc_ix(d, 2);
    double v0 = tv[0][p.l1 * NTAB + i0], v1 = tv[1][p.l2 * NTAB + i1], v2 = tv[2][p.l3 * NTAB + i2];
    double d0 = td[0][p.l1 * NTAB + i0], d1 = td[1][p.l2 * NTAB + i1], d2 = td[2][p.l3 * NTAB + i2];
    double v12 = v1 * v2;
    double gr[3] = {d0 * v12, v0 * d1 * v2, v0 * v1 * d2};
    double ph = v0 * v12;
    if (SPACE == 0) {
#pragma unroll
        for (int c = 0; c < 3; ++c) out[c] = gr[0] * gp[c] + gr[1] * gp[3 + c] + gr[2] * gp[6 + c];
        return ph * gp[9];
    } else {
        int cp = desc_comp(d);
        double sg = desc_neg(d) ? -1.0 : 1.0;
        double val = sg * ph;
        double cu[3], mg[3], fl;
        curl_mass<LOCAL>(cp, sg, gr, gp, cu, mg, fl);
#pragma unroll
        for (int c = 0; c < 3; ++c) {
            out[c] = gp[3 * c] * cu[0] + gp[3 * c + 1] * cu[1] + gp[3 * c + 2] * cu[2];
            out[3 + c] = val * mg[c];
        }
        return val * fl;
    }
}

// load-vector contribution of row d at one point
template <int SPACE>
__device__ __forceinline__ double load_entry(unsigned d, Pt p, const double (*tv)[NTAB * NPT], const double *gp)
{
    double v = tv[0][p.l1 * NTAB + desc_ix(d, 0)] * tv[1][p.l2 * NTAB + desc_ix(d, 1)] *
               tv[2][p.l3 * NTAB + desc_ix(d, 2)];
    if (SPACE == 0) return v * gp[9];
    return (desc_neg(d) ? -v : v) * gp[18 + desc_comp(d)];
}

// Make this wave's LDS writes visible to its other lanes: wait for the LDS
// counter only (s_waitcnt lgkmcnt(0), vmcnt/expcnt left at their maxima so
// outstanding global stores are not waited for) and keep the compiler from
// moving LDS accesses across.
__device__ __forceinline__ void wave_sync_lds()
{
    __builtin_amdgcn_s_waitcnt(0xC07F);
    __builtin_amdgcn_wave_barrier();
}

// Workgroup barrier for LDS traffic only: the same LDS-counter wait, then
// s_barrier, without the workgroup-scope fence of __syncthreads() that would
// also wait for every outstanding global store.  For phases whose threads
// share data through LDS alone (their global writes are outputs nobody in the
// workgroup reads back).
__device__ __forceinline__ void lds_barrier()
{
    __asm__ volatile("" ::: "memory");
    __builtin_amdgcn_s_waitcnt(0xC07F);
    __builtin_amdgcn_s_barrier();
    __asm__ volatile("" ::: "memory");
}

// element header: orders, counts, quadrature sizes (thread 0), plus status on
// failure.  Returns false (uniformly) if the element is invalid.
template <int SPACE>
__device__ __forceinline__ bool element_setup(const ElemArgs &a, int e, ElemShared &es, const int t = threadIdx.x)
{
    // The 37 header integers (norder, nori_e, nori_f) come in through the
    // scalar cache: wave-uniform loads into SGPRs, which the decoding lane
    // reads directly.  They are waited for on the scalar counter, so neither
    // an LDS round trip nor outstanding vector stores (a persistent grid's
    // previous element) hold them up.
    typedef const __attribute__((address_space(4))) int *sint_t;
    if (t < 64) {
        int hdr[37];
        const sint_t pn = (sint_t)(uintptr_t)(a.norder + 19 * (size_t)e);
        const sint_t pe = (sint_t)(uintptr_t)(a.nori_e + 12 * (size_t)e);
        const sint_t pf = (sint_t)(uintptr_t)(a.nori_f + 6 * (size_t)e);
#pragma unroll
        for (int i = 0; i < 19; ++i) hdr[i] = pn[i];
#pragma unroll
        for (int i = 0; i < 12; ++i) hdr[19 + i] = pe[i];
#pragma unroll
        for (int i = 0; i < 6; ++i) hdr[31 + i] = pf[i];
        if (t == 0) {
            ElemShared s;
            s.ok = decode_orders(hdr, hdr + 19, hdr + 31, s.o);
            s.n = s.nb = s.nint = s.nh = 0;
            s.nq[0] = s.nq[1] = s.nq[2] = 1;
            if (s.ok) {
                dof_counts(SPACE, s.o, s.n, s.nb);
                int nbh;
                dof_counts(0, s.o, s.nh, nbh);
                quad_counts(s.o, s.nq);
                s.nint = s.nq[0] * s.nq[1] * s.nq[2];
                if (a.ngeo > s.nh) s.ok = 0;
            }
            s.negjac = 0;
            es = s;
        }
    }
    lds_barrier();
    return es.ok != 0;
}

// A lane's FP64 value read by the whole wave (two 32-bit readlanes; `l` is
// wave-uniform).
__device__ __forceinline__ double readlane_f64(double v, int l)
{
    const long long b = __double_as_longlong(v);
    const int lo = __builtin_amdgcn_readlane((int)b, l);
    const int hi = __builtin_amdgcn_readlane((int)(b >> 32), l);
    return __longlong_as_double(((long long)hi << 32) | (unsigned)lo);
}

// 1 / d for d > 0 from the hardware reciprocal estimate (v_rcp_f64) and two
// Newton-Raphson steps r <- r + r (1 - d r): a few dependent operations where
// the IEEE division is a dozen, on the elimination's critical path (the
// result is within an ulp or two of 1 / d)
__device__ __forceinline__ double recip_pos(double d)
{
    double r = __builtin_amdgcn_rcp(d);
    r = fma(r, fma(-d, r, 1.0), r);
    return fma(r, fma(-d, r, 1.0), r);
}

// The steps of eliminate_wave over the first NR rows of the lane's column
// (nb <= NR; rows nb.. are padding, whose multiplier 0 leaves them as they
// are).  Step j updates row j + 1 first and starts the reciprocal of the
// next pivot, then the remaining rows, with no branch between them, so the
// division overlaps the row updates.
template <int NBX, int NR>
__device__ __forceinline__ bool eliminate_rows(double (&a)[NBX], int nb, int lane, double &rp_own)
{
    bool bad = false;
    const double piv0 = readlane_f64(a[0], 0);
    double rp = piv0 > 0.0 ? recip_pos(piv0) : 0.0;
    bad |= nb > 0 && !(piv0 > 0.0);
#pragma unroll
    for (int j = 0; j < NR; ++j) {
        if (j < nb) {
            if (lane == j) rp_own = rp;
            const double u = a[j] * rp; // this lane's entry of row j over D(j)
            double rpn = 0.0;
            if (j + 1 < NR) {
                a[j + 1] = fma(-readlane_f64(a[j + 1], j), u, a[j + 1]);
                const double pn = readlane_f64(a[j + 1], j + 1);
                rpn = pn > 0.0 ? recip_pos(pn) : 0.0;
                bad |= j + 1 < nb && !(pn > 0.0);
            }
#pragma unroll
            for (int i = j + 2; i < NR; ++i) a[i] = fma(-readlane_f64(a[i], j), u, a[i]);
            rp = rpn;
        }
    }
    return bad;
}

// Gaussian elimination without pivoting of [Abb | I] by one wave, in
// registers (Abb = L_u D L_u^T is SPD: every pivot D(j) > 0).  Lane c < NBX
// holds column c of Abb padded to NBX x NBX with the identity, lane NBX + c
// column c of I, so step j is one readlane of the pivot and one of each
// multiplier numerator from lane j, then a multiply-add per row in every
// lane: the row-wise elimination kb(i,:) -= kb(i,j) (kb(j,:) / D(j)), with
// no workgroup barrier.  Abb comes from `kb`
// (bubble row i at kb + i * ld, Abb columns from ni); the identity half ends
// as L_u^-1, written to li (rows < nb, leading dimension lw), and
// D(j)^-1/2 to rs[j].  Returns true when a pivot is not positive (its 1 / D is
// then taken as 0).
template <int NBX>
__device__ __forceinline__ bool eliminate_wave(const double *kb, int ld, int ni, int nb, double *li, int lw,
                                               double *rs, int lane)
{
    static_assert(2 * NBX <= 64, "one lane per column of [Abb | I]");
    const int c = lane % NBX;
    const bool left = lane < NBX;
    double a[NBX];
#pragma unroll
    for (int i = 0; i < NBX; ++i)
        a[i] = (left && i < nb && c < nb) ? kb[i * ld + ni + c] : (i == c ? 1.0 : 0.0);
    double rp_own = 0.0; // lane j keeps 1 / D(j)
    const bool bad = nb > NBX / 2 ? eliminate_rows<NBX, NBX>(a, nb, lane, rp_own)
                                  : eliminate_rows<NBX, NBX / 2>(a, nb, lane, rp_own);
    if (!left && lane < 2 * NBX)
#pragma unroll
        for (int i = 0; i < NBX; ++i)
            if (i < nb) li[i * lw + c] = a[i];
    if (lane < nb) rs[lane] = sqrt(rp_own);
    return bad;
}

// write an accumulator tile (rows R.., cols C..) of a symmetric matrix into
// column-major `out` (leading dimension ld, valid extent lim x lim), both the
// tile and its mirror; `stage` is this wave's 16 x 17 LDS scratch.
template <typename V, typename S, typename M>
__device__ inline void write_sym_tile(V acc, int R, int C, int lim, S *out, size_t ld, S *stage, int lane)
{
    int rr = C + (lane & 15);
#pragma unroll
    for (int j = 0; j < 4; ++j) {
        int cc = R + M::row(lane, j);
        if (rr < lim && cc < lim) out[(size_t)cc * ld + rr] = (S)acc[j];
    }
    if (R != C) {
#pragma unroll
        for (int j = 0; j < 4; ++j) stage[M::row(lane, j) * 17 + (lane & 15)] = (S)acc[j];
        wave_sync_lds();
        int r2 = R + (lane & 15);
#pragma unroll
        for (int j = 0; j < 4; ++j) {
            int c2 = C + 4 * (lane >> 4) + j;
            S v = stage[(lane & 15) * 17 + 4 * (lane >> 4) + j];
            if (r2 < lim && c2 < lim) out[(size_t)c2 * ld + r2] = v;
        }
        wave_sync_lds();
    }
}



// write_sym_tile without the LDS stage: the mirrored tile as there
// (coalesced over the 16 lanes of a row group), the tile itself as this
// lane's four entries of column C + lane % 16 (consecutive rows in the FP32
// layout: one 16- or 32-byte run per lane)
template <typename V, typename S, typename M>
__device__ inline void write_sym_tile_direct(V acc, int R, int C, int lim, S *out, size_t ld, int lane)
{
    const int rr = C + (lane & 15);
#pragma unroll
    for (int j = 0; j < 4; ++j) {
        const int cc = R + M::row(lane, j);
        if (rr < lim && cc < lim) out[(size_t)cc * ld + rr] = (S)acc[j];
    }
    if (R != C) {
#pragma unroll
        for (int j = 0; j < 4; ++j) {
            const int r2 = R + M::row(lane, j);
            if (r2 < lim && rr < lim) out[(size_t)rr * ld + r2] = (S)acc[j];
        }
    }
}

// FP32 accumulation with an error-free carry: every ~96 A columns the MFMA
// accumulator c is folded into a running FP32 sum h by TwoSum
// (h' = fl(h + c), c' = (h + c) - h' exactly), so c only ever accumulates a
// few chunks' products on top of a residual and the FP32 rounding of the long
// k-sum (k = 3Q or 6Q columns) no longer grows with Q.  The tile's value is
// h + c, used in FP64 by the condensation and rounded once for FP32 output.
// For FP64 accumulators the carry is absent (Hi is a placeholder).
template <typename T, typename A> struct Carry {
    typedef char hi_t;
    __device__ static inline void zero(hi_t &) {}
    __device__ static inline void fold(hi_t &, A &) {}
    __device__ static inline double val(const hi_t &, const A &c, int j) { return (double)c[j]; }
    __device__ static inline A sum(const hi_t &, const A &c) { return c; }
};
template <typename A> struct Carry<float, A> {
    typedef A hi_t;
    __device__ static inline void zero(hi_t &h) { h = A{0.f, 0.f, 0.f, 0.f}; }
    __device__ static inline void fold(hi_t &h, A &c)
    {
#pragma unroll
        for (int j = 0; j < 4; ++j) {
            float a = c[j], hh = h[j];
            float t = hh + a, bp = t - hh;
            c[j] = (hh - (t - bp)) + (a - bp);
            h[j] = t;
        }
    }
    __device__ static inline double val(const hi_t &h, const A &c, int j) { return (double)h[j] + (double)c[j]; }
    __device__ static inline A sum(const hi_t &h, const A &c) { return h + c; }
};

// bfloat16 operands for the BF16 trial: round to nearest even, 4 values
// packed per lane as v_mfma_f32_16x16x16_bf16 takes them (lane l supplies
// A[l % 16][4 (l / 16) + i] and B[4 (l / 16) + i][l % 16], i = 0..3; the
// result has the layout of the f32 16x16x4 instruction)
typedef short s4_t __attribute__((ext_vector_type(4)));
__device__ inline short f2bf(float f)
{
    unsigned u = __float_as_uint(f);
    u += 0x7FFFu + ((u >> 16) & 1u);
    return (short)(u >> 16);
}
__device__ inline s4_t bf4(const float *col, int ld, int off)
{
    return s4_t{f2bf(col[off]), f2bf(col[ld + off]), f2bf(col[2 * ld + off]), f2bf(col[3 * ld + off])};
}

// Geometry pass of the block-tiled path: every quadrature point of an
// element once (geom3D.F90:77-91, geom.F90:57-113, elem_opt.F90:101-122),
// into the element's slab of the workspace, which all of the element's K
// blocks then read; one workgroup per element.
template <int SPACE>
__global__ __launch_bounds__(WG) void geometry_kernel(ElemArgs a)
{
    constexpr int GS = geo_stride(SPACE);
    __shared__ ElemShared es;
    __shared__ double tv[3][NTAB * NPT], td[3][NTAB * NPT];
    __shared__ unsigned dgeo[MAXN_H1];
    __shared__ double xg[3 * XS_MAX];
    const int e = blockIdx.x, tid = threadIdx.x, wave = tid >> 6;
    if (!element_setup<SPACE>(a, e, es)) return; // the element kernel reports the orders
    if (wave < 3) fill_tables(es.nq[wave], tv[wave], td[wave]);
    for (int k = tid; k < a.ngeo; k += WG) dgeo[k] = desc_h1(k, es.o);
    const double *xn = a.xnod + 3 * (size_t)a.ngeo * e;
    const bool xs_ok = a.ngeo <= XS_MAX;
    if (xs_ok)
        for (int k = tid; k < 3 * a.ngeo; k += WG) xg[k] = xn[k];
    __syncthreads();
    double *gp = a.gws + (size_t)e * geo_slab(SPACE);
    const double coef = a.coef[e];
    bool neg = false;
    for (int l = tid; l < es.nint; l += WG)
        neg |= xs_ok ? point_geometry<SPACE>(es, l, tv, td, dgeo, xg, a.ngeo, coef, a.load, gp + l * GS)
                     : point_geometry<SPACE>(es, l, tv, td, dgeo, xn, a.ngeo, coef, a.load, gp + l * GS);
    if (neg) es.negjac = 1;
    __syncthreads();
    if (tid == 0) gp[(size_t)MAXQ * GS] = es.negjac ? 1.0 : 0.0;
}

// General element kernel: one workgroup of 8 waves per (element, 128 x 128
// block of K).  An off-diagonal block's 64 tiles of 16 x 16 are 8 per wave
// (wave w: tile row w); a diagonal block's 36 upper tiles are 5 or 4 per wave
// (waves p and p + 4 share tile rows p and 7 - p, 9 tiles).  With 8 tiles a
// wave fits in 128 VGPRs, so two workgroups give 4 waves per SIMD.
// The geometry comes from the element's slab (geometry_kernel).
// Each row of the block's A is built by 2 (off-diagonal) or 4 (diagonal)
// threads, each for its share of the chunk's points, with the row's
// descriptor read once; the A chunks are double buffered (FP32 24, FP64 12
// columns per chunk), so a chunk's build follows the MFMAs of the previous
// one with one barrier between chunks.
// BF (with T = float): the contraction's operands are rounded to bfloat16 and
// multiplied on v_mfma_f32_16x16x16_bf16 with FP32 accumulation (the BF16
// trial); the A chunk is padded from KC to 32 columns with zeros.
constexpr int WGE = 512; // threads per workgroup of the block-tiled kernel
template <typename T, int SPACE, bool BF = false>
__global__ __launch_bounds__(WGE, 4) void elem_kernel(ElemArgs a)
{
    typedef Mfma<T> M;
    typedef typename M::acc_t acc_t;
    constexpr int NC = SPACE == 0 ? 3 : 6;             // A columns per point
    constexpr int KCE = (BF || sizeof(T) == 4) ? KC : KC / 2; // A columns per chunk
    constexpr int NBUF = BF ? 1 : 2;                   // A buffers
    constexpr int QC = KCE / NC;                       // points per A chunk
    constexpr int GQ = 64;                             // points per geometry pass
    constexpr int GS = geo_stride(SPACE);              // geometry doubles per point
    constexpr int KCA = BF ? 32 : KCE;                 // allocated A columns
    constexpr int BUFE = KCA * LDA;                    // one A buffer, in T
    constexpr int NS = 8;                              // tile slots per wave
    static_assert(!BF || sizeof(T) == 4, "bfloat16 operands come from an FP32 chunk");
    static_assert(GQ % QC == 0 && QC % 2 == 0, "a geometry pass holds whole chunks");

    __shared__ ElemShared es;
    __shared__ double tv[3][NTAB * NPT], td[3][NTAB * NPT];
    __shared__ unsigned dsp[2 * BR];   // descriptors of the block's rows
    __shared__ unsigned pix[MAXQ];     // packed 1D point indices
    __shared__ double geo[GQ * GS];
    __shared__ double bv[3][BR];       // partial load sums of a diagonal block's rows
    __shared__ T ac[NBUF * BUFE];

    const int e = blockIdx.x;
    const int tid = threadIdx.x, lane = tid & 63, wave = tid >> 6;
    int I, J;
    block_ij(blockIdx.y, I, J);
    const bool diag = (I == J);

    if (!element_setup<SPACE>(a, e, es)) {
        if (tid == 0 && blockIdx.y == 0 && a.status) a.status[e] = HP3D_E_ORDER;
        return;
    }
    // the grid holds the block rows of the caller's max_ndof: an element
    // above it would be cut short, so it is an argument error
    if (es.n > a.max_ndof) {
        if (tid == 0 && blockIdx.y == 0 && a.status) a.status[e] = HP3D_E_ARG;
        return;
    }
    const int n = es.n, nint = es.nint;
    if (I * BR >= n || J * BR >= n) return; // block beyond this element's size

    if (wave < 3) fill_tables(es.nq[wave], tv[wave], td[wave]);
    {
        const ElemOrders ol = es.o; // in registers for the descriptor walks
        for (int r = tid; r < 2 * BR; r += WGE) {
            int g = (r < BR ? I * BR + r : J * BR + r - BR);
            dsp[r] = g < n ? (SPACE == 0 ? desc_h1(g, ol) : desc_hcurl(g, ol)) : 0u;
        }
    }
    for (int l = tid; l < nint; l += WGE) {
        Pt p = point_ix(l, es.nq);
        pix[l] = (unsigned)p.l1 | ((unsigned)p.l2 << 4) | ((unsigned)p.l3 << 8);
    }
    if (BF)
        for (int i = tid; i < (KCA - KC) * LDA; i += WGE) ac[KC * LDA + i] = (T)0;
    __syncthreads();

    const double *gsrc = a.gws + (size_t)e * geo_slab(SPACE);

    // tile slots of this wave: slots s < nA in tile row tiA (tile columns
    // tjA + s), the rest in tile row tiB (tile columns tjB + s - nA); all
    // wave-uniform (scalar registers), with the lane's row within a tile apart
    const int wv = __builtin_amdgcn_readfirstlane(wave);
    const int l15 = lane & 15;
    const int nrI = min(BR, n - I * BR), nrJ = min(BR, n - J * BR);
    const int offJ = diag ? 0 : BR;
    int tiA, tiB, tjA, tjB, nA, nslot;
    if (!diag) {
        tiA = tiB = wv;
        tjA = tjB = 0;
        nA = nslot = NS;
    } else if (wv < 4) {
        tiA = wv;
        tiB = 7 - wv;
        tjA = wv;
        tjB = 0;
        nA = nslot = 5;
    } else {
        const int p = wv - 4;
        tiA = p;
        tiB = 7 - p;
        tjA = p + 5;
        tjB = 7 - p;
        nA = 3 - p;
        nslot = 4;
    }
    int offS[NS]; // column offset of each slot's tile, wave-uniform
    unsigned valid = 0;
#pragma unroll
    for (int s = 0; s < NS; ++s) {
        const int ti = s < nA ? tiA : tiB, tj = s < nA ? tjA + s : tjB + (s - nA);
        offS[s] = offJ + tj * 16;
        const bool v = s < nslot && ti * 16 < nrI && tj * 16 < nrJ;
        valid |= (v ? 1u : 0u) << s;
    }
    const unsigned vmask = valid;
    typedef Carry<std::conditional_t<BF, double, T>, acc_t> CY; // no carry under bfloat16 rounding
    acc_t acc[NS];
    typename CY::hi_t hi[NS];
#pragma unroll
    for (int s = 0; s < NS; ++s) {
        acc[s] = acc_t{0, 0, 0, 0};
        CY::zero(hi[s]);
    }
    const bool do_load = diag && a.bout != nullptr;

    // this thread's row of A (row r of the block's row set: I rows, then J
    // rows off the diagonal) and its share sub of the chunk's points
    const int nrows = diag ? BR : 2 * BR;
    const int row = tid & (nrows - 1), sub = diag ? (wv >> 1) : (wv >> 2);
    const int grow = row < BR ? I * BR + row : J * BR + row - BR;
    const bool row_ok = grow < n;
    const unsigned drow = dsp[row];
    double bsum = 0.0;

    auto build_pts = [&](auto nshare, int g0, int l0, int gend, T *buf) {
        constexpr int P = decltype(nshare)::value, NP = (QC + P - 1) / P;
#pragma unroll
        for (int qq = 0; qq < NP; ++qq) {
            const int q = P * qq + sub, lq = l0 + q;
            if (QC % P != 0 && q >= QC) break;
            // branch-free: past the chunk or N, point 0 and the row's (valid)
            // descriptor give finite entries that are stored as zeros
            const bool ok = row_ok && lq < gend;
            const int lqc = lq < gend ? lq : 0;
            const unsigned pk = pix[g0 + lqc];
            const Pt p{(int)(pk & 15), (int)((pk >> 4) & 15), (int)(pk >> 8)};
            double out[NC];
            const double bl = a_entries_load<SPACE>(drow, p, tv, td, geo + lqc * GS, out);
            bsum += ok ? bl : 0.0;
#pragma unroll
            for (int c = 0; c < NC; ++c) buf[(NC * q + c) * LDA + row] = ok ? (T)out[c] : (T)0;
        }
    };
    auto build = [&](int g0, int l0, int gend, T *buf) {
        if (diag) build_pts(std::integral_constant<int, WGE / BR>{}, g0, l0, gend, buf);
        else build_pts(std::integral_constant<int, WGE / (2 * BR)>{}, g0, l0, gend, buf);
    };
    auto mfma = [&](const T *buf) {
        if constexpr (BF) {
#pragma unroll
            for (int k0 = 0; k0 < KCA; k0 += 16) {
                const float *col = (const float *)buf + (k0 + 4 * (lane >> 4)) * LDA + l15;
                s4_t aA = bf4(col, LDA, tiA * 16);
                s4_t aB = bf4(col, LDA, tiB * 16);
#pragma unroll
                for (int s = 0; s < NS; ++s)
                    if (vmask & (1u << s))
                        acc[s] = __builtin_amdgcn_mfma_f32_16x16x16bf16_1k(s < nA ? aA : aB, bf4(col, LDA, offS[s]),
                                                                          acc[s], 0, 0, 0);
            }
        } else {
#pragma unroll
            for (int k0 = 0; k0 < KCE; k0 += 4) {
                const T *col = buf + (k0 + (lane >> 4)) * LDA + l15;
                const T aA = col[tiA * 16], aB = col[tiB * 16];
#pragma unroll
                for (int s = 0; s < NS; ++s)
                    if (vmask & (1u << s)) acc[s] = M::mma(s < nA ? aA : aB, col[offS[s]], acc[s]);
            }
        }
    };

    int nfold = 0;
    for (int g0 = 0; g0 < nint; g0 += GQ) {
        const int gend = min(GQ, nint - g0);
        for (int i = tid; i < gend * GS; i += WGE) geo[i] = gsrc[(size_t)g0 * GS + i];
        __syncthreads();
        const int nch = (gend + QC - 1) / QC;
        build(g0, 0, gend, ac);
        __syncthreads();
        for (int c = 0; c < nch; ++c) {
            if constexpr (NBUF == 2) {
                // waves w and w + 4 share a SIMD: one issues its MFMAs while the
                // other builds the next chunk, then they swap
                T *cur = ac + (c & 1) * BUFE, *nxt = ac + ((c + 1) & 1) * BUFE;
                const bool first = wv >= 4 && c + 1 < nch, second = wv < 4 && c + 1 < nch;
                if (first) build(g0, (c + 1) * QC, gend, nxt);
                __asm__ volatile("" ::: "memory");
                mfma(cur);
                __asm__ volatile("" ::: "memory");
                if (second) build(g0, (c + 1) * QC, gend, nxt);
            } else {
                mfma(ac);
            }
            if ((++nfold & 3) == 0) { // FP32: every 4 chunks of 24 columns
#pragma unroll
                for (int s = 0; s < NS; ++s) CY::fold(hi[s], acc[s]);
            }
            __syncthreads();
            if constexpr (NBUF == 1) {
                if (c + 1 < nch) {
                    build(g0, (c + 1) * QC, gend, ac);
                    __syncthreads();
                }
            }
        }
    }

    // the load vector: the four partial sums of each diagonal row
    if (do_load && sub > 0) bv[sub - 1][row] = bsum;
    __syncthreads();
    if (do_load && sub == 0 && row_ok) a.bout[a.off_v[e] + grow] = bsum + bv[0][row] + bv[1][row] + bv[2][row];
    if (tid == 0 && blockIdx.y == 0 && a.status)
        a.status[e] = gsrc[(size_t)MAXQ * GS] != 0.0 ? HP3D_NEG_JACOBIAN : 0;

    T *kout = (T *)a.kout + a.off_k[e];
    const size_t ld = a.ld_k > 0 ? (size_t)a.ld_k : (size_t)n;
    T *stage = ac + wave * (16 * 17);
#pragma unroll
    for (int s = 0; s < NS; ++s) {
        if (!(valid & (1u << s))) continue;
        const int ti = s < nA ? tiA : tiB, tj = s < nA ? tjA + s : tjB + (s - nA);
        write_sym_tile<acc_t, T, M>(CY::sum(hi[s], acc[s]), I * BR + ti * 16, J * BR + tj * 16, n, kout, ld,
                                    stage, lane);
    }
}

// Single-block element kernel (N <= 128): K accumulates in registers over the
// whole quadrature; then, when the caller asks for it and nb <= NBF, the
// bubble dofs are condensed in the same launch (the algebra of stc_fwd_herm /
// stc_fwd_gen, stc.F90:338-509):
//   Gaussian elimination without pivoting of [Abb | I] (Abb SPD: Abb =
//   L_u D L_u^T, the pivots D > 0 are the SPD check) gives D and L_u^-1;
//   then on FP64 MFMA: S = D^-1/2 L_u^-1 [Abi | Bb];  Aii -= S^T S and
//   Bi -= S^T s_b on the register tiles (Aib Abb^-1 Abi = S^T S);
//   X = L_u^-T D^-1/2 S = Abb^-1 [Abi | Bb] = [ASchur | BSchur].
// The A chunks are double buffered: the MFMAs of chunk c are issued before
// the VALU/LDS work that builds chunk c+1, so the two overlap.
constexpr int NBF = 32;      // largest bubble block condensed in the fused path
struct CondArgs {
    const long long *off_c, *off_s;
    double *aii, *bi, *aschur, *bschur;
};

template <typename T> struct SingleCfg {
    static constexpr int KCS = sizeof(T) == 4 ? 48 : 24; // A columns per chunk
};

template <typename T, int SPACE, int R, bool BF = false>
__device__ __forceinline__ void single_element(const ElemArgs &a, const CondArgs &cd, const int e,
                                               const int tid)
{
    static_assert(!BF || sizeof(T) == 4, "bfloat16 operands accumulate in FP32");
    typedef Mfma<T> M;
    typedef typename M::acc_t acc_t;
    typedef Mfma<double> MD;
    constexpr int NC = SPACE == 0 ? 3 : 6;
    constexpr int KCS = SingleCfg<T>::KCS;
    constexpr int QC = KCS / NC;               // points per chunk
    // points per geometry pass (a multiple of QC): 128 for the 128-row H1
    // form, 64 where a third workgroup per CU has to fit in LDS
    constexpr int GQ = SPACE == 0 && R == BR ? 128 : 64;
    // 64-row H1 form: the Jacobian (wave 0) and the load f(x) (wave 2) of a
    // point are evaluated side by side; gp[9] = w det J, gp[10] = f
    constexpr bool SPLIT = SPACE == 0 && R == BR / 2;
    static_assert(!SPLIT || 2 * GQ <= WG, "load threads beside the geometry threads");
    constexpr int GS = SPACE == 0 ? (SPLIT ? 11 : 10) : 21;
    static_assert(R == BR || R == BR / 2, "128 or 64 rows");
    constexpr int LD1 = R + 16;                // LDS stride of an A column (in T)
    constexpr int LDK = LD1 + 1;               // LDS stride of a bubble row (doubles; odd, so
                                               // that a column read across rows is conflict-free)
    constexpr int NBX = R == BR ? NBF : NBF / 2; // largest bubble block condensed here
    constexpr int LW = NBX + 1;                // leading dimension of L_u^-1
    constexpr int NSLOT = R == BR ? 9 : 3;     // tile slots per wave
    constexpr int UMAX = R == BR ? 5 : 2;      // right-hand-side tiles per wave in the condensation
    // BF16 trial: the chunk in bfloat16, row-major (a row's KCS values
    // contiguous, LDB shorts apart), so an MFMA operand (4 consecutive k of
    // one row) is one 8-byte LDS read
    constexpr int LDB = KCS + 4;
    static_assert(!BF || (KCS % 16 == 0 && LDB % 4 == 0), "bf16 chunk layout");
    constexpr int BUF = BF ? R * LDB / 2 : KCS * LD1; // one A buffer, in T
    constexpr int AC_D = (2 * BUF * (int)sizeof(T) + 7) / 8;
    constexpr int CW_D = NBX * LDK + NBX * LW + NBX + LD1 + 4 * 16 * 17; // output / condensation workspace
    constexpr int SM_D = AC_D > CW_D ? AC_D : CW_D;
    static_assert(GQ % QC == 0 && AC_D >= QC * R, "geometry passes hold whole chunks");

    __shared__ ElemShared es;
    __shared__ double tv[3][NTAB * NPT], td[3][NTAB * NPT];
    __shared__ unsigned dsp[R];
    __shared__ unsigned dgeo[MAXN_H1];
    __shared__ unsigned pix[MAXQ];
    __shared__ double geo[GQ * GS];
    __shared__ double smem[SM_D];
    __shared__ double xg[3 * XS_MAX];
    __shared__ int s_fail;
    T *ac = reinterpret_cast<T *>(smem);

    const int lane = tid & 63, wave = tid >> 6;
    // the element's geometry nodes (up to 27-node maps, one value per thread)
    // and its coefficient are loaded before the header is decoded, so that
    // their latency overlaps the header's
    const double *xn = a.xnod + 3 * (size_t)a.ngeo * e;
    const bool xs_ok = a.ngeo <= XS_MAX;
    static_assert(3 * XS_MAX <= WG, "one staged node coordinate per thread");
    const double xv = xs_ok && tid < 3 * a.ngeo ? xn[tid] : 0.0;
    const double coef = a.coef[e];
    if (!element_setup<SPACE>(a, e, es, tid)) {
        if (tid == 0 && a.status) a.status[e] = HP3D_E_ORDER;
        return;
    }
    // an element above the caller's max_ndof (which chose R) is an argument
    // error: nothing of it is computed
    if (es.n > a.max_ndof || es.n > R) {
        if (tid == 0 && a.status) a.status[e] = HP3D_E_ARG;
        return;
    }
    const int n = es.n, nint = es.nint;
    // wave 0 fills the three 1D tables while waves 1-3 build the descriptors
    if (wave == 0) {
        fill_tables3(es.nq, lane, tv, td);
    } else {
        // the orders in registers: the descriptor walks over edges and faces
        // then run without a chain of LDS round trips
        const ElemOrders ol = es.o;
        for (int r = tid - 64; r < R; r += WG - 64)
            dsp[r] = r < n ? (SPACE == 0 ? desc_h1_scan(r, ol) : desc_hcurl(r, ol)) : 0u;
        for (int k = tid - 64; k < a.ngeo; k += WG - 64) dgeo[k] = desc_h1_scan(k, ol);
    }
    for (int l = tid; l < nint; l += WG) {
        Pt p = point_ix(l, es.nq);
        pix[l] = (unsigned)p.l1 | ((unsigned)p.l2 << 4) | ((unsigned)p.l3 << 8);
    }
    // geometry nodes of up to 27-node maps staged in LDS
    if (xs_ok && tid < 3 * a.ngeo) xg[tid] = xv;
    if (tid == 0) s_fail = 0;
    __syncthreads();

    // this thread's row of K's outputs (threads with sub == 0 hold one row each)
    const int row = tid & (R - 1), sub = tid / R;
    const bool row_ok = row < n;
    // tile slots: s < nA in tile row tiA (tile columns tjA + s), the rest in
    // tile row tiB (tile columns tjB + s - nA).  128 rows: wave w holds tile
    // rows w and 7 - w (9 tiles).  64 rows: waves p and p + 2 share tile rows
    // p and 3 - p (5 tiles: 3 + 2).
    int tiA, tjA, nA, tiB, tjB, nslot;
    if constexpr (R == BR) {
        tiA = tjA = wave;
        nA = 8 - wave;
        tiB = tjB = 7 - wave;
        nslot = 9;
    } else {
        const int p = wave & 1;
        tiA = p;
        tiB = tjB = 3 - p;
        if (wave < 2) {
            tjA = p;
            nA = nslot = 3;
        } else {
            tjA = p + 3;
            nA = 1 - p;
            nslot = 2;
        }
    }
    auto slot_tile = [&](int s, int &ti, int &tj) {
        ti = s < nA ? tiA : tiB;
        tj = s < nA ? tjA + s : tjB + (s - nA);
    };
    const int offA = tiA * 16 + (lane & 15), offB = tiB * 16 + (lane & 15);
    int offS[NSLOT];
#pragma unroll
    for (int s = 0; s < NSLOT; ++s) {
        int ti, tj;
        slot_tile(s, ti, tj);
        offS[s] = tj * 16 + (lane & 15);
    }
    typedef Carry<std::conditional_t<BF, double, T>, acc_t> CY; // no carry under bfloat16 rounding
    acc_t acc[NSLOT];
    typename CY::hi_t hi[NSLOT];
#pragma unroll
    for (int s = 0; s < NSLOT; ++s) {
        acc[s] = acc_t{0, 0, 0, 0};
        CY::zero(hi[s]);
    }
    unsigned valid = 0;
#pragma unroll
    for (int s = 0; s < NSLOT; ++s) {
        int ti, tj;
        slot_tile(s, ti, tj);
        valid |= ((s < nslot && ti * 16 < n && tj * 16 < n) ? 1u : 0u) << s;
    }
    // A chunk build: thread t takes point qt of the chunk and rows
    // rt + TPP * i (i < RPT), so the point's geometry is read once for all of
    // them; the load vector's partial sums per row meet after the main loop
    constexpr int TPP = WG / QC, RPT = R / TPP;
    static_assert(WG % QC == 0 && R % TPP == 0, "build split");
    const int qt = tid / TPP, rt = tid % TPP;
    double bacc[RPT];
#pragma unroll
    for (int i = 0; i < RPT; ++i) bacc[i] = 0.0;

    // Branch-free: beyond the chunk (lq >= gend) or N the entries come from
    // point 0 and descriptor 0 (valid table indexes) and are stored as zeros,
    // so no load sits under a predicate
    auto build = [&](int g0, int l0, int gend, T *buf) {
        const int lq = l0 + qt;
        const bool pok = lq < gend;
        const int lqc = pok ? lq : 0;
        const unsigned pk = pix[g0 + lqc];
        const Pt p{(int)(pk & 15), (int)((pk >> 4) & 15), (int)(pk >> 8)};
        double g[GS];
#pragma unroll
        for (int k = 0; k < GS; ++k) g[k] = geo[lqc * GS + k];
        if (SPLIT) g[9] = g[10] * g[9]; // f (w det J), as the unsplit pass forms it
#pragma unroll 4
        for (int i = 0; i < RPT; ++i) {
            const int r = rt + TPP * i;
            double out[NC];
            const double bl = a_entries_load<SPACE, true>(dsp[r], p, tv, td, g, out);
            const bool ok = pok && r < n;
            bacc[i] += ok ? bl : 0.0;
#pragma unroll
            for (int c = 0; c < NC; ++c) {
                if constexpr (BF)
                    reinterpret_cast<short *>(buf)[r * LDB + NC * qt + c] = ok ? f2bf((float)out[c]) : (short)0;
                else
                    buf[(NC * qt + c) * LD1 + r] = ok ? (T)out[c] : (T)0;
            }
        }
    };
    // The slot mask is wave-uniform.  When some of this wave's tiles lie beyond
    // N (N <= 112), a scalar branch skips their MFMAs; when all are live the
    // unguarded loop keeps the MFMAs back to back.
    const unsigned vmask = __builtin_amdgcn_readfirstlane(valid);
    const bool all_live = vmask == (1u << NSLOT) - 1;
    // FP32 H1: the loop specialised per wave, so each slot's operand rows are
    // compile-time offsets (no per-MFMA select or branch); the FP64 and
    // H(curl) forms would spill with it.  It runs every slot of the wave also
    // when some tiles lie beyond N (N <= 112, or 48 in the 64-row form): the
    // build stores zeros in the rows from N to R, such tiles come out zero,
    // and no output or condensation step reads a slot outside `valid`
    constexpr bool SPEC = sizeof(T) == 4 && SPACE == 0 && !BF;
    // wave W's slots (as set above): rows tiA / tiB, first columns tjA / tjB,
    // nA slots in row tiA, ns slots in all
    auto mfma_w = [&](auto wc, const T *buf) {
        constexpr int W = decltype(wc)::value, P = W & 1;
        constexpr int TA = R == BR ? W : P, TB = R == BR ? 7 - W : 3 - P;
        constexpr int JA = R == BR ? W : (W < 2 ? P : P + 3), JB = TB;
        constexpr int NA = R == BR ? 8 - W : (W < 2 ? 3 : 1 - P), NS = R == BR ? 9 : (W < 2 ? 3 : 2);
#pragma unroll
        for (int k0 = 0; k0 < KCS; k0 += 4) {
            const T *col = buf + (k0 + (lane >> 4)) * LD1 + (lane & 15);
            const T aA = col[TA * 16], aB = col[TB * 16];
#pragma unroll
            for (int s = 0; s < NS; ++s)
                acc[s] = M::mma(s < NA ? aA : aB, col[(s < NA ? JA + s : JB + (s - NA)) * 16], acc[s]);
        }
    };
    const int wv = __builtin_amdgcn_readfirstlane(wave);
    // The FP32 H1 chunk step with the next chunk's build woven in: after each
    // k step's MFMAs, one row of the next chunk (rows during the first RPT
    // k steps), so that the build's VALU and LDS work can issue while the
    // MFMAs run.  The 128-row form only: the 64-row form's chunk has a third
    // of the MFMAs to hide the build under, and measured slower with it.
    auto mfma_build_w = [&](auto wc, const T *buf, int g0, int l0, int gend, T *nbuf) {
        constexpr int W = decltype(wc)::value, P = W & 1;
        constexpr int TA = R == BR ? W : P, TB = R == BR ? 7 - W : 3 - P;
        constexpr int JA = R == BR ? W : (W < 2 ? P : P + 3), JB = TB;
        constexpr int NA = R == BR ? 8 - W : (W < 2 ? 3 : 1 - P), NS = R == BR ? 9 : (W < 2 ? 3 : 2);
        constexpr int NKS = KCS / 4;
        static_assert(RPT <= NKS, "a row per k step");
        const int lq = l0 + qt;
        const bool pok = lq < gend;
        const int lqc = pok ? lq : 0;
        const unsigned pk = pix[g0 + lqc];
        const Pt p{(int)(pk & 15), (int)((pk >> 4) & 15), (int)(pk >> 8)};
        double g[GS];
#pragma unroll
        for (int k = 0; k < GS; ++k) g[k] = geo[lqc * GS + k];
        if (SPLIT) g[9] = g[10] * g[9];
#pragma unroll
        for (int ks = 0; ks < NKS; ++ks) {
            const T *col = buf + (4 * ks + (lane >> 4)) * LD1 + (lane & 15);
            const T aA = col[TA * 16], aB = col[TB * 16];
#pragma unroll
            for (int s = 0; s < NS; ++s)
                acc[s] = M::mma(s < NA ? aA : aB, col[(s < NA ? JA + s : JB + (s - NA)) * 16], acc[s]);
            if (ks < RPT) {
                const int r = rt + TPP * ks;
                double out[NC];
                const double bl = a_entries_load<SPACE, true>(dsp[r], p, tv, td, g, out);
                const bool ok = pok && r < n;
                bacc[ks] += ok ? bl : 0.0;
#pragma unroll
                for (int c = 0; c < NC; ++c) nbuf[(NC * qt + c) * LD1 + r] = ok ? (T)out[c] : (T)0;
            }
        }
    };
    auto mfma = [&](const T *buf) {
        if constexpr (BF) {
            // lane l supplies A[l % 16][4 (l / 16) + i] and B[4 (l / 16) + i][l % 16]
            // (i < 4) of v_mfma_f32_16x16x16_bf16: row l % 16 of a tile, four
            // consecutive columns of the chunk
            const short *sb = reinterpret_cast<const short *>(buf) + (lane & 15) * LDB + 4 * (lane >> 4);
#pragma unroll
            for (int k0 = 0; k0 < KCS; k0 += 16) {
                const s4_t aA = *reinterpret_cast<const s4_t *>(sb + tiA * 16 * LDB + k0);
                const s4_t aB = *reinterpret_cast<const s4_t *>(sb + tiB * 16 * LDB + k0);
#pragma unroll
                for (int s = 0; s < NSLOT; ++s) {
                    int ti, tj;
                    slot_tile(s, ti, tj);
                    if (vmask & (1u << s))
                        acc[s] = __builtin_amdgcn_mfma_f32_16x16x16bf16_1k(
                            s < nA ? aA : aB, *reinterpret_cast<const s4_t *>(sb + tj * 16 * LDB + k0), acc[s], 0, 0,
                            0);
                }
            }
            return;
        }
        if (SPEC) {
            if constexpr (SPEC) {
                switch (wv) {
                case 0: mfma_w(std::integral_constant<int, 0>{}, buf); break;
                case 1: mfma_w(std::integral_constant<int, 1>{}, buf); break;
                case 2: mfma_w(std::integral_constant<int, 2>{}, buf); break;
                default: mfma_w(std::integral_constant<int, 3>{}, buf); break;
                }
            }
        } else if (all_live) {
#pragma unroll
            for (int k0 = 0; k0 < KCS; k0 += 4) {
                const T *col = buf + (k0 + (lane >> 4)) * LD1;
                const T aA = col[offA], aB = col[offB];
#pragma unroll
                for (int s = 0; s < NSLOT; ++s) acc[s] = M::mma(s < nA ? aA : aB, col[offS[s]], acc[s]);
            }
        } else {
#pragma unroll
            for (int k0 = 0; k0 < KCS; k0 += 4) {
                const T *col = buf + (k0 + (lane >> 4)) * LD1;
                const T aA = col[offA], aB = col[offB];
#pragma unroll
                for (int s = 0; s < NSLOT; ++s)
                    if (vmask & (1u << s)) acc[s] = M::mma(s < nA ? aA : aB, col[offS[s]], acc[s]);
            }
        }
    };

    // fold the FP32 accumulators every 96 columns (2 chunks)
    constexpr int FOLD_EVERY = KCS >= 48 ? 2 : 4;
    int nfold = 0;
    for (int g0 = 0; g0 < nint; g0 += GQ) {
        const int gend = min(GQ, nint - g0);
        if (tid < gend) {
            const bool neg = xs_ok ? point_geometry<SPACE, SPLIT>(es, g0 + tid, tv, td, dgeo, xg, a.ngeo, coef,
                                                                  a.load, geo + tid * GS)
                                   : point_geometry<SPACE, SPLIT>(es, g0 + tid, tv, td, dgeo, xn, a.ngeo, coef,
                                                                  a.load, geo + tid * GS);
            if (neg) es.negjac = 1;
        } else if (SPLIT && tid >= WG / 2 && tid - WG / 2 < gend) {
            const int l = tid - WG / 2;
            geo[l * GS + 10] = xs_ok ? point_load_h1(es, g0 + l, tv, dgeo, xg, a.ngeo, a.load)
                                     : point_load_h1(es, g0 + l, tv, dgeo, xn, a.ngeo, a.load);
        }
        __syncthreads();
        const int nch = (gend + QC - 1) / QC;
        build(g0, 0, gend, ac);
        __syncthreads();
        for (int c = 0; c < nch; ++c) {
            T *const cur = ac + (c & 1) * BUF, *const nxt = ac + ((c + 1) & 1) * BUF;
            if (SPEC && R == BR && c + 1 < nch) {
                if constexpr (SPEC && R == BR) {
                    switch (wv) {
                    case 0: mfma_build_w(std::integral_constant<int, 0>{}, cur, g0, (c + 1) * QC, gend, nxt); break;
                    case 1: mfma_build_w(std::integral_constant<int, 1>{}, cur, g0, (c + 1) * QC, gend, nxt); break;
                    case 2: mfma_build_w(std::integral_constant<int, 2>{}, cur, g0, (c + 1) * QC, gend, nxt); break;
                    default: mfma_build_w(std::integral_constant<int, 3>{}, cur, g0, (c + 1) * QC, gend, nxt); break;
                    }
                }
            } else {
                mfma(cur);
                if (c + 1 < nch) build(g0, (c + 1) * QC, gend, nxt);
            }
            if ((++nfold & (FOLD_EVERY - 1)) == 0) {
#pragma unroll
                for (int s = 0; s < NSLOT; ++s) CY::fold(hi[s], acc[s]);
            }
            __syncthreads();
        }
    }

    // ---- outputs: load vector, K, and the condensed system ----
    // (the main loop ended on a barrier: the A buffers are free; from here on
    // the threads share data through LDS only, and the barriers do not wait
    // for the K, Aii and ASchur stores)
    double *kb = smem;              // bubble rows b < nb: Abi cols 0..ni-1, Abb cols ni..n-1, Bb col n
    double *wm = kb + NBX * LDK;    // L_u^-1, leading dimension LW
    double *rs = wm + NBX * LW;     // D(j)^-1/2
    double *stg = rs + NBX + LD1;   // 4 x 16 x 17 wave stages
    T *kstage = reinterpret_cast<T *>(stg + wave * (16 * 17));
    // the FP32 128-row form hands K tiles over through geo[0, NSLOT * 128)
    // (as doubles' room); the load vector sits in geo after them
    constexpr bool KHAND = sizeof(T) == 4 && R == BR;
    constexpr int BV_OFF = KHAND ? NSLOT * 128 : 0;
    static_assert(GQ * GS >= BV_OFF + R, "load vector beside the K hand-over");
    double *bv = geo + BV_OFF;      // load vector
    // the load vector: the QC partial sums of each row, in the L_u^-1, D and
    // stage space (not read or written before the next barrier but one), so
    // that the bubble rows can go into kb right after the sums are read
    double bsum = 0.0;
    {
        static_assert(QC * R <= NBX * LW + NBX + LD1 + 4 * 16 * 17, "partial sums in the stage space");
        double *bq = wm; // QC x R
#pragma unroll
        for (int i = 0; i < RPT; ++i) bq[qt * R + rt + TPP * i] = bacc[i];
        lds_barrier();
        if (sub == 0)
#pragma unroll
            for (int q = 0; q < QC; ++q) bsum += bq[q * R + row];
    }
    const int nb = es.nb, ni = n - nb;
    const bool cond = cd.aii != nullptr && nb <= NBX;
    if (sub == 0 && row_ok) {
        if (a.bout) a.bout[a.off_v[e] + row] = bsum;
        bv[row] = bsum;
        if (cond && row >= ni) kb[(row - ni) * LDK + n] = bsum;
    }
    auto write_k = [&]() {
        T *kout = (T *)a.kout + a.off_k[e];
        const size_t ld = a.ld_k > 0 ? (size_t)a.ld_k : (size_t)n;
#pragma unroll
        for (int s = 0; s < NSLOT; ++s) {
            if (!(valid & (1u << s))) continue;
            int ti, tj;
            slot_tile(s, ti, tj);
            write_sym_tile<acc_t, T, M>(CY::sum(hi[s], acc[s]), ti * 16, tj * 16, n, kout, ld, kstage, lane);
        }
    };
    if (!cond) {
        lds_barrier(); // the partial sums are read before the stages are written
        if (a.kout) write_k();
        // an element beyond the bubble limit the caller declared is an error
        if (tid == 0 && a.status)
            a.status[e] = cd.aii != nullptr ? HP3D_E_ARG : (es.negjac ? HP3D_NEG_JACOBIAN : 0);
        return;
    }
    // bubble rows of K from the register tiles
#pragma unroll
    for (int s = 0; s < NSLOT; ++s) {
        if (!(valid & (1u << s))) continue;
        int ti, tj;
        slot_tile(s, ti, tj);
#pragma unroll
        for (int j = 0; j < 4; ++j) {
            int r = ti * 16 + M::row(lane, j), c = tj * 16 + (lane & 15);
            double v = CY::val(hi[s], acc[s], j);
            if (r < n && c < n) {
                if (c >= ni) kb[(c - ni) * LDK + r] = v;
                if (r >= ni) kb[(r - ni) * LDK + c] = v;
            }
        }
    }
    // Gaussian elimination of [Abb | I] in the registers of wave EW (L_u^-1
    // and D^-1/2), while the other waves store their tiles of K.  In the
    // FP32 128-row form EW first hands its own K tiles (the carry folded to
    // FP32) to the other waves through the geometry buffer, free after the
    // main loop: its stores would otherwise follow its elimination on the
    // critical path.  Elsewhere EW stores its own afterwards.
    constexpr int EW = 3;
    T *khand = reinterpret_cast<T *>(geo);
    static_assert(!KHAND || GQ * GS * 2 >= NSLOT * 256, "K hand-over buffer");
    if constexpr (KHAND) {
        if (a.kout && wv == EW) {
#pragma unroll
            for (int s = 0; s < NSLOT; ++s) {
                if (!(valid & (1u << s))) continue;
                const acc_t v = CY::sum(hi[s], acc[s]);
#pragma unroll
                for (int j = 0; j < 4; ++j) khand[s * 256 + 4 * lane + j] = v[j];
            }
        }
    }
    lds_barrier();
    double *li = wm; // L_u^-1, leading dimension LW (unit lower; zero above)
    if (wv == EW && eliminate_wave<NBX>(kb, LDK, ni, nb, li, LW, rs, lane) && lane == 0) s_fail = 1;
    if (a.kout) {
        if (!(KHAND && wv == EW)) write_k();
        if constexpr (KHAND) {
            if (wv != EW) {
                // EW's slots s = wv, wv + 3, wv + 6: tile row 3 (columns 3..7),
                // then tile row 4 (columns 4..7)
                T *kout = (T *)a.kout + a.off_k[e];
                const size_t ld = a.ld_k > 0 ? (size_t)a.ld_k : (size_t)n;
                for (int s = wv; s < NSLOT; s += 3) {
                    const int ti = s < 5 ? 3 : 4, tj = s < 5 ? 3 + s : s - 1;
                    if (ti * 16 >= n || tj * 16 >= n) continue;
                    acc_t v;
#pragma unroll
                    for (int j = 0; j < 4; ++j) v[j] = khand[s * 256 + 4 * lane + j];
                    write_sym_tile<acc_t, T, M>(v, ti * 16, tj * 16, n, kout, ld, kstage, lane);
                }
            }
        }
    }
    lds_barrier();
    // GEMMs against the right-hand block [Abi | Bb] on FP64 MFMA.  Output
    // tiles: rows 16*tr.. (bubble index), columns 16*tc.. (ni Abi columns,
    // then Bb).  Z = L_u^-1 R, stored as S = D^-1/2 Z; X = L_u^-T (D^-1/2 S).
    const int ntr = (nb + 15) / 16, nct = (ni + 16) / 16, ntile = ntr * nct;
    auto rhs_col = [&](int c) { return c < ni ? c : (c == ni ? n : -1); };
    // This wave's tiles t = wave + 4u, u < nu (wave-uniform), as independent
    // MFMA chains interleaved over k: one straight-line body per tile count,
    // loads from clamped indexes masked at use (no predicated loads)
    auto gemm_li = [&](bool second, d4_t *out) {
        const int nu = wv < ntile ? (ntile - 1 - wv) / 4 + 1 : 0;
        auto run = [&](auto nuc) {
            constexpr int NU = decltype(nuc)::value;
            int arc[NU], pcc[NU];
            bool aok[NU], bok[NU];
#pragma unroll
            for (int u = 0; u < NU; ++u) {
                const int t = wv + 4 * u, tr = t / nct, tc = t % nct;
                const int ar = tr * 16 + (lane & 15), pc = rhs_col(tc * 16 + (lane & 15));
                aok[u] = ar < nb;
                bok[u] = pc >= 0;
                arc[u] = aok[u] ? ar : 0;
                pcc[u] = bok[u] ? pc : 0;
                out[u] = d4_t{0, 0, 0, 0};
            }
#pragma unroll
            for (int k0 = 0; k0 < NBX; k0 += 4) {
                const int k = k0 + (lane >> 4);
                const bool kok = k < nb;
                const int kc = kok ? k : 0;
                const double rk = second ? rs[kc] : 1.0;
#pragma unroll
                for (int u = 0; u < NU; ++u) {
                    double av = second ? li[kc * LW + arc[u]] : li[arc[u] * LW + kc];
                    double bw = kb[kc * LDK + pcc[u]] * rk;
                    av = aok[u] && kok ? av : 0.0;
                    bw = bok[u] && kok ? bw : 0.0;
                    out[u] = MD::mma(av, bw, out[u]);
                }
            }
        };
        switch (nu) {
        case 1: run(std::integral_constant<int, 1>{}); break;
        case 2: run(std::integral_constant<int, 2>{}); break;
        case 3: if constexpr (UMAX >= 3) run(std::integral_constant<int, 3>{}); break;
        case 4: if constexpr (UMAX >= 4) run(std::integral_constant<int, 4>{}); break;
        case 5: if constexpr (UMAX >= 5) run(std::integral_constant<int, 5>{}); break;
        default: break;
        }
    };
    auto store_rhs = [&](bool first, const d4_t *in) {
#pragma unroll
        for (int u = 0; u < UMAX; ++u) {
            int t = wave + 4 * u;
            if (t >= ntile) continue;
            int tr = t / nct, tc = t % nct;
            int pc = rhs_col(tc * 16 + (lane & 15));
#pragma unroll
            for (int j = 0; j < 4; ++j) {
                int b = tr * 16 + MD::row(lane, j);
                if (b < nb && pc >= 0) kb[b * LDK + pc] = first ? in[u][j] * rs[b] : in[u][j];
            }
        }
    };
    d4_t yt[UMAX];
    if (nb > 0) gemm_li(false, yt);
    lds_barrier();
    if (nb > 0) store_rhs(true, yt);
    lds_barrier();
    // S = D^-1/2 Z now sits in the bubble rows.  Bi = b_i - Z^T D^-1 z_b = b_i - S^T s_b
    double *aii = cd.aii + cd.off_c[e];
    double *bi = cd.bi + a.off_v[e];
    if (tid < ni) {
        double sacc = bv[tid];
        for (int k = 0; k < nb; ++k) sacc -= kb[k * LDK + tid] * kb[k * LDK + n];
        bi[tid] = sacc;
    }
    // Aii = Kii - Z^T D^-1 Z = Kii - S^T S on the register tiles, in FP64
#pragma unroll
    for (int s = 0; s < NSLOT; ++s) {
        if (!(valid & (1u << s))) continue;
        int ti, tj;
        slot_tile(s, ti, tj);
        const int R0 = ti * 16, C0 = tj * 16;
        if (R0 >= ni || C0 >= ni) continue;
        d4_t da; // the tile in the accumulator's layout, promoted to FP64
#pragma unroll
        for (int j = 0; j < 4; ++j) da[j] = CY::val(hi[s], acc[s], j);
        d4_t up = d4_t{0, 0, 0, 0};
        // The FP64 MFMA returns row i of its A operand's rows in register
        // i / 4 of lane group i % 4, the FP32 accumulator holds row
        // 4 (l / 16) + j in register j of lane l.  For FP32 tiles, feeding
        // operand row i = a + 4b (a, b < 4) with tile row 4a + b makes up[j]
        // of lane l the entry (R0 + 4 (l / 16) + j, C0 + l % 16), da's own
        // layout; FP64 tiles already share the MFMA's.
        const int li16 = lane & 15;
        const int ra = R0 + (sizeof(T) == 4 ? 4 * (li16 & 3) + (li16 >> 2) : li16), cb = C0 + li16;
        const bool rok = ra < ni, cok = cb < ni;
        const int rac = rok ? ra : 0, cbc = cok ? cb : 0;
#pragma unroll
        for (int k0 = 0; k0 < NBX; k0 += 4) {
            const int k = k0 + (lane >> 4);
            const bool kok = k < nb;
            const int kc = kok ? k : 0;
            double av = kb[kc * LDK + rac], bw = kb[kc * LDK + cbc];
            av = kok && rok ? av : 0.0;
            bw = kok && cok ? bw : 0.0;
            up = MD::mma(av, bw, up);
        }
#pragma unroll
        for (int j = 0; j < 4; ++j) da[j] -= up[j];
        write_sym_tile_direct<d4_t, double, M>(da, R0, C0, ni, aii, (size_t)ni, lane);
    }
    // X = L_u^-T D^-1/2 S = Abb^-1 [Abi | Bb], stored from the MFMA tiles
    // straight into ASchur (nb x ni, leading dimension nb) and BSchur: the
    // Aii step above and this one only read S, so neither waits for the other
    if (nb > 0) {
        gemm_li(true, yt);
        double *xs = cd.aschur + cd.off_s[e];
        double *xb = cd.bschur + a.off_v[e];
#pragma unroll
        for (int u = 0; u < UMAX; ++u) {
            const int t = wave + 4 * u;
            if (t >= ntile) continue;
            const int tr = t / nct, tc = t % nct;
            const int pc = rhs_col(tc * 16 + (lane & 15));
#pragma unroll
            for (int j = 0; j < 4; ++j) {
                const int b = tr * 16 + MD::row(lane, j);
                if (b < nb && pc >= 0) {
                    if (pc < ni) xs[(size_t)pc * nb + b] = yt[u][j];
                    else xb[b] = yt[u][j];
                }
            }
        }
    }
    if (tid == 0 && a.status) a.status[e] = s_fail ? HP3D_E_FACTOR : (es.negjac ? HP3D_NEG_JACOBIAN : 0);
}

// one workgroup per element; R = 64 for elements of at most 64 dofs and 16
// bubbles: half the LDS and a third of the tile registers, so three
// workgroups share a CU
template <typename T, int SPACE, int R, bool BF = false>
__global__ __launch_bounds__(WG, R == BR ? 2 : 3) void elem_single_kernel(ElemArgs a, CondArgs cd)
{
    single_element<T, SPACE, R, BF>(a, cd, blockIdx.x, threadIdx.x);
}

// arguments of the persistent form, as one kernel argument at offset 0
struct SingleArgs {
    ElemArgs a;
    CondArgs cd;
};
typedef const __attribute__((address_space(4))) SingleArgs *kargs_t;

// persistent form: a fixed grid (2 workgroups per CU) walks the elements with
// a grid stride; the barrier after each element keeps the next element's
// setup from overwriting LDS a slower wave still reads.  Per element the
// arguments are read again from the kernel-argument segment (scalar loads)
// and the thread index passes through an empty asm, so the compiler keeps
// nothing derived from them live from one element to the next: the element
// body needs every register.
template <typename T, int SPACE, int R>
__global__ __launch_bounds__(WG, R == BR ? 2 : 3) void elem_single_persistent(SingleArgs sa)
{
#if defined(__HIP_DEVICE_COMPILE__)
    const int nel = sa.a.nel, stride = gridDim.x;
    for (int e = blockIdx.x; e < nel; e += stride) {
        kargs_t kp = (kargs_t)__builtin_amdgcn_kernarg_segment_ptr();
        __asm__ volatile("" : "+s"(kp));
        const ElemArgs al = kp->a;
        const CondArgs cl = kp->cd;
        int tid = threadIdx.x;
        __asm__ volatile("" : "+v"(tid));
        single_element<T, SPACE, R>(al, cl, e, tid);
        __syncthreads();
    }
#endif
}

// ------------------------------------------------------------ condensation

__host__ __device__ inline int rnd16(int x) { return (x + 15) & ~15; }

struct StcArgs {
    int space, nel, ld_k, ld_c, nbp_max, n_max, nb_max; // nbp_max sizes the slabs; n_max, nb_max: the caller's bounds
    const int *norder; // order codes, or NULL when nsz gives (n, nb) per element
    const int *nsz;
    const long long *off_k, *off_v, *off_c, *off_s;
    const void *kin;
    const double *bin;
    double *aii, *bi, *aschur, *bschur, *ws;
    int *status;
};

// ------------------------------------------------- strip condensation
//
// Static condensation of one element per workgroup of 8 waves, for elements
// whose K is in HBM (the two-pass bins and the drop-in); all arithmetic FP64,
// K read as FP32 or FP64.  Replaces stc_fwd_wrapper + stc_fwd_herm
// (stc.F90:182-416), with its Schur update Aii = Aii - Aib (Abb^-1 Abi)
// (stc.F90:404-416):
//   1. Cholesky Abb = L L^T, left-looking by 16-column block panels.  Block
//      row i of a panel is owned by wave i mod 8, which gathers
//      A(i,p) - sum_k L(i,k) L(p,k)^T on MFMA; the owner of the diagonal
//      block factors it in LDS and inverts it; the panel below is
//      L(i,p) = A~(i,p) L_pp^-T on MFMA.  L goes to the workspace as 16 x 16
//      column-major tiles (lower triangle), with L_pp^-1 and L_pp^-T.
//   2. Z = [Abi | Bb] in strips of CW columns held in registers (block row i
//      of a strip by wave i mod 8).  Forward S = L^-1 Z and backward
//      X = L^-T S sweeps, right-looking, one block row per step: its owner
//      applies the inverted diagonal block and broadcasts the solved rows
//      through LDS, every wave updates its own rows with the streamed L
//      tiles.  While the backward sweep broadcasts X's rows, the waves
//      accumulate Aib X on the strip's columns of Aii (upper triangle and the
//      strip's diagonal block) and, on the load column, Aib BSchur.
//   3. The strip's X goes to ASchur / BSchur, Aii = Kii - Aib X (both
//      triangles) and Bi = Bi - Aib BSchur.
// Each L tile is read once per sweep per strip; nothing else is re-read.

// workspace tiles of one element with up to nbk 16-row bubble blocks: the
// lower triangle of L (after the factorisation Mb(i,p) = L_ii^-1 L(i,p) in
// its place), M(i,p) = L(i,p) L_pp^-1, then L_pp^-1 and L_pp^-T
__host__ __device__ inline int stc_tl(int i, int k) { return i * (i + 1) / 2 + k; }
__host__ __device__ inline size_t stc_slab(int nbp_max)
{
    const int nbk = nbp_max / 16;
    return (size_t)(nbk * (nbk + 1) + 2 * nbk) * 256;
}

template <typename T, int CW, int MAXR, int NW, int RG, bool RL>
__global__ __launch_bounds__(64 * NW) void stc_strip_kernel(StcArgs a)
{
    typedef Mfma<double> MD;
    constexpr int WGS = 64 * NW;        // NW waves
    constexpr int CT = CW / 16;         // tile columns of a strip
    constexpr int STR = CW + 2;         // LDS row stride of the strip
    constexpr int NBPX = 16 * NW * MAXR; // largest padded bubble block
    __shared__ double xs[NBPX * STR];   // the strip: solved block rows as they are broadcast
    __shared__ double dg[16 * 17], li[16 * 17], rsd[16];
    __shared__ double stg[NW][16 * 17];
    __shared__ int s_n, s_nb, s_fail;
    const int e = blockIdx.x, tid = threadIdx.x, lane = tid & 63;
    const int wave = __builtin_amdgcn_readfirstlane(tid >> 6);
    const int lr = lane & 15, lq = lane >> 4; // A/B operand: row lr, k-slice lq
    if (tid == 0) {
        int n = 0, nb = 0;
        bool ok = true;
        if (a.nsz) {
            n = a.nsz[2 * e];
            nb = a.nsz[2 * e + 1];
        } else {
            int zero[18] = {0};
            ElemOrders o;
            ok = decode_orders(a.norder + 19 * (size_t)e, zero, zero + 12, o);
            if (ok) dof_counts(a.space, o, n, nb);
        }
        s_n = ok ? n : -1;
        s_nb = nb;
        s_fail = 0;
    }
    __syncthreads();
    const int n = s_n, nb = s_nb;
    if (n < 0) {
        if (tid == 0 && a.status) a.status[e] = HP3D_E_ORDER;
        return;
    }
    if (n > a.n_max || nb > a.nb_max) { // above the caller's declared bounds
        if (tid == 0 && a.status) a.status[e] = HP3D_E_ARG;
        return;
    }
    const int ni = n - nb;
    const T *K = (const T *)a.kin + a.off_k[e];
    const size_t ldk = a.ld_k > 0 ? (size_t)a.ld_k : (size_t)n;
    const size_t ldc = a.ld_c > 0 ? (size_t)a.ld_c : (size_t)ni;
    const double *b = a.bin + a.off_v[e];
    double *aii = a.aii + a.off_c[e];
    double *bi = a.bi + a.off_v[e];
    if (nb == 0) { // nothing to condense: copy through (stc.F90:207)
        for (int idx = tid; idx < ni * ni; idx += WGS) {
            int i = idx % ni, j = idx / ni;
            aii[j * ldc + i] = (double)K[j * ldk + i];
        }
        for (int i = tid; i < ni; i += WGS) bi[i] = b[i];
        if (tid == 0 && a.status) a.status[e] = a.status[e] == HP3D_NEG_JACOBIAN ? HP3D_NEG_JACOBIAN : 0;
        return;
    }
    if (ni == 0) {
        if (tid == 0 && a.status) a.status[e] = HP3D_E_NOIFACE;
        return;
    }
    const int nbp = rnd16(nb), nbk = nbp / 16;
    const int nib = (ni + 15) / 16; // block rows of Aii
    if (nbp > a.nbp_max || nbk > NW * MAXR) { // slab or strip too small
        if (tid == 0 && a.status) a.status[e] = HP3D_E_ARG;
        return;
    }
    // the register-resident Cholesky holds up to RLB block rows
    constexpr int RLB = 15;
    if (RL && nbk > RLB) {
        if (tid == 0 && a.status) a.status[e] = HP3D_E_ARG;
        return;
    }
    const int nt = nbk * (nbk + 1) / 2;
    double *W = a.ws + (size_t)e * stc_slab(a.nbp_max);
    double *Mt = W + (size_t)nt * 256;
    double *Linv = Mt + (size_t)nt * 256, *LinvT = Linv + (size_t)nbk * 256;
    double *st = stg[wave];
    const int ab = lr + 16 * lq;    // a tile's A(row lr, col lq) in column-major, k-step 0
    const int abT = lq + 16 * lr;   // the transposed tile's
    // Loads are unconditional, from an index clamped to 0 where the value is
    // not wanted, and their values are converted or masked only where they
    // are used: a load whose value is used inside a lane-predicated block is
    // waited for right there, which serialises the loads.

    // the diagonal block p, symmetric in dg: factored by the calling wave,
    // L_pp^-1 into li and to the workspace with its transpose
    auto factor_diag = [&](int p) {
        // Abb_pp = L_u D L_u^T by the one-wave elimination of [Abb_pp | I]
        // (L_u^-1 into li, D^-1/2 into rsd), then L_pp^-1 = D^-1/2 L_u^-1
        const bool bad = eliminate_wave<16>(dg, 17, 0, 16, li, 17, rsd, lane);
        if (bad && lane == 0) s_fail = 1;
        wave_sync_lds();
        for (int idx = lane; idx < 256; idx += 64) {
            const int m = idx >> 4, c = idx & 15;
            li[m * 17 + c] *= rsd[m];
        }
        wave_sync_lds();
        for (int idx = lane; idx < 256; idx += 64) {
            const int r = idx & 15, c = idx >> 4; // column-major tiles
            Linv[(size_t)p * 256 + idx] = li[r * 17 + c];
            LinvT[(size_t)p * 256 + idx] = li[c * 17 + r];
        }
    };

    if constexpr (!RL) {
    // ---- 1: Cholesky, left-looking by block panels (more block rows than
    // the registers can hold)
    for (int p = 0; p < nbk; ++p) {
        // A(i,p) of Abb (K symmetric: read along rows), padded with the identity
        T araw[MAXR][4];
#pragma unroll
        for (int r = 0; r < MAXR; ++r) {
            const int i = wave + NW * r;
#pragma unroll
            for (int j = 0; j < 4; ++j) {
                const int rr = 16 * i + MD::row(lane, j), cc = 16 * p + lr;
                const bool v = i >= p && i < nbk && rr < nb && cc < nb;
                araw[r][j] = K[v ? (size_t)(ni + cc) + (size_t)(ni + rr) * ldk : 0];
            }
        }
        // sum_k L(i,k) L(p,k)^T, the tiles of step k + 1 loaded during step k
        d4_t csum[MAXR];
#pragma unroll
        for (int r = 0; r < MAXR; ++r) csum[r] = d4_t{0.0, 0.0, 0.0, 0.0};
        double ta[MAXR][4], tb[4];
        auto ld_k = [&](int k) {
            const double *Lpk = W + (size_t)stc_tl(p, k) * 256;
#pragma unroll
            for (int s = 0; s < 4; ++s) tb[s] = Lpk[ab + 64 * s];
#pragma unroll
            for (int r = 0; r < MAXR; ++r) {
                const int i = wave + NW * r;
                const double *Lik = (i >= p && i < nbk) ? W + (size_t)stc_tl(i, k) * 256 : Lpk;
#pragma unroll
                for (int s = 0; s < 4; ++s) ta[r][s] = Lik[ab + 64 * s];
            }
        };
        ld_k(0); // (p = 0: the diagonal tile, unused)
        for (int k = 0; k < p; ++k) {
            double ca[MAXR][4], cb[4];
#pragma unroll
            for (int s = 0; s < 4; ++s) {
                cb[s] = tb[s];
#pragma unroll
                for (int r = 0; r < MAXR; ++r) ca[r][s] = ta[r][s];
            }
            ld_k(min(k + 1, p - 1));
#pragma unroll
            for (int r = 0; r < MAXR; ++r) {
                const int i = wave + NW * r;
                if (i < p || i >= nbk) continue;
#pragma unroll
                for (int s = 0; s < 4; ++s) csum[r] = MD::mma(ca[r][s], cb[s], csum[r]);
            }
        }
        d4_t cacc[MAXR];
#pragma unroll
        for (int r = 0; r < MAXR; ++r) {
            const int i = wave + NW * r;
#pragma unroll
            for (int j = 0; j < 4; ++j) {
                const int rr = 16 * i + MD::row(lane, j), cc = 16 * p + lr;
                const double av = (rr < nb && cc < nb) ? (double)araw[r][j] : (rr == cc ? 1.0 : 0.0);
                cacc[r][j] = av - csum[r][j];
            }
        }
#pragma unroll
        for (int r = 0; r < MAXR; ++r)
            if (wave + NW * r == p) {
#pragma unroll
                for (int j = 0; j < 4; ++j) dg[MD::row(lane, j) * 17 + lr] = cacc[r][j];
            }
        __syncthreads();
        if (wave == p % NW) factor_diag(p);
        __syncthreads();
        // the panel below: L(i,p) = A~(i,p) L_pp^-T
#pragma unroll
        for (int r = 0; r < MAXR; ++r) {
            const int i = wave + NW * r;
            if (i <= p || i >= nbk) continue;
#pragma unroll
            for (int j = 0; j < 4; ++j) st[MD::row(lane, j) * 17 + lr] = cacc[r][j];
            wave_sync_lds();
            d4_t d = {0.0, 0.0, 0.0, 0.0};
#pragma unroll
            for (int s = 0; s < 4; ++s) {
                const int kk = 4 * s + lq;
                d = MD::mma(st[lr * 17 + kk], li[lr * 17 + kk], d);
            }
            wave_sync_lds();
            double *Lip = W + (size_t)stc_tl(i, p) * 256;
#pragma unroll
            for (int j = 0; j < 4; ++j) Lip[MD::row(lane, j) + 16 * lr] = d[j];
        }
        __syncthreads();
    }
    } else {
    // ---- 1: Cholesky, right-looking, the lower triangle of Abb held in
    // registers as 16 x 16 tiles (tile t = (i, j) of the lower triangle by
    // wave t mod 8): per block column p the owner of the diagonal tile
    // factors it, the owners of the tiles below form the panel
    // L(i,p) = A~(i,p) L_pp^-T and broadcast it through LDS, and every wave
    // updates its trailing tiles A~(i,j) -= L(i,p) L(j,p)^T
    constexpr int NSL = (RLB * (RLB + 1) / 2 + NW - 1) / NW; // tile slots per wave
    d4_t tl[NSL];
    // (i, j) of slot u's tile t = wave + 8u, walked incrementally
    auto slot_ij = [&](auto&& f) {
        int i = 0, j = wave;
        while (j > i) {
            j -= i + 1;
            ++i;
        }
#pragma unroll
        for (int u = 0; u < NSL; ++u) {
            f(u, wave + NW * u < nt ? i : -1, j);
            j += NW;
            while (j > i) {
                j -= i + 1;
                ++i;
            }
        }
    };
    // A(i,j) of Abb (K symmetric: read along rows), loaded without
    // conditions, then the padding set to the identity
    slot_ij([&](int u, int i, int j) {
#pragma unroll
        for (int q = 0; q < 4; ++q) {
            const int rr = 16 * i + MD::row(lane, q), cc = 16 * j + lr;
            const bool v = i >= 0 && rr < nb && cc < nb;
            tl[u][q] = (double)K[v ? (size_t)(ni + cc) + (size_t)(ni + rr) * ldk : 0];
        }
    });
    slot_ij([&](int u, int i, int j) {
#pragma unroll
        for (int q = 0; q < 4; ++q) {
            const int rr = 16 * i + MD::row(lane, q), cc = 16 * j + lr;
            if (!(i >= 0 && rr < nb && cc < nb)) tl[u][q] = rr == cc ? 1.0 : 0.0;
        }
    });
    double *pan = xs; // the panel's tiles L(i,p), column-major, at pan + 256 i
    for (int p = 0; p < nbk; ++p) {
        if (wave == stc_tl(p, p) % NW) {
            slot_ij([&](int u, int i, int j) {
                if (i == p && j == p) {
#pragma unroll
                    for (int q = 0; q < 4; ++q) dg[MD::row(lane, q) * 17 + lr] = tl[u][q];
                }
            });
            wave_sync_lds();
            factor_diag(p);
        }
        lds_barrier();
        slot_ij([&](int u, int i, int j) {
            if (j == p && i > p) {
#pragma unroll
                for (int q = 0; q < 4; ++q) st[MD::row(lane, q) * 17 + lr] = tl[u][q];
                wave_sync_lds();
                d4_t d = {0.0, 0.0, 0.0, 0.0};
#pragma unroll
                for (int s2 = 0; s2 < 4; ++s2) {
                    const int kk = 4 * s2 + lq;
                    d = MD::mma(st[lr * 17 + kk], li[lr * 17 + kk], d);
                }
                wave_sync_lds();
                double *Lip = W + (size_t)stc_tl(i, p) * 256, *Pp = pan + (size_t)256 * i;
#pragma unroll
                for (int q = 0; q < 4; ++q) {
                    Lip[MD::row(lane, q) + 16 * lr] = d[q];
                    Pp[MD::row(lane, q) + 16 * lr] = d[q];
                }
            }
        });
        lds_barrier();
        slot_ij([&](int u, int i, int j) {
            if (i >= 0 && j > p) {
                const double *Pi = pan + (size_t)256 * i, *Pj = pan + (size_t)256 * j;
                d4_t sum = {0.0, 0.0, 0.0, 0.0};
#pragma unroll
                for (int s2 = 0; s2 < 4; ++s2) sum = MD::mma(Pi[ab + 64 * s2], Pj[ab + 64 * s2], sum);
                tl[u] -= sum;
            }
        });
    }
    __syncthreads();
    }

    // ---- the sweeps' tiles: -M(i,p) = -L(i,p) L_pp^-1 (forward) and
    // -Mb(i,p) = -L_ii^-1 L(i,p) in place of L(i,p) (backward), i > p, so that
    // a step broadcasts its block row before the diagonal solve (the row's
    // owner solves it off the critical path)
    for (int tix = wave; tix < nt; tix += NW) {
        int i = 0;
        while ((i + 1) * (i + 2) / 2 <= tix) ++i;
        const int pcol = tix - i * (i + 1) / 2;
        if (pcol == i) continue;
        double *Lt = W + (size_t)tix * 256;
        const double *Dp = Linv + (size_t)pcol * 256, *Di = Linv + (size_t)i * 256;
        d4_t m = {0.0, 0.0, 0.0, 0.0}, mb = {0.0, 0.0, 0.0, 0.0};
#pragma unroll
        for (int s2 = 0; s2 < 4; ++s2) {
            m = MD::mma(Lt[ab + 64 * s2], Dp[abT + 4 * s2], m);
            mb = MD::mma(Di[ab + 64 * s2], Lt[abT + 4 * s2], mb);
        }
#pragma unroll
        for (int j = 0; j < 4; ++j) { // stored negated: the sweeps subtract
            Mt[(size_t)tix * 256 + MD::row(lane, j) + 16 * lr] = -m[j];
            Lt[MD::row(lane, j) + 16 * lr] = -mb[j];
        }
    }
    __syncthreads();

    // ---- 2: strips of [Abi | Bb]
    const int ncol = ni + 1;
    for (int c0 = 0; c0 < ncol; c0 += CW) {
        d4_t acc[MAXR][CT];
        // Z(r, c) = Abi(r, c) (K symmetric: read along rows of Aib), Bb in column ni
        {
            T zraw[MAXR][CT][4];
            double braw[MAXR][4];
#pragma unroll
            for (int r = 0; r < MAXR; ++r) {
                const int i = wave + NW * r;
#pragma unroll
                for (int j = 0; j < 4; ++j) {
                    const int rb = 16 * i + MD::row(lane, j);
                    const bool vr = i < nbk && rb < nb;
                    braw[r][j] = b[vr ? ni + rb : 0];
#pragma unroll
                    for (int t = 0; t < CT; ++t) {
                        const int c = c0 + 16 * t + lr;
                        zraw[r][t][j] = K[(vr && c < ni) ? (size_t)c + (size_t)(ni + rb) * ldk : 0];
                    }
                }
            }
#pragma unroll
            for (int r = 0; r < MAXR; ++r) {
                const int i = wave + NW * r;
#pragma unroll
                for (int j = 0; j < 4; ++j) {
                    const int rb = 16 * i + MD::row(lane, j);
                    const bool vr = i < nbk && rb < nb;
#pragma unroll
                    for (int t = 0; t < CT; ++t) {
                        const int c = c0 + 16 * t + lr;
                        acc[r][t][j] = !vr ? 0.0 : (c < ni ? (double)zraw[r][t][j] : (c == ni ? braw[r][j] : 0.0));
                    }
                }
            }
        }
        // the owner's updated block row p goes to rows 16p.. of the strip in
        // LDS, then is solved in registers against a diagonal tile (column-
        // major D): acc <- D acc
        auto put_row = [&](const d4_t (&row)[CT], int p) {
            double *buf = xs + (size_t)16 * p * STR;
#pragma unroll
            for (int t = 0; t < CT; ++t)
#pragma unroll
                for (int j = 0; j < 4; ++j) buf[MD::row(lane, j) * STR + 16 * t + lr] = row[t][j];
        };
        auto solve_row = [&](d4_t (&row)[CT], const double (&dA)[4], int p) {
            const double *buf = xs + (size_t)16 * p * STR;
            wave_sync_lds();
#pragma unroll
            for (int t = 0; t < CT; ++t) {
                d4_t d = {0.0, 0.0, 0.0, 0.0};
#pragma unroll
                for (int s2 = 0; s2 < 4; ++s2) d = MD::mma(dA[s2], buf[(4 * s2 + lq) * STR + 16 * t + lr], d);
                row[t] = d;
            }
        };
        // one sweep: step p broadcasts the owner's updated row p, every wave
        // subtracts T(i,p) row_p from its rows i (forward: i > p, T = M;
        // backward: i < p, T = Mb(p,i)^T) and the owner of the next row
        // updates and broadcasts it first; the tiles of step p + 1 are loaded
        // during step p.  No row is needed solved during the sweep: each wave
        // solves its own rows at the end, against their diagonal tiles.
        auto sweep = [&](bool fwd) {
            double nL[MAXR][4];
            auto ld_step = [&](int p) {
#pragma unroll
                for (int r = 0; r < MAXR; ++r) {
                    const int i = wave + NW * r;
                    const bool v = fwd ? (i > p && i < nbk) : (i < p);
                    const double *Tt = !v ? W : fwd ? Mt + (size_t)stc_tl(i, p) * 256 : W + (size_t)stc_tl(p, i) * 256;
#pragma unroll
                    for (int s2 = 0; s2 < 4; ++s2) nL[r][s2] = Tt[(fwd ? ab + 64 * s2 : abT + 4 * s2)];
                }
            };
            const int p0 = fwd ? 0 : nbk - 1, dp = fwd ? 1 : -1;
            const double *Dg = fwd ? Linv : LinvT;
            double dS[MAXR][4]; // diagonal tiles of this wave's rows
#pragma unroll
            for (int r = 0; r < MAXR; ++r) {
                const int i = min(wave + NW * r, nbk - 1);
#pragma unroll
                for (int s2 = 0; s2 < 4; ++s2) dS[r][s2] = Dg[(size_t)i * 256 + ab + 64 * s2];
            }
            ld_step(p0);
            if (wave == p0 % NW) {
#pragma unroll
                for (int r = 0; r < MAXR; ++r)
                    if (r == p0 / NW) put_row(acc[r], p0);
            }
            for (int p = p0; p >= 0 && p < nbk; p += dp) {
                double cL[MAXR][4];
#pragma unroll
                for (int r = 0; r < MAXR; ++r)
#pragma unroll
                    for (int s2 = 0; s2 < 4; ++s2) cL[r][s2] = nL[r][s2];
                const int pn = p + dp;
                const bool more = pn >= 0 && pn < nbk;
                ld_step(more ? pn : p); // loads stay unconditional
                const bool own_next = more && wave == pn % NW;
                lds_barrier();
                const double *buf = xs + (size_t)16 * p * STR;
                auto upd = [&](int r) {
#pragma unroll
                    for (int s2 = 0; s2 < 4; ++s2) {
#pragma unroll
                        for (int t = 0; t < CT; ++t)
                            acc[r][t] = MD::mma(cL[r][s2], buf[(4 * s2 + lq) * STR + 16 * t + lr], acc[r][t]);
                    }
                };
                // the next row first: update, broadcast
#pragma unroll
                for (int r = 0; r < MAXR; ++r)
                    if (own_next && r == pn / NW) {
                        upd(r);
                        put_row(acc[r], pn);
                    }
#pragma unroll
                for (int r = 0; r < MAXR; ++r) {
                    const int i = wave + NW * r;
                    if (own_next && r == pn / NW) continue;
                    if (fwd ? (i <= p || i >= nbk) : (i >= p)) continue;
                    upd(r);
                }
            }
            // the wave's own rows: their final broadcast values in LDS, solved
#pragma unroll
            for (int r = 0; r < MAXR; ++r)
                if (wave + NW * r < nbk) solve_row(acc[r], dS[r], wave + NW * r);
        };
        sweep(true);   // S = L^-1 Z (rows of S in registers)
        lds_barrier(); // every wave has read the forward rows
        sweep(false);  // X = L^-T S
        lds_barrier(); // every wave has read the backward rows
        // the solved X into the strip in LDS (for Aib X)
#pragma unroll
        for (int r = 0; r < MAXR; ++r) {
            const int i = wave + NW * r;
            if (i < nbk) put_row(acc[r], i);
        }
        // outputs: X -> ASchur (column-major, leading dimension nb), BSchur
        double *X = a.aschur + a.off_s[e];
        double *BS = a.bschur + a.off_v[e];
#pragma unroll
        for (int r = 0; r < MAXR; ++r) {
            const int i = wave + NW * r;
            if (i >= nbk) continue;
#pragma unroll
            for (int t = 0; t < CT; ++t) {
#pragma unroll
                for (int j = 0; j < 4; ++j) st[MD::row(lane, j) * 17 + lr] = acc[r][t][j];
                wave_sync_lds();
                const int rb = 16 * i + lr;
#pragma unroll
                for (int j = 0; j < 4; ++j) {
                    const int c = c0 + 16 * t + lq + 4 * j; // column of element (lr, lq + 4j)
                    const double v = st[lr * 17 + lq + 4 * j];
                    if (rb < nb) {
                        if (c < ni) X[(size_t)c * nb + rb] = v;
                        else if (c == ni) BS[rb] = v;
                    }
                }
                wave_sync_lds();
            }
        }
        lds_barrier(); // the whole strip X in LDS
        // Aib X on this strip's columns: Aii rows up to the strip's last
        // column (upper triangle and the strip's diagonal block), all rows on
        // the load column's strip.  Block row I of Aii goes to wave I mod 8,
        // RG rows at a time with all CT tile columns, so each Aib value feeds
        // CT MFMAs.
        const int nI = min(nib, (c0 + CW) / 16);
        for (int I0 = wave; I0 < nI; I0 += NW * RG) {
            d4_t accA[RG][CT];
#pragma unroll
            for (int g = 0; g < RG; ++g)
#pragma unroll
                for (int t = 0; t < CT; ++t) accA[g][t] = d4_t{0.0, 0.0, 0.0, 0.0};
            // Aib(ri, kb) = K(ri, ni + kb), the values of step p + 1 loaded
            // during step p
            T nA[RG][4];
            auto ld_A = [&](int p) {
#pragma unroll
                for (int g = 0; g < RG; ++g) {
                    const int ri = 16 * (I0 + NW * g) + lr;
#pragma unroll
                    for (int s2 = 0; s2 < 4; ++s2) {
                        const int kb = 16 * p + 4 * s2 + lq;
                        nA[g][s2] = K[(ri < ni && kb < nb) ? (size_t)ri + (size_t)(ni + kb) * ldk : 0];
                    }
                }
            };
            ld_A(0);
            for (int p = 0; p < nbk; ++p) {
                double cA[RG][4];
#pragma unroll
                for (int g = 0; g < RG; ++g) {
                    const int ri = 16 * (I0 + NW * g) + lr;
#pragma unroll
                    for (int s2 = 0; s2 < 4; ++s2) {
                        const int kb = 16 * p + 4 * s2 + lq;
                        cA[g][s2] = (ri < ni && kb < nb) ? (double)nA[g][s2] : 0.0;
                    }
                }
                ld_A(min(p + 1, nbk - 1));
#pragma unroll
                for (int s2 = 0; s2 < 4; ++s2) {
#pragma unroll
                    for (int t = 0; t < CT; ++t) {
                        const double bv = xs[(size_t)(16 * p + 4 * s2 + lq) * STR + 16 * t + lr];
#pragma unroll
                        for (int g = 0; g < RG; ++g)
                            if (I0 + NW * g < nI) accA[g][t] = MD::mma(cA[g][s2], bv, accA[g][t]);
                    }
                }
            }
            // Aii = Kii - Aib X (both triangles), Bi = Bi - Aib BSchur
#pragma unroll
            for (int g = 0; g < RG; ++g) {
                const int I = I0 + NW * g;
                if (I >= nI) continue;
#pragma unroll
                for (int t = 0; t < CT; ++t) {
                    d4_t v4 = accA[g][t];
                    T kraw[4];
#pragma unroll
                    for (int j = 0; j < 4; ++j) {
                        const int r = 16 * I + MD::row(lane, j), c = c0 + 16 * t + lr;
                        kraw[j] = K[(r < ni && c < ni) ? (size_t)c + (size_t)r * ldk : 0];
                    }
                    // mirrored entries (c, r) first, straight from the
                    // accumulator: lanes run along c, i.e. along the rows of column r
#pragma unroll
                    for (int j = 0; j < 4; ++j) {
                        const int r = 16 * I + MD::row(lane, j), c = c0 + 16 * t + lr;
                        if (r >= ni) continue;
                        if (c < ni) {
                            const double v = (double)kraw[j] - v4[j];
                            v4[j] = v;
                            if (r < c0) aii[(size_t)c + (size_t)r * ldc] = v; // (c, r) of the lower part
                        } else if (c == ni) {
                            bi[r] = b[r] - v4[j];
                        }
                    }
                    // the entries (r, c) themselves, transposed through LDS so lanes run along r
#pragma unroll
                    for (int j = 0; j < 4; ++j) st[MD::row(lane, j) * 17 + lr] = v4[j];
                    wave_sync_lds();
#pragma unroll
                    for (int j = 0; j < 4; ++j) {
                        const int r = 16 * I + lr, c = c0 + 16 * t + lq + 4 * j;
                        if (r < ni && c < ni) aii[(size_t)r + (size_t)c * ldc] = st[lr * 17 + lq + 4 * j];
                    }
                    wave_sync_lds();
                }
            }
        }
        lds_barrier(); // the strip's LDS is read before the next strip writes it
    }
    __syncthreads();
    // the element pass's negative-Jacobian warning survives a successful factorisation
    if (tid == 0 && a.status)
        a.status[e] = s_fail ? HP3D_E_FACTOR : (a.status[e] == HP3D_NEG_JACOBIAN ? HP3D_NEG_JACOBIAN : 0);
}

template <typename T>
static void launch_stc(StcArgs a, int max_nbubble, hipStream_t st)
{
    // 64-column strips (2 block rows per wave) up to 256 bubbles, with the
    // register-resident Cholesky up to 240; 32-column strips (4 block rows
    // per wave) beyond
    const int nbp = rnd16(max_nbubble);
    if (nbp <= 240)
        hipLaunchKernelGGL((stc_strip_kernel<T, 64, 2, 8, 2, true>), dim3(a.nel), dim3(512), 0, st, a);
    else if (nbp <= 256)
        hipLaunchKernelGGL((stc_strip_kernel<T, 64, 2, 8, 2, false>), dim3(a.nel), dim3(512), 0, st, a);
    else
        hipLaunchKernelGGL((stc_strip_kernel<T, 32, 4, 8, 2, false>), dim3(a.nel), dim3(512), 0, st, a);
}

struct BwdArgs {
    int space, nel;
    const int *norder;
    const long long *off_s, *off_v;
    const double *aschur, *bschur;
    double *x;
};

__global__ __launch_bounds__(WG) void stc_bwd_kernel(BwdArgs a)
{
    const int e = blockIdx.x;
    const int *no = a.norder + 19 * (size_t)e;
    int zero[18] = {0};
    ElemOrders o;
    if (!decode_orders(no, zero, zero + 12, o)) return;
    int n, nb;
    dof_counts(a.space, o, n, nb);
    int ni = n - nb;
    const double *A = a.aschur + a.off_s[e];
    double *x = a.x + a.off_v[e];
    for (int q = threadIdx.x; q < nb; q += WG) {
        double s = 0.0;
        for (int i = 0; i < ni; ++i) s += A[(size_t)i * nb + q] * x[i];
        x[ni + q] = a.bschur[a.off_v[e] + q] - s;
    }
}

// Dirichlet lift of the condensed element loads (celem_systemI.F90:794-808,
// ZBMOD(:) -= ZAMOD(:,k2) ZDOFD(k2) for every Dirichlet dof k2), on the
// condensed system: bi[k] -= sum_k2 Aii[k, k2] ud[k2], with ud the element's
// interface vector of Dirichlet values (0 at free dofs).  One workgroup per
// element, a thread per row.
struct LiftArgs {
    int space, nel;
    const int *norder;
    const long long *off_c, *off_v;
    const double *aii, *ud;
    double *bi;
};

__global__ __launch_bounds__(WG) void dirichlet_lift_kernel(LiftArgs a)
{
    const int e = blockIdx.x;
    int zero[18] = {0};
    ElemOrders o;
    if (!decode_orders(a.norder + 19 * (size_t)e, zero, zero + 12, o)) return;
    int n, nb;
    dof_counts(a.space, o, n, nb);
    const int ni = n - nb;
    const double *A = a.aii + a.off_c[e], *u = a.ud + a.off_v[e];
    double *b = a.bi + a.off_v[e];
    for (int k = threadIdx.x; k < ni; k += WG) {
        double s = 0.0;
        for (int k2 = 0; k2 < ni; ++k2) s += A[(size_t)k2 * ni + k] * u[k2];
        b[k] -= s;
    }
}

// ------------------------------------------------------- field evaluation
//
// One workgroup per element evaluates finite-element fields from element dof
// vectors at the quadrature points (soleval, element/util/soleval.F90, with
// the dofs in element order) and integrates one of:
//   HP3D_FIELD_ERROR    element_error (element/util/compute_error.F90:226-400):
//                       set_3Dint rule with INTEGRATION = 2; out[2e] = errorH
//                       (int |grad(u - u_h)|^2 + |u - u_h|^2) or errorE (curl
//                       and value), out[2e+1] = the same norm of the exact u
//   HP3D_FIELD_ADJOINT  out[e] = -coef_e int grad v_h . grad u_h (H1) or
//                       -coef_e int v_h . u_h (H(curl)): lambda_e^T (dK_e /
//                       dlog coef_e) u_e with a minus sign, i.e. dJ/dtheta_e
//                       for coef = exp(theta), recomputed at the points of the
//                       element rule without forming K_e
//   HP3D_FIELD_L2       out[e] = int u_h . u_h and r[off_v[e] + k] = int
//                       phi_k . u_h (the element mass matrix times u_e)
// Threads own quadrature points; the dof loop reads descriptors and
// coefficients from LDS (same address across the wave: broadcast).

struct FieldArgs {
    int space, mode, nel, ngeo, exact;
    const int *norder, *nori_e, *nori_f;
    const double *xnod, *coef;
    const long long *off_v;
    const double *u, *v;
    double *out, *r;
    int *status;
};

// exact solutions of the manufactured problems (value, gradient); the
// selector is the load selector of the same problem
__device__ inline void exact_h1(int sel, const double *x, double &u, double *du)
{
    const double pi = 3.14159265358979323846;
    u = du[0] = du[1] = du[2] = 0.0;
    if (sel == HP3D_LOAD_SINSINSIN) {
        double sx = sin(pi * x[0]), sy = sin(pi * x[1]), sz = sin(pi * x[2]);
        u = sx * sy * sz;
        du[0] = pi * cos(pi * x[0]) * sy * sz;
        du[1] = pi * sx * cos(pi * x[1]) * sz;
        du[2] = pi * sx * sy * cos(pi * x[2]);
    } else if (sel == HP3D_LOAD_CONV) {
        double sx = sin(pi * x[0]), sh = sinh(pi * x[1]), z2 = x[2] * x[2];
        u = sx * sh * z2;
        du[0] = pi * cos(pi * x[0]) * sh * z2;
        du[1] = pi * sx * cosh(pi * x[1]) * z2;
        du[2] = 2.0 * sx * sh * x[2];
    } else if (sel == HP3D_LOAD_POLY2) {
        double a0 = x[0] * (1.0 - x[0]), a1 = x[1] * (1.0 - x[1]), a2 = x[2] * (1.0 - x[2]);
        u = a0 * a1 * a2;
        du[0] = (1.0 - 2.0 * x[0]) * a1 * a2;
        du[1] = a0 * (1.0 - 2.0 * x[1]) * a2;
        du[2] = a0 * a1 * (1.0 - 2.0 * x[2]);
    } else if (sel == HP3D_LOAD_XYZ) {
        u = x[0] * x[1] * x[2];
        du[0] = x[1] * x[2];
        du[1] = x[0] * x[2];
        du[2] = x[0] * x[1];
    }
}

__device__ inline void exact_hcurl(int sel, const double *x, double *e, double *curl)
{
    const double pi = 3.14159265358979323846;
    for (int i = 0; i < 3; ++i) e[i] = curl[i] = 0.0;
    if (sel == HP3D_LOAD_CURL_SIN) {
        e[0] = sin(pi * x[1]) * sin(pi * x[2]);
        curl[1] = pi * sin(pi * x[1]) * cos(pi * x[2]);
        curl[2] = -pi * cos(pi * x[1]) * sin(pi * x[2]);
    }
}

// set_3Dint brick rule (set_3D_int.F90:8-21,203-244): face orders as coded
// (no swap), orders raised by INTEGRATION and capped at MAXP
__device__ inline void quad_counts_plain(const ElemOrders &o, int integration, int nq[3])
{
    int m[3] = {o.b[0], o.b[1], o.b[2]};
    for (int e = 0; e < 12; ++e) {
        int d = edge_t(e).dir;
        m[d] = o.pe[e] > m[d] ? o.pe[e] : m[d];
    }
    for (int f = 0; f < 6; ++f) {
        FaceT ft = face_t(f);
        m[ft.ax0] = o.fh[f] > m[ft.ax0] ? o.fh[f] : m[ft.ax0];
        m[ft.ax1] = o.fv[f] > m[ft.ax1] ? o.fv[f] : m[ft.ax1];
    }
    for (int d = 0; d < 3; ++d) {
        int q = m[d] + integration;
        nq[d] = (q < MAXP ? q : MAXP) + 1;
    }
}

// reference-coordinate value and gradient of the scalar dof d at point p
__device__ __forceinline__ void dof_ref(unsigned d, Pt p, const double (*tv)[NTAB * NPT],
                                        const double (*td)[NTAB * NPT], double &ph, double gr[3])
{
    int i0 = desc_ix(d, 0), i1 = desc_ix(d, 1), i2 = desc_ix(d, 2);
    double v0 = tv[0][p.l1 * NTAB + i0], v1 = tv[1][p.l2 * NTAB + i1], v2 = tv[2][p.l3 * NTAB + i2];
    double d0 = td[0][p.l1 * NTAB + i0], d1 = td[1][p.l2 * NTAB + i1], d2 = td[2][p.l3 * NTAB + i2];
    ph = v0 * v1 * v2;
    gr[0] = d0 * v1 * v2;
    gr[1] = v0 * d1 * v2;
    gr[2] = v0 * v1 * d2;
}

// field of coefficients c at point p in reference form: H1 value f[0] and
// gradient g; H(curl) E_ref in f[0..2] and curl_ref in g
template <int SPACE>
__device__ __forceinline__ void eval_ref(int n, const unsigned *dsc, const double *c, Pt p,
                                         const double (*tv)[NTAB * NPT], const double (*td)[NTAB * NPT],
                                         double f[3], double g[3])
{
    for (int i = 0; i < 3; ++i) f[i] = g[i] = 0.0;
    for (int k = 0; k < n; ++k) {
        unsigned d = dsc[k];
        double ph, gr[3];
        dof_ref(d, p, tv, td, ph, gr);
        double ck = c[k];
        if (SPACE == 0) {
            f[0] += ck * ph;
            for (int i = 0; i < 3; ++i) g[i] += ck * gr[i];
        } else {
            int cp = desc_comp(d);
            double sg = desc_neg(d) ? -ck : ck;
            // E = s phi e_cp, curl E = s grad(phi) x e_cp
            f[cp] += sg * ph;
            g[(cp + 1) % 3] += sg * gr[(cp + 2) % 3];
            g[(cp + 2) % 3] -= sg * gr[(cp + 1) % 3];
        }
    }
}

// physical value / gradient (H1) or E / curl E (H(curl)) from the reference form
template <int SPACE>
__device__ __forceinline__ void to_phys(const double f[3], const double g[3], double jm[3][3], double inv[3][3],
                                        double det, double fp[3], double gp[3])
{
    if (SPACE == 0) {
        fp[0] = f[0];
        fp[1] = fp[2] = 0.0;
        for (int j = 0; j < 3; ++j) gp[j] = g[0] * inv[0][j] + g[1] * inv[1][j] + g[2] * inv[2][j];
    } else {
        for (int j = 0; j < 3; ++j) {
            fp[j] = f[0] * inv[0][j] + f[1] * inv[1][j] + f[2] * inv[2][j];
            gp[j] = (jm[j][0] * g[0] + jm[j][1] * g[1] + jm[j][2] * g[2]) / det;
        }
    }
}

template <int SPACE>
__global__ __launch_bounds__(WG) void field_kernel(FieldArgs a)
{
    constexpr int MAXN = SPACE == 0 ? MAXN_H1 : MAXN_E;
    __shared__ ElemShared es;
    __shared__ double tv[3][NTAB * NPT], td[3][NTAB * NPT];
    __shared__ unsigned dsc[MAXN];
    __shared__ unsigned dgeo[MAXN_H1];
    __shared__ double cu[MAXN], cv[MAXN];
    __shared__ double pw[3 * MAXQ];
    __shared__ double red[2][WG / 64];

    const int e = blockIdx.x, tid = threadIdx.x, lane = tid & 63, wave = tid >> 6;
    if (tid == 0) {
        ElemShared s;
        s.ok = decode_orders(a.norder + 19 * (size_t)e, a.nori_e + 12 * (size_t)e,
                             a.nori_f + 6 * (size_t)e, s.o);
        s.n = s.nb = s.nint = s.nh = 0;
        s.nq[0] = s.nq[1] = s.nq[2] = 1;
        if (s.ok) {
            dof_counts(SPACE, s.o, s.n, s.nb);
            int nbh;
            dof_counts(0, s.o, s.nh, nbh);
            if (a.mode == HP3D_FIELD_ERROR) quad_counts_plain(s.o, 2, s.nq);
            else quad_counts(s.o, s.nq);
            s.nint = s.nq[0] * s.nq[1] * s.nq[2];
            if (a.ngeo > s.nh) s.ok = 0;
        }
        s.negjac = 0;
        es = s;
    }
    __syncthreads();
    if (!es.ok) {
        if (tid == 0 && a.status) a.status[e] = HP3D_E_ORDER;
        return;
    }
    const int n = es.n, nint = es.nint;
    if (wave < 3) fill_tables(es.nq[wave], tv[wave], td[wave]);
    const double *ue = a.u + a.off_v[e];
    const double *ve = a.mode == HP3D_FIELD_ADJOINT ? a.v + a.off_v[e] : nullptr;
    const ElemOrders ol = es.o; // in registers for the descriptor walks
    for (int k = tid; k < n; k += WG) {
        dsc[k] = SPACE == 0 ? desc_h1(k, ol) : desc_hcurl(k, ol);
        cu[k] = ue[k];
        if (ve) cv[k] = ve[k];
    }
    for (int k = tid; k < a.ngeo; k += WG) dgeo[k] = desc_h1(k, es.o);
    __syncthreads();

    const double *xn = a.xnod + 3 * (size_t)a.ngeo * e;
    const double coef = a.coef ? a.coef[e] : 1.0;
    double acc0 = 0.0, acc1 = 0.0;
    bool neg = false;
    for (int l = tid; l < nint; l += WG) {
        Pt p = point_ix(l, es.nq);
        double x[3], jm[3][3], inv[3][3], det, w;
        point_map(p, es.nq, tv, td, dgeo, xn, a.ngeo, x, jm, inv, det, w);
        neg = neg || det < 0.0;
        const double wd = w * det;
        double f[3], g[3], fu[3], gu[3];
        eval_ref<SPACE>(n, dsc, cu, p, tv, td, f, g);
        to_phys<SPACE>(f, g, jm, inv, det, fu, gu);
        if (a.mode == HP3D_FIELD_ERROR) {
            double fe[3], ge[3];
            if (SPACE == 0) {
                exact_h1(a.exact, x, fe[0], ge);
                fe[1] = fe[2] = 0.0;
            } else {
                exact_hcurl(a.exact, x, fe, ge);
            }
            for (int j = 0; j < 3; ++j) {
                acc0 += ((ge[j] - gu[j]) * (ge[j] - gu[j]) + (fe[j] - fu[j]) * (fe[j] - fu[j])) * wd;
                acc1 += (ge[j] * ge[j] + fe[j] * fe[j]) * wd;
            }
        } else if (a.mode == HP3D_FIELD_ADJOINT) {
            double fv[3], gv[3];
            eval_ref<SPACE>(n, dsc, cv, p, tv, td, f, g);
            to_phys<SPACE>(f, g, jm, inv, det, fv, gv);
            if (SPACE == 0) acc0 -= coef * (gu[0] * gv[0] + gu[1] * gv[1] + gu[2] * gv[2]) * wd;
            else acc0 -= coef * (fu[0] * fv[0] + fu[1] * fv[1] + fu[2] * fv[2]) * wd;
        } else {
            acc0 += (fu[0] * fu[0] + fu[1] * fu[1] + fu[2] * fu[2]) * wd;
            // weighted field pulled back to reference components: phi_k . u_h w det
            if (SPACE == 0) {
                pw[l] = fu[0] * wd;
            } else {
                for (int i = 0; i < 3; ++i)
                    pw[3 * l + i] = (inv[i][0] * fu[0] + inv[i][1] * fu[1] + inv[i][2] * fu[2]) * wd;
            }
        }
    }
    if (neg) es.negjac = 1;
    // workgroup sums
    for (int off = 32; off > 0; off >>= 1) {
        acc0 += __shfl_down(acc0, off);
        acc1 += __shfl_down(acc1, off);
    }
    if (lane == 0) {
        red[0][wave] = acc0;
        red[1][wave] = acc1;
    }
    __syncthreads();
    if (tid == 0) {
        double s0 = 0.0, s1 = 0.0;
        for (int w = 0; w < WG / 64; ++w) {
            s0 += red[0][w];
            s1 += red[1][w];
        }
        if (a.mode == HP3D_FIELD_ERROR) {
            a.out[2 * (size_t)e] = s0;
            a.out[2 * (size_t)e + 1] = s1;
        } else {
            a.out[e] = s0;
        }
        if (a.status) a.status[e] = es.negjac ? HP3D_NEG_JACOBIAN : 0;
    }
    if (a.mode == HP3D_FIELD_L2 && a.r) {
        double *re = a.r + a.off_v[e];
        for (int k = tid; k < n; k += WG) {
            unsigned d = dsc[k];
            double s = 0.0;
            for (int l = 0; l < nint; ++l) {
                Pt p = point_ix(l, es.nq);
                double ph, gr[3];
                dof_ref(d, p, tv, td, ph, gr);
                if (SPACE == 0) s += ph * pw[l];
                else s += (desc_neg(d) ? -ph : ph) * pw[3 * l + desc_comp(d)];
            }
            re[k] = s;
        }
    }
}

// Segmented gather-sum: out[t] = sum over j in [ptr[t], ptr[t+1]) of in[src[j]].
// With src listing, for every nonzero of a global matrix (or entry of a global
// vector), the positions of its element contributions in the packed condensed
// matrices (vectors), this is the assembly of the element systems into CSR
// values (mumps_sc.F90:266-323 accumulates the same triplets, duplicates
// summed); the fixed summation order makes it deterministic.
__global__ __launch_bounds__(WG) void gather_sum_kernel(long long nout, const long long *ptr,
                                                       const long long *src, const double *in,
                                                       double *out)
{
    long long t = (long long)blockIdx.x * WG + threadIdx.x;
    if (t >= nout) return;
    double s = 0.0;
    for (long long j = ptr[t]; j < ptr[t + 1]; ++j) s += in[src[j]];
    out[t] = s;
}

// y = A x for a CSR matrix (FP64 values, int32 column indices): 32 lanes per
// row, each lane striding over the row's nonzeros (coalesced reads of
// indices and values), then a 32-lane shuffle reduction.  HBM bound: 12
// bytes per nonzero plus the gathered x.  The global condensed systems have
// tens to hundreds of nonzeros per row.
constexpr int SPMV_LPR = 32;
__global__ __launch_bounds__(WG) void csr_spmv_kernel(long long nrows, const long long *indptr,
                                                     const int *indices, const double *vals,
                                                     const double *x, double *y)
{
    long long row = (long long)blockIdx.x * (WG / SPMV_LPR) + threadIdx.x / SPMV_LPR;
    const int sub = threadIdx.x % SPMV_LPR;
    double s = 0.0;
    if (row < nrows) {
        for (long long j = indptr[row] + sub; j < indptr[row + 1]; j += SPMV_LPR)
            s += vals[j] * x[indices[j]];
    }
    for (int off = SPMV_LPR / 2; off > 0; off >>= 1) s += __shfl_down(s, off, SPMV_LPR);
    if (row < nrows && sub == 0) y[row] = s;
}

// ------------------------------------------------------------ host state

// Per-element entries: each host thread that calls them gets its own context
// (stream, device staging, pinned host staging), created on its first call
// after hp3d_gpu_init, so the calls of an OpenMP element loop overlap on the
// GPU.  hp3d_gpu_finalize frees every context; a thread's context from before
// a finalize is recognised by its generation and replaced on its next call.
struct Ctx {
    hipStream_t stream = nullptr;
    // one device buffer and its pinned host mirror, laid out so that one copy
    // carries a call's inputs and one copy its outputs:
    //   [0, IO_XA)            geometry dofs and coef, right-aligned at IO_XA
    //   [IO_XA, +256)         header: status[0], codes from int 16 on
    //   [IO_XA + 256, ...)    outputs (K and load, or Aii, Bi, ASchur, BSchur)
    char *d_io = nullptr, *h_io = nullptr;
    double *d_k = nullptr;  // K and load of the two-pass path; stc_fwd's input
    double *h_k = nullptr;  // pinned staging of stc_fwd's input
    double *d_ws = nullptr;
    long long *d_off = nullptr; // zeros
};
constexpr size_t IO_XA = ((3 * (size_t)MAXN_H1 + 1) * sizeof(double) + 255) / 256 * 256;
constexpr size_t IO_HDR = 256;
constexpr int IO_CODES = 16; // int offset of the codes in the header

struct Lib {
    std::mutex mu;          // init, finalize and context creation
    int device = -1;
    int max_ndof = 0;
    int nbm = 0;            // largest bubble block a context's workspace holds
    unsigned gen = 1;       // bumped by every finalize
    std::vector<Ctx *> ctxs;
};
static Lib g;
static thread_local Ctx *t_ctx = nullptr;
static thread_local unsigned t_gen = 0;
static thread_local char t_err[256] = {0}; // hp3d_gpu_last_error: per calling thread

static int hip_fail(hipError_t r, const char *what)
{
    snprintf(t_err, sizeof(t_err), "%s: %s", what, hipGetErrorString(r));
    return HP3D_E_HIP;
}
#define HIPCHK(x)                                     \
    do {                                              \
        hipError_t _r = (x);                          \
        if (_r != hipSuccess) return hip_fail(_r, #x); \
    } while (0)

// launch options (hp3d_gpu_set_option): read when a launch is made, so a
// captured graph keeps the form it was captured with
static int g_persistent = 0;

static int nblocks_for(int max_ndof)
{
    int t = (max_ndof + BR - 1) / BR;
    return t * (t + 1) / 2;
}

template <int R>
static int launch_single_r(int space, int precision, const ElemArgs &args, const CondArgs &cd,
                           hipStream_t st)
{
    dim3 block(WG);
    if (precision == HP3D_BF16) {
        dim3 grid(args.nel);
        if (space == HP3D_H1) hipLaunchKernelGGL((elem_single_kernel<float, 0, R, true>), grid, block, 0, st, args, cd);
        else hipLaunchKernelGGL((elem_single_kernel<float, 1, R, true>), grid, block, 0, st, args, cd);
    } else if (g_persistent) {
        int dev = 0, ncu = 0;
        HIPCHK(hipGetDevice(&dev));
        HIPCHK(hipDeviceGetAttribute(&ncu, hipDeviceAttributeMultiprocessorCount, dev));
        dim3 grid(std::min(args.nel, (R == BR ? 2 : 3) * std::max(ncu, 1)));
        const SingleArgs sa{args, cd};
        if (space == HP3D_H1) {
            if (precision == HP3D_FP64) hipLaunchKernelGGL((elem_single_persistent<double, 0, R>), grid, block, 0, st, sa);
            else hipLaunchKernelGGL((elem_single_persistent<float, 0, R>), grid, block, 0, st, sa);
        } else {
            if (precision == HP3D_FP64) hipLaunchKernelGGL((elem_single_persistent<double, 1, R>), grid, block, 0, st, sa);
            else hipLaunchKernelGGL((elem_single_persistent<float, 1, R>), grid, block, 0, st, sa);
        }
    } else {
        dim3 grid(args.nel);
        if (space == HP3D_H1) {
            if (precision == HP3D_FP64) hipLaunchKernelGGL((elem_single_kernel<double, 0, R>), grid, block, 0, st, args, cd);
            else hipLaunchKernelGGL((elem_single_kernel<float, 0, R>), grid, block, 0, st, args, cd);
        } else {
            if (precision == HP3D_FP64) hipLaunchKernelGGL((elem_single_kernel<double, 1, R>), grid, block, 0, st, args, cd);
            else hipLaunchKernelGGL((elem_single_kernel<float, 1, R>), grid, block, 0, st, args, cd);
        }
    }
    HIPCHK(hipGetLastError());
    return 0;
}

// the single-block kernel for a batch whose elements have at most max_ndof
// dofs and (when condensing) at most max_nbubble bubbles: 64 rows when they fit
static int launch_single(int space, int precision, const ElemArgs &args, const CondArgs &cd,
                         int max_ndof, int max_nbubble, hipStream_t st)
{
    if (max_ndof <= BR / 2 && (cd.aii == nullptr || max_nbubble <= NBF / 2))
        return launch_single_r<BR / 2>(space, precision, args, cd, st);
    return launch_single_r<BR>(space, precision, args, cd, st);
}

static int launch_elem(int space, int precision, const ElemArgs &args, int max_ndof,
                       hipStream_t st)
{
    if (max_ndof <= BR) {
        CondArgs cd{nullptr, nullptr, nullptr, nullptr, nullptr, nullptr};
        return launch_single(space, precision, args, cd, max_ndof, 0, st);
    }
    if (!args.gws) return HP3D_E_ARG; // the block-tiled path needs the geometry workspace
    if (space == HP3D_H1) hipLaunchKernelGGL((geometry_kernel<0>), dim3(args.nel), dim3(WG), 0, st, args);
    else hipLaunchKernelGGL((geometry_kernel<1>), dim3(args.nel), dim3(WG), 0, st, args);
    dim3 grid(args.nel, nblocks_for(max_ndof)), block(WGE);
    if (space == HP3D_H1) {
        if (precision == HP3D_FP64) hipLaunchKernelGGL((elem_kernel<double, 0>), grid, block, 0, st, args);
        else if (precision == HP3D_BF16) hipLaunchKernelGGL((elem_kernel<float, 0, true>), grid, block, 0, st, args);
        else hipLaunchKernelGGL((elem_kernel<float, 0>), grid, block, 0, st, args);
    } else {
        if (precision == HP3D_FP64) hipLaunchKernelGGL((elem_kernel<double, 1>), grid, block, 0, st, args);
        else if (precision == HP3D_BF16) hipLaunchKernelGGL((elem_kernel<float, 1, true>), grid, block, 0, st, args);
        else hipLaunchKernelGGL((elem_kernel<float, 1>), grid, block, 0, st, args);
    }
    HIPCHK(hipGetLastError());
    return 0;
}

// ------------------------------------------------ allocation-free scratch

// Device scratch of hp3d_gpu_celem_batch_noalloc: the workspace of its two-pass
// bins (per-element geometry slabs, then the condensation slabs), one copy per
// device, held from module load until the process ends.  In host code the
// array's name is only a stand-in: kernels get the address that
// hipGetSymbolAddress returns for the current device (scratch_dev).
constexpr size_t SCRATCH_BYTES = HP3D_SCRATCH_BYTES;
static __device__ __align__(256) double scratch[SCRATCH_BYTES / sizeof(double)];

constexpr int MAX_DEVICES = 64;
static std::mutex g_scratch_mu;
static double *g_scratch[MAX_DEVICES] = {nullptr};

// the current device's copy of `scratch`, resolved once per device
static int scratch_dev(double **ptr)
{
    int dev = -1;
    HIPCHK(hipGetDevice(&dev));
    if (dev < 0 || dev >= MAX_DEVICES) {
        snprintf(t_err, sizeof(t_err), "device %d: no scratch slot", dev);
        return HP3D_E_ARG;
    }
    std::lock_guard<std::mutex> lk(g_scratch_mu);
    if (!g_scratch[dev]) {
        void *q = nullptr;
        HIPCHK(hipGetSymbolAddress(&q, HIP_SYMBOL(scratch)));
        if (!q) return HP3D_E_HIP;
        g_scratch[dev] = (double *)q;
    }
    *ptr = g_scratch[dev];
    return 0;
}

static int max_bubbles(int space, int max_ndof)
{
    // the bubble block is bounded by the interior space of order MAXP but also
    // by the element size
    int nb = space == HP3D_H1 ? (MAXP - 1) * (MAXP - 1) * (MAXP - 1) : 3 * MAXP * (MAXP - 1) * (MAXP - 1);
    return nb < max_ndof ? nb : max_ndof;
}

} // namespace hp3d

using namespace hp3d;

// ================================================================== C ABI

extern "C" {

const char *hp3d_gpu_last_error(void) { return t_err; }

int hp3d_gpu_set_option(int option, int value)
{
    switch (option) {
    case HP3D_OPT_PERSISTENT: g_persistent = value ? 1 : 0; return 0;
    default: return HP3D_E_ARG;
    }
}

int hp3d_gpu_ndof(int space, const int norder[19], int *nrdof, int *nrdofb)
{
    if (!norder || (space != HP3D_H1 && space != HP3D_HCURL)) return HP3D_E_ARG;
    int zero[18] = {0};
    ElemOrders o;
    if (!decode_orders(norder, zero, zero + 12, o)) return HP3D_E_ORDER;
    int n, nb;
    dof_counts(space, o, n, nb);
    if (nrdof) *nrdof = n;
    if (nrdofb) *nrdofb = nb;
    return 0;
}

static size_t io_bytes() { return IO_XA + IO_HDR + ((size_t)g.max_ndof * g.max_ndof + 2 * g.max_ndof) * sizeof(double); }

static void ctx_free(Ctx *c)
{
    if (c->stream) hipStreamSynchronize(c->stream);
    void *dev[] = {c->d_io, c->d_k, c->d_ws, c->d_off};
    for (void *p : dev)
        if (p) hipFree(p);
    void *host[] = {c->h_io, c->h_k};
    for (void *p : host)
        if (p) hipHostFree(p);
    if (c->stream) hipStreamDestroy(c->stream);
    delete c;
}

// allocate a context on the library's device (caller holds g.mu)
static int ctx_alloc(Ctx *c)
{
    HIPCHK(hipSetDevice(g.device));
    HIPCHK(hipStreamCreateWithFlags(&c->stream, hipStreamNonBlocking));
    const size_t kb = ((size_t)g.max_ndof * g.max_ndof + g.max_ndof) * sizeof(double);
    HIPCHK(hipMalloc(&c->d_io, io_bytes()));
    HIPCHK(hipMalloc(&c->d_k, kb));
    HIPCHK(hipMalloc(&c->d_ws, std::max(stc_slab(rnd16(g.nbm)), geo_slab(HP3D_HCURL)) *
                                   sizeof(double)));
    HIPCHK(hipMalloc(&c->d_off, sizeof(long long)));
    HIPCHK(hipHostMalloc(&c->h_io, io_bytes()));
    HIPCHK(hipHostMalloc(&c->h_k, kb));
    HIPCHK(hipMemsetAsync(c->d_off, 0, sizeof(long long), c->stream));
    HIPCHK(hipStreamSynchronize(c->stream));
    return 0;
}

// the calling thread's context, created on its first call after hp3d_gpu_init
static int ctx_get(Ctx **out)
{
    if (t_ctx && t_gen == g.gen) {
        *out = t_ctx;
        return 0;
    }
    std::lock_guard<std::mutex> lk(g.mu);
    if (g.device < 0) {
        snprintf(t_err, sizeof(t_err), "hp3d_gpu_init not called");
        return HP3D_E_INIT;
    }
    Ctx *c = new Ctx;
    if (int r = ctx_alloc(c)) {
        ctx_free(c);
        return r;
    }
    g.ctxs.push_back(c);
    t_ctx = c;
    t_gen = g.gen;
    *out = c;
    return 0;
}

int hp3d_gpu_init(int device, int max_ndof)
{
    {
        std::lock_guard<std::mutex> lk(g.mu);
        if (g.device < 0) {
            int cnt = 0;
            if (hipGetDeviceCount(&cnt) != hipSuccess || cnt <= device || device < 0) {
                snprintf(t_err, sizeof(t_err), "no HIP device %d", device);
                return HP3D_E_INIT;
            }
            HIPCHK(hipSetDevice(device));
            double *scr = nullptr;
            if (int r = scratch_dev(&scr)) return r;
            if (max_ndof <= 0) max_ndof = MAXN_E;
            g.max_ndof = max_ndof;
            g.nbm = max_bubbles(HP3D_HCURL, max_ndof);
            g.device = device;
        }
    }
    // the calling thread's context now, so that allocation failures show here
    Ctx *c;
    return ctx_get(&c);
}

int hp3d_gpu_finalize(void)
{
    std::lock_guard<std::mutex> lk(g.mu);
    if (g.device < 0) return 0;
    for (Ctx *c : g.ctxs) ctx_free(c);
    g.ctxs.clear();
    ++g.gen;
    g.device = -1;
    return 0;
}

int hp3d_gpu_elem_batch(int space, int precision, int nel, const int *norder, const int *nori_e,
                        const int *nori_f, const double *xnod, int ngeo, const double *coef,
                        int load, const long long *off_k, int ld_k, void *kout,
                        const long long *off_v, double *bout, void *workspace, int *status,
                        int max_ndof, void *stream)
{
    if ((space != HP3D_H1 && space != HP3D_HCURL) || (precision != HP3D_FP64 && precision != HP3D_FP32 && precision != HP3D_BF16))
        return HP3D_E_ARG;
    if (nel < 0 || ngeo < 1 || ngeo > MAXN_H1 || max_ndof < 1 || max_ndof > MAXN_E || ld_k < 0)
        return HP3D_E_ARG;
    if (nel == 0) return 0;
    if (!norder || !nori_e || !nori_f || !xnod || !coef || !off_k || !kout) return HP3D_E_ARG;
    if (!workspace && hp3d_gpu_elem_workspace(space, precision, nel, max_ndof) > 0) return HP3D_E_ARG;
    if (bout && !off_v) return HP3D_E_ARG;
    if (ld_k > 0 && ld_k < max_ndof) return HP3D_E_ARG;
    ElemArgs a;
    a.space = space;
    a.nel = nel;
    a.ngeo = ngeo;
    a.load = load;
    a.ld_k = ld_k;
    a.max_ndof = max_ndof;
    a.norder = norder;
    a.nori_e = nori_e;
    a.nori_f = nori_f;
    a.xnod = xnod;
    a.coef = coef;
    a.off_k = off_k;
    a.off_v = off_v;
    a.kout = kout;
    a.bout = bout;
    a.status = status;
    a.gws = (double *)workspace;
    return launch_elem(space, precision, a, max_ndof, (hipStream_t)stream);
}

size_t hp3d_gpu_elem_workspace(int space, int precision, int nel, int max_ndof)
{
    if (nel <= 0 || max_ndof <= 0 || (space != HP3D_H1 && space != HP3D_HCURL)) return 0;
    if (max_ndof <= BR) return 0; // single-block kernel
    return (size_t)nel * geo_slab(space) * sizeof(double);
}

size_t hp3d_gpu_stc_workspace(int nel, int max_ndof, int max_nbubble)
{
    if (nel <= 0 || max_nbubble <= 0 || max_ndof <= 0) return 0;
    return (size_t)nel * stc_slab(rnd16(max_nbubble)) * sizeof(double);
}

size_t hp3d_gpu_celem_workspace(int space, int precision, int nel, int max_ndof, int max_nbubble)
{
    if (max_ndof <= BR && max_nbubble <= NBF) return 0; // fused
    // the geometry pass and the condensation use the same memory one after the other
    return std::max(hp3d_gpu_elem_workspace(space, precision, nel, max_ndof),
                    hp3d_gpu_stc_workspace(nel, max_ndof, max_nbubble));
}

int hp3d_gpu_stc_batch(int space, int precision, int nel, const int *norder,
                       const long long *off_k, int ld_k, const void *kin,
                       const long long *off_v, const double *bin, const long long *off_c,
                       int ld_c, double *aii, double *bi, const long long *off_s,
                       double *aschur, double *bschur, void *workspace, int *status,
                       int max_ndof, int max_nbubble, void *stream)
{
    if ((space != HP3D_H1 && space != HP3D_HCURL) || (precision != HP3D_FP64 && precision != HP3D_FP32 && precision != HP3D_BF16))
        return HP3D_E_ARG;
    if (nel < 0 || max_ndof < 1 || max_ndof > MAXN_E || max_nbubble < 0 || max_nbubble > max_ndof ||
        ld_k < 0 || ld_c < 0)
        return HP3D_E_ARG;
    if (nel == 0) return 0;
    if (!norder || !off_k || !kin || !off_v || !bin || !off_c || !aii || !bi || !off_s ||
        !aschur || !bschur)
        return HP3D_E_ARG;
    if (hp3d_gpu_stc_workspace(nel, max_ndof, max_nbubble) > 0 && !workspace) return HP3D_E_ARG;
    if (precision == HP3D_FP32 && (const void *)aii == kin) return HP3D_E_ARG;
    StcArgs a;
    a.space = space;
    a.nel = nel;
    a.ld_k = ld_k;
    a.ld_c = ld_c;
    a.nbp_max = rnd16(max_nbubble);
    a.n_max = max_ndof;
    a.nb_max = max_nbubble;
    a.norder = norder;
    a.nsz = nullptr;
    a.off_k = off_k;
    a.off_v = off_v;
    a.off_c = off_c;
    a.off_s = off_s;
    a.kin = kin;
    a.bin = bin;
    a.aii = aii;
    a.bi = bi;
    a.aschur = aschur;
    a.bschur = bschur;
    a.ws = (double *)workspace;
    a.status = status;
    hipStream_t st = (hipStream_t)stream;
    if (precision == HP3D_FP64) launch_stc<double>(a, max_nbubble, st);
    else launch_stc<float>(a, max_nbubble, st);
    HIPCHK(hipGetLastError());
    return 0;
}

int hp3d_gpu_celem_batch(int space, int precision, int nel, const int *norder, const int *nori_e,
                         const int *nori_f, const double *xnod, int ngeo, const double *coef,
                         int load, const long long *off_k, int ld_k, void *kout,
                         const long long *off_v, double *bout, const long long *off_c,
                         double *aii, double *bi, const long long *off_s, double *aschur,
                         double *bschur, void *workspace, int *status, int max_ndof,
                         int max_nbubble, void *stream)
{
    if ((space != HP3D_H1 && space != HP3D_HCURL) || (precision != HP3D_FP64 && precision != HP3D_FP32 && precision != HP3D_BF16))
        return HP3D_E_ARG;
    if (nel < 0 || ngeo < 1 || ngeo > MAXN_H1 || max_ndof < 1 || max_ndof > MAXN_E || max_nbubble < 0 ||
        max_nbubble > max_ndof || ld_k < 0)
        return HP3D_E_ARG;
    if (nel == 0) return 0;
    if (!norder || !nori_e || !nori_f || !xnod || !coef || !off_v || !off_c || !aii || !bi ||
        !off_s || !aschur || !bschur)
        return HP3D_E_ARG;
    if (kout && !off_k) return HP3D_E_ARG;
    if (ld_k > 0 && ld_k < max_ndof) return HP3D_E_ARG;
    ElemArgs a;
    a.space = space;
    a.nel = nel;
    a.ngeo = ngeo;
    a.load = load;
    a.ld_k = ld_k;
    a.max_ndof = max_ndof;
    a.norder = norder;
    a.nori_e = nori_e;
    a.nori_f = nori_f;
    a.xnod = xnod;
    a.coef = coef;
    a.off_k = off_k;
    a.off_v = off_v;
    a.kout = kout;
    a.bout = bout;
    a.status = status;
    a.gws = (double *)workspace;
    hipStream_t st = (hipStream_t)stream;
    if (max_ndof <= BR && max_nbubble <= NBF) {
        CondArgs cd{off_c, off_s, aii, bi, aschur, bschur};
        return launch_single(space, precision, a, cd, max_ndof, max_nbubble, st);
    }
    // two passes: element matrices, then condensation from HBM
    if (!kout || !bout) return HP3D_E_ARG;
    int r = launch_elem(space, precision, a, max_ndof, st);
    if (r) return r;
    return hp3d_gpu_stc_batch(space, precision, nel, norder, off_k, ld_k, kout, off_v, bout, off_c,
                              0, aii, bi, off_s, aschur, bschur, workspace, status, max_ndof,
                              max_nbubble, stream);
}

int hp3d_gpu_celem_batch_noalloc(int space, int precision, int nel, const int *norder,
                                 const int *nori_e, const int *nori_f, const double *xnod, int ngeo,
                                 const double *coef, int load, const long long *off_k, int ld_k,
                                 void *kout, const long long *off_v, double *bout,
                                 const long long *off_c, double *aii, double *bi,
                                 const long long *off_s, double *aschur, double *bschur,
                                 int *status, int max_ndof, int max_nbubble, void *stream)
{
    if (nel <= 0 || max_ndof < 1 || max_ndof > MAXN_E || max_nbubble < 0 || max_nbubble > max_ndof)
        return hp3d_gpu_celem_batch(space, precision, nel, norder, nori_e, nori_f, xnod, ngeo, coef,
                                    load, off_k, ld_k, kout, off_v, bout, off_c, aii, bi, off_s,
                                    aschur, bschur, nullptr, status, max_ndof, max_nbubble, stream);
    size_t per = hp3d_gpu_celem_workspace(space, precision, 1, max_ndof, max_nbubble);
    if (per == 0) // fused path: no workspace
        return hp3d_gpu_celem_batch(space, precision, nel, norder, nori_e, nori_f, xnod, ngeo, coef,
                                    load, off_k, ld_k, kout, off_v, bout, off_c, aii, bi, off_s,
                                    aschur, bschur, nullptr, status, max_ndof, max_nbubble, stream);
    if (per > SCRATCH_BYTES) {
        snprintf(t_err, sizeof(t_err),
                 "celem_batch_noalloc: %zu bytes of workspace per element (max_ndof %d, "
                 "max_nbubble %d) exceed the %zu-byte scratch",
                 per, max_ndof, max_nbubble, SCRATCH_BYTES);
        return HP3D_E_ARG;
    }
    if (!norder || !nori_e || !nori_f || !xnod || !coef) return HP3D_E_ARG;
    double *ws = nullptr;
    if (int r = scratch_dev(&ws)) return r;
    // elements in chunks whose workspace fits the scratch, one after the other
    // on `stream`; the offset arrays index the caller's outputs directly
    const int chunk = (int)std::min<size_t>(SCRATCH_BYTES / per, (size_t)nel);
    for (int c0 = 0; c0 < nel; c0 += chunk) {
        const int cn = std::min(chunk, nel - c0);
        int r = hp3d_gpu_celem_batch(space, precision, cn, norder + 19 * (size_t)c0,
                                     nori_e + 12 * (size_t)c0, nori_f + 6 * (size_t)c0,
                                     xnod + 3 * (size_t)ngeo * c0, ngeo, coef + c0, load,
                                     off_k ? off_k + c0 : nullptr, ld_k, kout,
                                     off_v ? off_v + c0 : nullptr, bout,
                                     off_c ? off_c + c0 : nullptr, aii, bi,
                                     off_s ? off_s + c0 : nullptr, aschur, bschur, ws,
                                     status ? status + c0 : nullptr, max_ndof, max_nbubble, stream);
        if (r) return r;
    }
    return 0;
}

int hp3d_gpu_stc_bwd_batch(int space, int nel, const int *norder, const long long *off_s,
                           const double *aschur, const long long *off_v, const double *bschur,
                           double *x, void *stream)
{
    if ((space != HP3D_H1 && space != HP3D_HCURL) || nel < 0) return HP3D_E_ARG;
    if (nel == 0) return 0;
    if (!norder || !off_s || !aschur || !off_v || !bschur || !x) return HP3D_E_ARG;
    BwdArgs a{space, nel, norder, off_s, off_v, aschur, bschur, x};
    hipLaunchKernelGGL(stc_bwd_kernel, dim3(nel), dim3(WG), 0, (hipStream_t)stream, a);
    HIPCHK(hipGetLastError());
    return 0;
}

int hp3d_gpu_dirichlet_lift(int space, int nel, const int *norder, const long long *off_c,
                            const double *aii, const long long *off_v, const double *ud, double *bi,
                            void *stream)
{
    if ((space != HP3D_H1 && space != HP3D_HCURL) || nel < 0) return HP3D_E_ARG;
    if (nel == 0) return 0;
    if (!norder || !off_c || !aii || !off_v || !ud || !bi) return HP3D_E_ARG;
    LiftArgs a{space, nel, norder, off_c, off_v, aii, ud, bi};
    hipLaunchKernelGGL(dirichlet_lift_kernel, dim3(nel), dim3(WG), 0, (hipStream_t)stream, a);
    HIPCHK(hipGetLastError());
    return 0;
}

int hp3d_gpu_shape(int space, const int norder[19], const int nori_e[12], const int nori_f[6],
                   int npts, const double *xi, double *val, double *grad)
{
    if ((space != HP3D_H1 && space != HP3D_HCURL) || npts < 0) return HP3D_E_ARG;
    if (!norder || !nori_e || !nori_f || (npts > 0 && (!xi || !val || !grad))) return HP3D_E_ARG;
    ElemOrders o;
    if (!decode_orders(norder, nori_e, nori_f, o)) return HP3D_E_ORDER;
    int n, nb;
    dof_counts(space, o, n, nb);
    for (int p = 0; p < npts; ++p) {
        double tv[3][NTAB], td[3][NTAB];
        for (int d = 0; d < 3; ++d) table_at(xi[3 * p + d], tv[d], td[d]);
        for (int k = 0; k < n; ++k) {
            const unsigned dsc = space == HP3D_H1 ? desc_h1(k, o) : desc_hcurl(k, o);
            const int i0 = desc_ix(dsc, 0), i1 = desc_ix(dsc, 1), i2 = desc_ix(dsc, 2);
            const double v0 = tv[0][i0], v1 = tv[1][i1], v2 = tv[2][i2];
            const double gr[3] = {td[0][i0] * v1 * v2, v0 * td[1][i1] * v2, v0 * v1 * td[2][i2]};
            const size_t pk = (size_t)p * n + k;
            if (space == HP3D_H1) {
                val[pk] = v0 * v1 * v2;
                for (int c = 0; c < 3; ++c) grad[3 * pk + c] = gr[c];
            } else {
                // one component cp of sign sg: E = sg phi e_cp, curl E = sg grad phi x e_cp
                const int cp = desc_comp(dsc);
                const double sg = desc_neg(dsc) ? -1.0 : 1.0;
                for (int c = 0; c < 3; ++c) val[3 * pk + c] = c == cp ? sg * v0 * v1 * v2 : 0.0;
                grad[3 * pk + 0] = cp == 0 ? 0.0 : (cp == 1 ? -sg * gr[2] : sg * gr[1]);
                grad[3 * pk + 1] = cp == 0 ? sg * gr[2] : (cp == 1 ? 0.0 : -sg * gr[0]);
                grad[3 * pk + 2] = cp == 0 ? -sg * gr[1] : (cp == 1 ? sg * gr[0] : 0.0);
            }
        }
    }
    return 0;
}

int hp3d_gpu_field_batch(int space, int mode, int nel, const int *norder, const int *nori_e,
                         const int *nori_f, const double *xnod, int ngeo, const double *coef,
                         const long long *off_v, const double *u, const double *v, int exact,
                         double *out, double *r, int *status, void *stream)
{
    if ((space != HP3D_H1 && space != HP3D_HCURL) || nel < 0 || ngeo < 1 || ngeo > MAXN_H1)
        return HP3D_E_ARG;
    if (mode != HP3D_FIELD_ERROR && mode != HP3D_FIELD_ADJOINT && mode != HP3D_FIELD_L2)
        return HP3D_E_ARG;
    if (nel == 0) return 0;
    if (!norder || !nori_e || !nori_f || !xnod || !off_v || !u || !out) return HP3D_E_ARG;
    if (mode == HP3D_FIELD_ADJOINT && (!v || !coef)) return HP3D_E_ARG;
    FieldArgs a{space, mode, nel, ngeo, exact, norder, nori_e, nori_f, xnod, coef, off_v, u, v,
                out, r, status};
    hipStream_t st = (hipStream_t)stream;
    if (space == HP3D_H1) hipLaunchKernelGGL(field_kernel<0>, dim3(nel), dim3(WG), 0, st, a);
    else hipLaunchKernelGGL(field_kernel<1>, dim3(nel), dim3(WG), 0, st, a);
    HIPCHK(hipGetLastError());
    return 0;
}

int hp3d_gpu_gather_sum(long long nout, const long long *ptr, const long long *src,
                        const double *in, double *out, void *stream)
{
    if (nout < 0) return HP3D_E_ARG;
    if (nout == 0) return 0;
    if (!ptr || !src || !in || !out) return HP3D_E_ARG;
    long long nblk = (nout + WG - 1) / WG;
    if (nblk > 0x7fffffffLL) return HP3D_E_ARG;
    hipLaunchKernelGGL(gather_sum_kernel, dim3((unsigned)nblk), dim3(WG), 0, (hipStream_t)stream,
                       nout, ptr, src, in, out);
    HIPCHK(hipGetLastError());
    return 0;
}

int hp3d_gpu_csr_spmv(long long nrows, const long long *indptr, const int *indices,
                      const double *vals, const double *x, double *y, void *stream)
{
    if (nrows < 0) return HP3D_E_ARG;
    if (nrows == 0) return 0;
    if (!indptr || !indices || !vals || !x || !y) return HP3D_E_ARG;
    long long nblk = (nrows + WG / SPMV_LPR - 1) / (WG / SPMV_LPR);
    if (nblk > 0x7fffffffLL) return HP3D_E_ARG;
    hipLaunchKernelGGL(csr_spmv_kernel, dim3((unsigned)nblk), dim3(WG), 0, (hipStream_t)stream,
                       nrows, indptr, indices, vals, x, y);
    HIPCHK(hipGetLastError());
    return 0;
}

// ---------------------------------------------------- per-element drop-in

static int *hdr_d(Ctx &c) { return (int *)(c.d_io + IO_XA); }
static int *hdr_h(Ctx &c) { return (int *)(c.h_io + IO_XA); }
static double *out_d(Ctx &c) { return (double *)(c.d_io + IO_XA + IO_HDR); }
static double *out_h(Ctx &c) { return (double *)(c.h_io + IO_XA + IO_HDR); }

// stage one element's codes, geometry and coef (status cleared) with one copy
static int stage_element(Ctx &c, int space, const int norder[19], const int nori_e[12],
                         const int nori_f[6], const double *xnod, int ngeo, double coef, int &n,
                         int &nb, const double *&d_xnod)
{
    if (!norder || !nori_e || !nori_f || !xnod || ngeo < 1 || ngeo > MAXN_H1) return HP3D_E_ARG;
    ElemOrders o;
    if (!decode_orders(norder, nori_e, nori_f, o)) return HP3D_E_ORDER;
    dof_counts(space, o, n, nb);
    int nh, nbh;
    dof_counts(HP3D_H1, o, nh, nbh);
    if (n > g.max_ndof || ngeo > nh) return HP3D_E_ARG;
    int *h = hdr_h(c);
    h[0] = 0;
    memcpy(h + IO_CODES, norder, 19 * sizeof(int));
    memcpy(h + IO_CODES + 19, nori_e, 12 * sizeof(int));
    memcpy(h + IO_CODES + 31, nori_f, 6 * sizeof(int));
    const size_t xb = (3 * (size_t)ngeo + 1) * sizeof(double);
    double *hx = (double *)(c.h_io + IO_XA - xb);
    memcpy(hx, xnod, 3 * (size_t)ngeo * sizeof(double));
    hx[3 * ngeo] = coef;
    HIPCHK(hipMemcpyAsync(c.d_io + IO_XA - xb, c.h_io + IO_XA - xb, xb + IO_HDR,
                          hipMemcpyHostToDevice, c.stream));
    d_xnod = (const double *)(c.d_io + IO_XA - xb);
    return 0;
}

// one copy back of the header and the first `nout` doubles of the outputs
static int fetch_outputs(Ctx &c, size_t nout)
{
    HIPCHK(hipMemcpyAsync(c.h_io + IO_XA, c.d_io + IO_XA, IO_HDR + nout * sizeof(double),
                          hipMemcpyDeviceToHost, c.stream));
    HIPCHK(hipStreamSynchronize(c.stream));
    return 0;
}

static void put_matrix(double *dst, int ld, const double *src, int rows, int cols)
{
    for (int j = 0; j < cols; ++j) memcpy(dst + (size_t)j * ld, src + (size_t)j * rows, rows * sizeof(double));
}

// K and load of the staged element into the outputs (K packed, then the load)
static int elem_staged(Ctx &c, int space, int n, int ngeo, const double *d_xnod, int load)
{
    const int *codes = hdr_d(c) + IO_CODES;
    return hp3d_gpu_elem_batch(space, HP3D_FP64, 1, codes, codes + 19, codes + 31, d_xnod, ngeo,
                               d_xnod + 3 * ngeo, load, c.d_off, 0, out_d(c), c.d_off,
                               out_d(c) + (size_t)n * n, c.d_ws, hdr_d(c), n, c.stream);
}

static int copy_elem_outputs(Ctx &c, int n, double *aloc, int ld_aloc, double *bloc)
{
    if (int r = fetch_outputs(c, (size_t)n * n + n)) return r;
    put_matrix(aloc, ld_aloc, out_h(c), n, n);
    memcpy(bloc, out_h(c) + (size_t)n * n, n * sizeof(double));
    return hdr_h(c)[0];
}

// the condensed system, packed in the outputs as Aii (ni x ni), Bi, ASchur
// (nb x ni), BSchur, into the caller's ALOC/BLOC and CLOC arrays
static int copy_condensed(Ctx &c, int ni, int nb, double *aloc, int ld_aloc, double *bloc,
                          double *aschur, double *bschur)
{
    const size_t ai = (size_t)ni * ni;
    if (int r = fetch_outputs(c, ai + ni + (size_t)nb * ni + nb)) return r;
    const double *o = out_h(c);
    put_matrix(aloc, ld_aloc, o, ni, ni);
    memcpy(bloc, o + ai, ni * sizeof(double));
    if (aschur) memcpy(aschur, o + ai + ni, (size_t)nb * ni * sizeof(double));
    if (bschur) memcpy(bschur, o + ai + ni + (size_t)nb * ni, nb * sizeof(double));
    return hdr_h(c)[0];
}

int hp3d_gpu_elem(int space, const int norder[19], const int nori_e[12], const int nori_f[6],
                  const double *xnod, int ngeo, double coef, int load, double *aloc,
                  int ld_aloc, double *bloc)
{
    if (space != HP3D_H1 && space != HP3D_HCURL) return HP3D_E_ARG;
    Ctx *cp;
    if (int rc = ctx_get(&cp)) return rc;
    Ctx &c = *cp;
    int n, nb;
    const double *d_xnod;
    int r = stage_element(c, space, norder, nori_e, nori_f, xnod, ngeo, coef, n, nb, d_xnod);
    if (r) return r;
    if (!aloc || !bloc || ld_aloc < n) return HP3D_E_ARG;
    if ((r = elem_staged(c, space, n, ngeo, d_xnod, load))) return r;
    return copy_elem_outputs(c, n, aloc, ld_aloc, bloc);
}

int hp3d_gpu_stc_fwd(int ni, int nb, double *aloc, int ld_aloc, double *bloc, double *aschur,
                     double *bschur)
{
    Ctx *cp;
    if (int rc = ctx_get(&cp)) return rc;
    Ctx &c = *cp;
    int n = ni + nb;
    if (ni < 0 || nb < 0 || n > g.max_ndof || !aloc || !bloc || ld_aloc < n) return HP3D_E_ARG;
    if (nb == 0) return 0;
    if (ni == 0) return HP3D_E_NOIFACE;
    if (nb > max_bubbles(HP3D_HCURL, g.max_ndof)) return HP3D_E_ARG;
    // ALOC (ld_aloc) packed and BLOC behind it: one copy; the header (status
    // cleared, sizes n and nb): a second
    for (int j = 0; j < n; ++j) memcpy(c.h_k + (size_t)j * n, aloc + (size_t)j * ld_aloc, n * sizeof(double));
    memcpy(c.h_k + (size_t)n * n, bloc, n * sizeof(double));
    int *h = hdr_h(c);
    h[0] = 0;
    h[IO_CODES] = n;
    h[IO_CODES + 1] = nb;
    HIPCHK(hipMemcpyAsync(c.d_k, c.h_k, ((size_t)n * n + n) * sizeof(double), hipMemcpyHostToDevice, c.stream));
    HIPCHK(hipMemcpyAsync(c.d_io + IO_XA, c.h_io + IO_XA, IO_HDR, hipMemcpyHostToDevice, c.stream));
    StcArgs a;
    a.space = HP3D_H1;
    a.nel = 1;
    a.ld_k = 0;
    a.ld_c = 0;
    a.nbp_max = rnd16(nb);
    a.n_max = n;
    a.nb_max = nb;
    a.norder = nullptr;
    a.nsz = hdr_d(c) + IO_CODES;
    a.off_k = a.off_v = a.off_c = a.off_s = c.d_off;
    a.kin = c.d_k;
    a.bin = c.d_k + (size_t)n * n;
    const size_t ai = (size_t)ni * ni;
    a.aii = out_d(c);
    a.bi = out_d(c) + ai;
    a.aschur = out_d(c) + ai + ni;
    a.bschur = out_d(c) + ai + ni + (size_t)nb * ni;
    a.ws = c.d_ws;
    a.status = hdr_d(c);
    launch_stc<double>(a, nb, c.stream);
    HIPCHK(hipGetLastError());
    return copy_condensed(c, ni, nb, aloc, ld_aloc, bloc, aschur, bschur);
}

int hp3d_gpu_celem(int space, const int norder[19], const int nori_e[12], const int nori_f[6],
                   const double *xnod, int ngeo, double coef, int load, double *aloc,
                   int ld_aloc, double *bloc, double *aschur, double *bschur)
{
    if (space != HP3D_H1 && space != HP3D_HCURL) return HP3D_E_ARG;
    Ctx *cp;
    if (int rc = ctx_get(&cp)) return rc;
    Ctx &c = *cp;
    int n, nb;
    const double *d_xnod;
    int r = stage_element(c, space, norder, nori_e, nori_f, xnod, ngeo, coef, n, nb, d_xnod);
    if (r) return r;
    if (!aloc || !bloc || ld_aloc < n) return HP3D_E_ARG;
    if (nb == 0) {
        if ((r = elem_staged(c, space, n, ngeo, d_xnod, load))) return r;
        return copy_elem_outputs(c, n, aloc, ld_aloc, bloc);
    }
    const int ni = n - nb;
    if (ni == 0) return HP3D_E_NOIFACE;
    // the batched entry on a batch of one: the fused kernel for N <= 128 and
    // nb <= 32, else the element pass and the strip condensation
    const int *codes = hdr_d(c) + IO_CODES;
    const size_t ai = (size_t)ni * ni;
    double *o = out_d(c);
    r = hp3d_gpu_celem_batch(space, HP3D_FP64, 1, codes, codes + 19, codes + 31, d_xnod, ngeo,
                             d_xnod + 3 * ngeo, load, c.d_off, 0, c.d_k, c.d_off,
                             c.d_k + (size_t)n * n, c.d_off, o, o + ai, c.d_off, o + ai + ni,
                             o + ai + ni + (size_t)nb * ni, c.d_ws, hdr_d(c), n, nb, c.stream);
    if (r) return r;
    return copy_condensed(c, ni, nb, aloc, ld_aloc, bloc, aschur, bschur);
}

} // extern "C"
