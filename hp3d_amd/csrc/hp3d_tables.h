// hp3d_tables.h -- hexahedron topology and quadrature tables shared by the
// device kernels and the host entry points of libhp3d_gpu.
//
// The tables encode hp3D's element conventions, which a drop-in has to keep
// identical (vertex/edge/face numbering, orientation codes, Gauss tables):
//   vertices  element_data / BlendHexaV       src/element/shape_1/BlendProject.F90:197-253
//   edges     BlendProjectHexaE               BlendProject.F90:255-363
//   faces     BlendProjectHexaF               BlendProject.F90:365-439
//   face orientation codes (OrientQuad)       src/element/shape_1/Orient.F90:39-116
//   NFAXES(3,:) (axes swapped)                src/modules/element_data.F90:246-249
//   1D Gauss nodes/weights on [-1,1]          src/element/quadrature/gauss_quadrature.F90:518-660
#pragma once

#define HP3D_HD __host__ __device__

namespace hp3d {

constexpr int MAXP = 6;
constexpr int MODORDER = 10;

// affine index (0: 1-xi, 1: xi) per direction for each vertex
HP3D_HD constexpr int vtx_mu(int v, int d)
{
    // v1..v8 = (000,100,110,010,001,101,111,011)
    return d == 0 ? ((v == 1 || v == 2 || v == 5 || v == 6) ? 1 : 0)
         : d == 1 ? ((v == 2 || v == 3 || v == 6 || v == 7) ? 1 : 0)
                  : (v >= 4 ? 1 : 0);
}

// edge e: direction of its coordinate, and the (affine index, direction) of
// its two blending factors
struct EdgeT { int dir, a1, d1, a2, d2; };
HP3D_HD constexpr EdgeT edge_t(int e)
{
    switch (e) {
    case 0: return {0, 0, 1, 0, 2};
    case 1: return {1, 1, 0, 0, 2};
    case 2: return {0, 1, 1, 0, 2};
    case 3: return {1, 0, 0, 0, 2};
    case 4: return {0, 0, 1, 1, 2};
    case 5: return {1, 1, 0, 1, 2};
    case 6: return {0, 1, 1, 1, 2};
    case 7: return {1, 0, 0, 1, 2};
    case 8: return {2, 0, 0, 0, 1};
    case 9: return {2, 1, 0, 0, 1};
    case 10: return {2, 1, 0, 1, 1};
    default: return {2, 0, 0, 1, 1};
    }
}

// face f: (affine index, direction) of its blending factor and its two local
// axes (element directions of local s and t)
struct FaceT { int a, dn, ax0, ax1; };
HP3D_HD constexpr FaceT face_t(int f)
{
    switch (f) {
    case 0: return {0, 2, 0, 1};
    case 1: return {1, 2, 0, 1};
    case 2: return {0, 1, 0, 2};
    case 3: return {1, 0, 1, 2};
    case 4: return {1, 1, 0, 2};
    default: return {0, 0, 1, 2};
    }
}

// face orientation code n in 0..7 -> which local axis becomes the face-node
// s and t axes, and whether each is reversed
struct FaceOri { int ps, pt, rs, rt; };
HP3D_HD constexpr FaceOri face_ori(int n)
{
    // (OrPa(1)-1, OrPa(2)-1, OrGS(1)==0, OrGT(1)==0)
    switch (n) {
    case 0: return {0, 1, 0, 0};
    case 1: return {1, 0, 0, 1};
    case 2: return {0, 1, 1, 1};
    case 3: return {1, 0, 1, 0};
    case 4: return {1, 0, 0, 0};
    case 5: return {0, 1, 1, 0};
    case 6: return {1, 0, 1, 1};
    default: return {0, 1, 0, 1};
    }
}

HP3D_HD inline int face_swapped(int n) { return (n == 1 || n == 3 || n == 4 || n == 6) ? 1 : 0; }

// 1D Gauss-Legendre rules on [-1,1] with the reference's tabulated digits
// (15 significant digits for n <= 9); rescaled to [0,1] as 0.5*(1+x), 0.5*w.
#define HP3D_GAUSS_X                                                                          \
    {{0.0},                                                                                   \
     {-.577350269189626, .577350269189626},                                                   \
     {-.774596669241483, .000000000000000, .774596669241483},                                 \
     {-.861136311594053, -.339981043584856, .339981043584856, .861136311594053},              \
     {-.906179845938664, -.538469310105683, .000000000000000, .538469310105683,               \
      .906179845938664},                                                                      \
     {-.932469514203152, -.661209386466265, -.238619186083197, .238619186083197,              \
      .661209386466265, .932469514203152},                                                    \
     {-.949107912342759, -.741531185599394, -.405845151377397, .000000000000000,              \
      .405845151377397, .741531185599394, .949107912342759},                                  \
     {-.960289856497536, -.796666477413627, -.525532409916329, -.183434642495650,             \
      .183434642495650, .525532409916329, .796666477413627, .960289856497536},                \
     {-.968160239507626, -.836031107326636, -.613371432700590, -.324253423403809,             \
      .000000000000000, .324253423403809, .613371432700590, .836031107326636,                 \
      .968160239507626},                                                                      \
     {-.9739065285171717200779640, -.8650633666889845107320967, -.6794095682990244062343274, \
      -.4333953941292471907992659, -.1488743389816312108848260, .1488743389816312108848260,  \
      .4333953941292471907992659, .6794095682990244062343274, .8650633666889845107320967,    \
      .9739065285171717200779640}}
#define HP3D_GAUSS_W                                                                          \
    {{2.000000000000000},                                                                     \
     {1.000000000000000, 1.000000000000000},                                                  \
     {.555555555555556, .888888888888889, .555555555555556},                                  \
     {.347854845137454, .652145154862546, .652145154862546, .347854845137454},                \
     {.236926885056189, .478628670499366, .568888888888889, .478628670499366,                 \
      .236926885056189},                                                                      \
     {.171324492379170, .360761573048139, .467913934572691, .467913934572691,                 \
      .360761573048139, .171324492379170},                                                    \
     {.129484966168870, .279705391489277, .381830050505119, .417959183673469,                 \
      .381830050505119, .279705391489277, .129484966168870},                                  \
     {.101228536290376, .222381034453374, .313706645877887, .362683783378362,                 \
      .362683783378362, .313706645877887, .222381034453374, .101228536290376},                \
     {.081274388361574, .180648160694857, .260610696402935, .312347077040003,                 \
      .330239355001260, .312347077040003, .260610696402935, .180648160694857,                 \
      .081274388361574},                                                                      \
     {.0666713443086881375935688, .1494513491505805931457763, .2190863625159820439955349,    \
      .2692667193099963550912269, .2955242247147528701738930, .2955242247147528701738930,    \
      .2692667193099963550912269, .2190863625159820439955349, .1494513491505805931457763,    \
      .0666713443086881375935688}}

// Decoded element orders and dof layout of one hexahedron.
struct ElemOrders {
    int pe[12];     // edge orders
    int fh[6], fv[6]; // face orders along the face-node s and t axes
    int b[3];       // interior orders (x, y, z)
    int oe[12], of[6];
};

HP3D_HD inline bool decode_orders(const int *norder, const int *nori_e, const int *nori_f,
                                  ElemOrders &o)
{
    bool ok = true;
    for (int e = 0; e < 12; ++e) {
        o.pe[e] = norder[e];
        o.oe[e] = nori_e[e];
        ok = ok && o.pe[e] >= 1 && o.pe[e] <= MAXP && (o.oe[e] == 0 || o.oe[e] == 1);
    }
    for (int f = 0; f < 6; ++f) {
        o.fh[f] = norder[12 + f] / MODORDER;
        o.fv[f] = norder[12 + f] % MODORDER;
        o.of[f] = nori_f[f];
        ok = ok && o.fh[f] >= 1 && o.fh[f] <= MAXP && o.fv[f] >= 1 && o.fv[f] <= MAXP &&
             o.of[f] >= 0 && o.of[f] <= 7;
    }
    o.b[0] = norder[18] / 100;
    o.b[1] = (norder[18] / 10) % 10;
    o.b[2] = norder[18] % 10;
    for (int d = 0; d < 3; ++d) ok = ok && o.b[d] >= 1 && o.b[d] <= MAXP;
    return ok;
}

// dof counts: total and bubble (space 0 = H1, 1 = H(curl))
HP3D_HD inline void dof_counts(int space, const ElemOrders &o, int &n, int &nb)
{
    if (space == 0) {
        n = 8;
        for (int e = 0; e < 12; ++e) n += o.pe[e] - 1;
        for (int f = 0; f < 6; ++f) n += (o.fh[f] - 1) * (o.fv[f] - 1);
        nb = (o.b[0] - 1) * (o.b[1] - 1) * (o.b[2] - 1);
    } else {
        n = 0;
        for (int e = 0; e < 12; ++e) n += o.pe[e];
        for (int f = 0; f < 6; ++f) n += o.fh[f] * (o.fv[f] - 1) + (o.fh[f] - 1) * o.fv[f];
        nb = o.b[0] * (o.b[1] - 1) * (o.b[2] - 1) + (o.b[0] - 1) * o.b[1] * (o.b[2] - 1) +
             (o.b[0] - 1) * (o.b[1] - 1) * o.b[2];
    }
    n += nb;
}

// Gauss points per direction of the brick rule: max order seen in each
// element direction (faces after the swap of their axes), INTEGRATION = 0.
HP3D_HD inline void quad_counts(const ElemOrders &o, int nq[3])
{
    int m[3] = {o.b[0], o.b[1], o.b[2]};
    for (int e = 0; e < 12; ++e) {
        int d = edge_t(e).dir;
        m[d] = o.pe[e] > m[d] ? o.pe[e] : m[d];
    }
    for (int f = 0; f < 6; ++f) {
        FaceT ft = face_t(f);
        int sw = face_swapped(o.of[f]);
        int o0 = sw ? o.fv[f] : o.fh[f], o1 = sw ? o.fh[f] : o.fv[f];
        m[ft.ax0] = o0 > m[ft.ax0] ? o0 : m[ft.ax0];
        m[ft.ax1] = o1 > m[ft.ax1] ? o1 : m[ft.ax1];
    }
    for (int d = 0; d < 3; ++d) nq[d] = (m[d] < MAXP ? m[d] : MAXP) + 1;
}

// 1D factor tables: index layout per direction
//   0: 1-xi, 1: xi, 2..6: phi_i(xi), 7..11: phi_i(1-xi),
//   12..17: P_i(xi), 18..23: P_i(1-xi)   (i = 0..5)
constexpr int NTAB = 24;
HP3D_HD constexpr int t_phi(int i, int flip) { return flip ? 5 + i : i; }
HP3D_HD constexpr int t_leg(int i, int flip) { return flip ? 18 + i : 12 + i; }

// dof descriptor: table index per direction, vector component, sign
HP3D_HD inline unsigned pack_desc(int ix, int iy, int iz, int comp, int neg)
{
    return (unsigned)ix | ((unsigned)iy << 5) | ((unsigned)iz << 10) | ((unsigned)comp << 15) |
           ((unsigned)neg << 17);
}

// Descriptor of H1 dof k in hp3D's element order (vertices, edges, faces,
// bubbles; Hexahedron.F90:73-156).
HP3D_HD inline unsigned desc_h1(int k, const ElemOrders &o)
{
    int idx[3];
    if (k < 8) return pack_desc(vtx_mu(k, 0), vtx_mu(k, 1), vtx_mu(k, 2), 0, 0);
    k -= 8;
    for (int e = 0; e < 12; ++e) {
        int ne = o.pe[e] - 1;
        if (k < ne) {
            EdgeT et = edge_t(e);
            idx[et.dir] = t_phi(k + 2, o.oe[e]);
            idx[et.d1] = et.a1;
            idx[et.d2] = et.a2;
            return pack_desc(idx[0], idx[1], idx[2], 0, 0);
        }
        k -= ne;
    }
    for (int f = 0; f < 6; ++f) {
        int h = o.fh[f], v = o.fv[f], nf = (h - 1) * (v - 1);
        if (k < nf) {
            FaceT ft = face_t(f);
            FaceOri fo = face_ori(o.of[f]);
            int ds = fo.ps ? ft.ax1 : ft.ax0, dt = fo.pt ? ft.ax1 : ft.ax0;
            int i = k % (h - 1) + 2, j = k / (h - 1) + 2;
            idx[ds] = t_phi(i, fo.rs);
            idx[dt] = t_phi(j, fo.rt);
            idx[ft.dn] = ft.a;
            return pack_desc(idx[0], idx[1], idx[2], 0, 0);
        }
        k -= nf;
    }
    int bx = o.b[0] - 1, by = o.b[1] - 1;
    return pack_desc(k % bx + 2, (k / bx) % by + 2, k / (bx * by) + 2, 0, 0);
}

// Descriptor of H(curl) dof k (Hexahedron.F90:283-388): one vector component,
// sign from the orientation of the coordinate carrying the Legendre factor.
HP3D_HD inline unsigned desc_hcurl(int k, const ElemOrders &o)
{
    int idx[3];
    for (int e = 0; e < 12; ++e) {
        int ne = o.pe[e];
        if (k < ne) {
            EdgeT et = edge_t(e);
            idx[et.dir] = t_leg(k, o.oe[e]);
            idx[et.d1] = et.a1;
            idx[et.d2] = et.a2;
            return pack_desc(idx[0], idx[1], idx[2], et.dir, o.oe[e]);
        }
        k -= ne;
    }
    for (int f = 0; f < 6; ++f) {
        int h = o.fh[f], v = o.fv[f];
        int n0 = h * (v - 1), n1 = (h - 1) * v;
        if (k < n0 + n1) {
            FaceT ft = face_t(f);
            FaceOri fo = face_ori(o.of[f]);
            int ds = fo.ps ? ft.ax1 : ft.ax0, dt = fo.pt ? ft.ax1 : ft.ax0;
            idx[ft.dn] = ft.a;
            if (k < n0) { // family 0: P_i along s, phi_j along t; j outer
                int i = k % h, j = k / h + 2;
                idx[ds] = t_leg(i, fo.rs);
                idx[dt] = t_phi(j, fo.rt);
                return pack_desc(idx[0], idx[1], idx[2], ds, fo.rs);
            }
            k -= n0; // family 1: phi_i along s (inner), P_j along t (outer)
            int i = k % (h - 1) + 2, j = k / (h - 1);
            idx[ds] = t_phi(i, fo.rs);
            idx[dt] = t_leg(j, fo.rt);
            return pack_desc(idx[0], idx[1], idx[2], dt, fo.rt);
        }
        k -= n0 + n1;
    }
    for (int a = 0; a < 3; ++a) {
        int cnt[3];
        for (int d = 0; d < 3; ++d) cnt[d] = (d == a) ? o.b[d] : o.b[d] - 1;
        int nfam = cnt[0] * cnt[1] * cnt[2];
        if (k < nfam) {
            int g[3] = {k % cnt[0], (k / cnt[0]) % cnt[1], k / (cnt[0] * cnt[1])};
            for (int d = 0; d < 3; ++d) idx[d] = (d == a) ? t_leg(g[d], 0) : t_phi(g[d] + 2, 0);
            return pack_desc(idx[0], idx[1], idx[2], a, 0);
        }
        k -= nfam;
    }
    return 0;
}

HP3D_HD inline int desc_ix(unsigned d, int dir) { return (d >> (5 * dir)) & 31; }
HP3D_HD inline int desc_comp(unsigned d) { return (d >> 15) & 3; }
HP3D_HD inline int desc_neg(unsigned d) { return (d >> 17) & 1; }

// The NTAB entries of the 1D factor table at one coordinate x in [0, 1]
// (values v, d/dx dv), with the formulas and evaluation order of the
// kernels' table fill: 1 - x and x; phi_i(s) = (P_i(2s-1) - P_{i-2}(2s-1)) /
// (4i - 2) with d phi_i/ds = P_{i-1}(2s-1), s = x or 1 - x (PolyILegendre,
// Idec = .true.); the scaled Legendre P_i(s; t), t = (1 - x) + x, of the
// oriented pair (HomLegendre).
HP3D_HD inline void table_at(double x, double *v, double *dv)
{
    v[0] = 1.0 - x;
    dv[0] = -1.0;
    v[1] = x;
    dv[1] = 1.0;
    for (int flip = 0; flip < 2; ++flip) {
        double s = flip ? 1.0 - x : x;
        double y = 2.0 * s - 1.0, p[MAXP + 1];
        p[0] = 1.0;
        p[1] = y;
        for (int i = 2; i <= MAXP; ++i)
            p[i] = ((double)(2 * i - 1) * y * p[i - 1] - (double)(i - 1) * p[i - 2]) / (double)i;
        for (int i = 2; i <= MAXP; ++i) {
            v[t_phi(i, flip)] = (p[i] - p[i - 2]) / (double)(4 * i - 2);
            dv[t_phi(i, flip)] = flip ? -p[i - 1] : p[i - 1];
        }
        double t = (1.0 - x) + x;
        double yy = 2.0 * s - t, tt = t * t, r[MAXP];
        r[0] = 1.0;
        r[1] = yy;
        for (int i = 2; i < MAXP; ++i)
            r[i] = ((double)(2 * i - 1) * yy * r[i - 1] - (double)(i - 1) * tt * r[i - 2]) / (double)i;
        for (int i = 0; i < MAXP; ++i) {
            v[t_leg(i, flip)] = r[i];
            dv[t_leg(i, flip)] = 0.0;
        }
    }
}

// Descriptor pass of the fused kernels: desc_h1 with the dof's entity found
// by a branch-free scan over the edge and face dof counts (every lane runs the
// same instructions; the reference walk exits at a different entity per lane,
// so a wave ran every exit), the edge, face and orientation tables as packed
// constants, and the face and bubble index splits by a float reciprocal, so a
// wave takes each entity kind's path once.  Same descriptors as desc_h1.
constexpr unsigned long long pack_edges(int lo)
{
    unsigned long long w = 0;
    for (int e = lo; e < lo + 8 && e < 12; ++e) {
        const EdgeT t = edge_t(e);
        w |= (unsigned long long)(t.dir | t.a1 << 2 | t.d1 << 3 | t.a2 << 5 | t.d2 << 6) << (8 * (e - lo));
    }
    return w;
}
constexpr unsigned long long pack_faces()
{
    unsigned long long w = 0;
    for (int f = 0; f < 6; ++f) {
        const FaceT t = face_t(f);
        w |= (unsigned long long)(t.a | t.dn << 1 | t.ax0 << 3 | t.ax1 << 5) << (8 * f);
    }
    return w;
}
constexpr unsigned pack_face_ori()
{
    unsigned w = 0;
    for (int c = 0; c < 8; ++c) {
        const FaceOri t = face_ori(c);
        w |= (unsigned)(t.ps | t.pt << 1 | t.rs << 2 | t.rt << 3) << (4 * c);
    }
    return w;
}
constexpr unsigned long long EDGE_LO = pack_edges(0), EDGE_HI = pack_edges(8), FACES = pack_faces();
constexpr unsigned FACE_ORI = pack_face_ori();

// floor(k / d) for 0 <= k < 1024, 1 <= d <= 36: (k + 1/2) / d lies at least
// 1/72 from an integer, far beyond the reciprocal's rounding (the hardware
// estimate on the device, the IEEE quotient on the host)
HP3D_HD inline int div_small(int k, int d)
{
#if defined(__HIP_DEVICE_COMPILE__)
    return (int)(((float)k + 0.5f) * __builtin_amdgcn_rcpf((float)d));
#else
    return (int)(((float)k + 0.5f) * (1.0f / (float)d));
#endif
}

HP3D_HD inline unsigned desc_h1_scan(int k, const ElemOrders &o)
{
    if (k < 8) return pack_desc(vtx_mu(k, 0), vtx_mu(k, 1), vtx_mu(k, 2), 0, 0);
    k -= 8;
    // the entity's orientation and face width are picked up during the scan,
    // so nothing is indexed by a lane-dependent value
    int acc = 0, ent = -1, kl = 0, ori = 0, h1 = 1;
#pragma unroll
    for (int e = 0; e < 12; ++e) {
        const int ne = o.pe[e] - 1;
        const bool hit = ent < 0 && k < acc + ne;
        ent = hit ? e : ent;
        kl = hit ? k - acc : kl;
        ori = hit ? o.oe[e] : ori;
        acc += ne;
    }
#pragma unroll
    for (int f = 0; f < 6; ++f) {
        const int nf = (o.fh[f] - 1) * (o.fv[f] - 1);
        const bool hit = ent < 0 && k < acc + nf;
        ent = hit ? 12 + f : ent;
        kl = hit ? k - acc : kl;
        ori = hit ? o.of[f] : ori;
        h1 = hit ? o.fh[f] - 1 : h1;
        acc += nf;
    }
    // (direction, index) pairs that together cover the three directions
    auto put = [](int d0, int v0, int d1, int v1, int d2, int v2) {
        const int i0 = d0 == 0 ? v0 : (d1 == 0 ? v1 : v2);
        const int i1 = d0 == 1 ? v0 : (d1 == 1 ? v1 : v2);
        const int i2 = d0 == 2 ? v0 : (d1 == 2 ? v1 : v2);
        return pack_desc(i0, i1, i2, 0, 0);
    };
    if (ent < 0) { // bubble
        const int kb = k - acc, bx = o.b[0] - 1, by = o.b[1] - 1;
        const int q1 = div_small(kb, bx), q2 = div_small(kb, bx * by);
        return pack_desc(kb - q1 * bx + 2, q1 - div_small(q1, by) * by + 2, q2 + 2, 0, 0);
    }
    if (ent < 12) { // edge
        const unsigned b = (unsigned)(((ent < 8 ? EDGE_LO : EDGE_HI) >> (8 * (ent & 7))) & 0xffu);
        const int dir = b & 3, a1 = (b >> 2) & 1, d1 = (b >> 3) & 3, a2 = (b >> 5) & 1, d2 = (b >> 6) & 3;
        return put(dir, t_phi(kl + 2, ori), d1, a1, d2, a2);
    }
    const int f = ent - 12; // face
    const unsigned b = (unsigned)((FACES >> (8 * f)) & 0xffu);
    const int fa = b & 1, dn = (b >> 1) & 3, ax0 = (b >> 3) & 3, ax1 = (b >> 5) & 3;
    const unsigned oo = (FACE_ORI >> (4 * ori)) & 15u;
    const int ps = oo & 1, pt = (oo >> 1) & 1, rs = (oo >> 2) & 1, rt = (oo >> 3) & 1;
    const int j0 = div_small(kl, h1);
    return put(ps ? ax1 : ax0, t_phi(kl - j0 * h1 + 2, rs), pt ? ax1 : ax0, t_phi(j0 + 2, rt), dn, fa);
}

} // namespace hp3d
