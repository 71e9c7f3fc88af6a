"""Dirichlet data of the boundary nodes for H1 problems with non-zero
boundary values: update_Ddof (src/hpinterp/update_Ddof.F90) with the
projection-based interpolation of dhpvert / dhpedgeH / dhpfaceH
(src/hpinterp/dhpvert.F90, edge/dhpedgeH.F90, face/dhpfaceH.F90):

* a vertex takes the value of the exact solution;
* a boundary edge's dofs minimise the H1 seminorm along the edge,
  int_e |d/ds (u - u_V - sum_i c_i psi_i)|^2 ds, with u_V the vertex
  interpolant;
* a boundary face's dofs minimise int_F |grad_F (u - u_V - u_E - sum_j c_j
  psi_j)|^2 dS, the tangential gradient on the face, with u_E the edge
  interpolant.

The reference takes these derivatives in the parametric coordinates of the
GMP block the element came from; here they are taken in physical
coordinates of the element's trilinear map, which is the same thing for
elements of a single affine block (the unit cube of test/poly_pois.F90 and
test/conv_pois.F90) and an equally admissible interpolant elsewhere.  The
shape functions come from hp3d_gpu_shape (the kernels' own descriptors,
evaluated on the host): this is set-up work on the boundary, as
update_Ddof is on the reference's host.

dirichlet_values returns, per element, the Dirichlet values of its interface
dofs (element order, 0 at free dofs): the ZDOFD of celem_systemI.
"""
import ctypes

import numpy as np

from . import _lib
from ._lib import H1, LOAD_CONV, LOAD_POLY2, LOAD_SINSINSIN, LOAD_XYZ, check
from .mesh import EDGE_VERTS, FACE_AXES, FACE_VERTS, VERT_XYZ

NG = 10  # Gauss points per direction of the projections


def exact_h1(sel):
    """u(x) and grad u(x) (vectorised over points) of the manufactured
    solutions the load selectors belong to."""
    pi = np.pi

    def f(x):
        x0, x1, x2 = x[:, 0], x[:, 1], x[:, 2]
        if sel == LOAD_SINSINSIN:
            sx, sy, sz = np.sin(pi * x0), np.sin(pi * x1), np.sin(pi * x2)
            u = sx * sy * sz
            g = np.stack([pi * np.cos(pi * x0) * sy * sz, pi * sx * np.cos(pi * x1) * sz,
                          pi * sx * sy * np.cos(pi * x2)], 1)
        elif sel == LOAD_CONV:
            sx, sh = np.sin(pi * x0), np.sinh(pi * x1)
            u = sx * sh * x2 * x2
            g = np.stack([pi * np.cos(pi * x0) * sh * x2 * x2, pi * sx * np.cosh(pi * x1) * x2 * x2,
                          2.0 * sx * sh * x2], 1)
        elif sel == LOAD_POLY2:
            a, b, c = x0 * (1 - x0), x1 * (1 - x1), x2 * (1 - x2)
            u = a * b * c
            g = np.stack([(1 - 2 * x0) * b * c, a * (1 - 2 * x1) * c, a * b * (1 - 2 * x2)], 1)
        elif sel == LOAD_XYZ:
            u = x0 * x1 * x2
            g = np.stack([x1 * x2, x0 * x2, x0 * x1], 1)
        else:
            raise ValueError("no exact solution for selector %d" % sel)
        return u, g
    return f


def shape_h1(norder, nori_e, nori_f, xi):
    """Values (npts, N) and reference gradients (npts, N, 3) of an element's
    H1 shape functions at master points xi (npts, 3), by hp3d_gpu_shape."""
    lib = _lib.load()
    no = np.ascontiguousarray(norder, dtype=np.int32)
    oe = np.ascontiguousarray(nori_e, dtype=np.int32)
    of = np.ascontiguousarray(nori_f, dtype=np.int32)
    xi = np.ascontiguousarray(xi, dtype=np.float64)
    n = ctypes.c_int()
    check(lib.hp3d_gpu_ndof(H1, no.ctypes.data_as(ctypes.c_void_p), ctypes.byref(n), None),
          "hp3d_gpu_ndof")
    npts = xi.shape[0]
    val = np.zeros((npts, n.value))
    grad = np.zeros((npts, n.value, 3))
    p = lambda a: a.ctypes.data_as(ctypes.c_void_p)
    check(lib.hp3d_gpu_shape(H1, p(no), p(oe), p(of), npts, p(xi), p(val), p(grad)),
          "hp3d_gpu_shape")
    return val, grad


def _gauss(n):
    x, w = np.polynomial.legendre.leggauss(n)
    return 0.5 * (x + 1.0), 0.5 * w


def dirichlet_values(mesh, dofmap, exact):
    """ud (nel, width): each element's interface dofs' Dirichlet values, in
    element order, 0 at free dofs (H1).  `exact` is a selector or a callable
    x -> (u, grad u)."""
    if dofmap.space != H1:
        raise ValueError("Dirichlet data are built for H1 here")
    u_fn = exact_h1(exact) if isinstance(exact, (int, np.integer)) else exact
    m = mesh
    bvert, bedge, bface = dofmap.fixed
    vval = np.zeros(m.verts.shape[0])
    ix = np.nonzero(bvert)[0]
    if ix.size:
        vval[ix] = u_fn(m.verts[ix])[0]
    evals, fvals = {}, {}
    tq, wq = _gauss(NG)
    no_all, ne_all, nf_all = m.norder, m.nori_e, m.nori_f
    # element dof ranges of each node: vertices, edges, faces (element order)
    def ranges(no):
        pe = no[:12] - 1
        fh, fv = no[12:18] // 10, no[12:18] % 10
        nf = (fh - 1) * (fv - 1)
        e0 = 8 + np.concatenate([[0], np.cumsum(pe)[:-1]])
        f0 = 8 + pe.sum() + np.concatenate([[0], np.cumsum(nf)[:-1]])
        return e0, pe, f0, nf
    xn = m.xnod()
    els, lfs = np.nonzero(bface[m.elem_faces])
    # edges first (every boundary edge lies on a boundary face)
    for e, lf in zip(els, lfs):
        no, oe, of = no_all[e], ne_all[e], nf_all[e]
        e0, pe, f0, nf = ranges(no)
        X = xn[e]
        for le in np.nonzero([set(EDGE_VERTS[k]) <= set(FACE_VERTS[lf]) for k in range(12)])[0]:
            g = int(m.elem_edges[e, le])
            if g in evals or not bedge[g]:
                continue
            if pe[le] == 0:
                evals[g] = np.zeros(0)
                continue
            a, b = VERT_XYZ[EDGE_VERTS[le][0]], VERT_XYZ[EDGE_VERTS[le][1]]
            dxi = (b - a).astype(np.float64)
            xi = a[None, :] + tq[:, None] * dxi[None, :]
            val, grad = shape_h1(no, oe, of, xi)
            dphi = grad @ dxi                                  # d/dt of every function
            x = val[:, :8] @ X                                 # physical points
            xt = dphi[:, :8] @ X                               # dx/dt
            js = np.linalg.norm(xt, axis=1)
            uu, gu = u_fn(x)
            r = (gu * xt).sum(1) - dphi[:, :8] @ vval[m.elems[e]]
            psi = dphi[:, e0[le]:e0[le] + pe[le]]
            wt = wq / js
            A = (psi * wt[:, None]).T @ psi
            evals[g] = np.linalg.solve(A, (psi * wt[:, None]).T @ r)
    for e, lf in zip(els, lfs):
        g = int(m.elem_faces[e, lf])
        if g in fvals:
            continue
        no, oe, of = no_all[e], ne_all[e], nf_all[e]
        e0, pe, f0, nf = ranges(no)
        if nf[lf] == 0:
            fvals[g] = np.zeros(0)
            continue
        X = xn[e]
        ax0, ax1 = FACE_AXES[lf]
        corner = VERT_XYZ[FACE_VERTS[lf][0]].astype(np.float64)
        s, t = np.meshgrid(tq, tq, indexing="ij")
        w2 = np.outer(wq, wq).ravel()
        xi = np.repeat(corner[None, :], NG * NG, 0)
        xi[:, ax0], xi[:, ax1] = s.ravel(), t.ravel()
        val, grad = shape_h1(no, oe, of, xi)
        x = val[:, :8] @ X
        xa, xb = grad[:, :8, ax0] @ X, grad[:, :8, ax1] @ X   # tangents
        G = np.stack([np.stack([(xa * xa).sum(1), (xa * xb).sum(1)], 1),
                      np.stack([(xb * xa).sum(1), (xb * xb).sum(1)], 1)], 1)
        detg = G[:, 0, 0] * G[:, 1, 1] - G[:, 0, 1] * G[:, 1, 0]
        Gi = np.stack([np.stack([G[:, 1, 1], -G[:, 0, 1]], 1),
                       np.stack([-G[:, 1, 0], G[:, 0, 0]], 1)], 1) / detg[:, None, None]
        uu, gu = u_fn(x)
        # u - u_V - u_E, its derivatives along the two face axes
        coef = np.zeros(val.shape[1])
        coef[:8] = vval[m.elems[e]]
        for le in range(12):
            if set(EDGE_VERTS[le]) <= set(FACE_VERTS[lf]) and pe[le] > 0:
                coef[e0[le]:e0[le] + pe[le]] = evals[int(m.elem_edges[e, le])]
        ra = (gu * xa).sum(1) - grad[:, :, ax0] @ coef
        rb = (gu * xb).sum(1) - grad[:, :, ax1] @ coef
        pa = grad[:, f0[lf]:f0[lf] + nf[lf], ax0]
        pb = grad[:, f0[lf]:f0[lf] + nf[lf], ax1]
        wt = w2 * np.sqrt(detg)
        A = ((pa * (Gi[:, 0, 0] * wt)[:, None]).T @ pa + (pa * (Gi[:, 0, 1] * wt)[:, None]).T @ pb
             + (pb * (Gi[:, 1, 0] * wt)[:, None]).T @ pa + (pb * (Gi[:, 1, 1] * wt)[:, None]).T @ pb)
        rhs = (pa.T @ ((Gi[:, 0, 0] * ra + Gi[:, 0, 1] * rb) * wt)
               + pb.T @ ((Gi[:, 1, 0] * ra + Gi[:, 1, 1] * rb) * wt))
        fvals[g] = np.linalg.solve(A, rhs)
    # per element: the values of its Dirichlet interface dofs
    width = dofmap.gid.shape[1]
    ud = np.zeros((m.nel, width))
    kind, nid, nj, gid = dofmap.node_kind, dofmap.node_id, dofmap.node_j, dofmap.gid
    fix = gid == -1
    e_i, k_i = np.nonzero(fix & (kind == 0))
    ud[e_i, k_i] = vval[nid[e_i, k_i]]
    for knd, table in ((1, evals), (2, fvals)):
        e_i, k_i = np.nonzero(fix & (kind == knd))
        for e, k in zip(e_i, k_i):
            ud[e, k] = table[int(nid[e, k])][nj[e, k]]
    return ud
