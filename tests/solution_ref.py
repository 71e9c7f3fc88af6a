"""FP64 CPU reference of the solution path, built from the oracle: element
loop with condensation (oracle.element_loop), a global numbering of its own
(Numbering below: hp3D's first-visit NFIRST_DOF, written here independently
of the product's), assembly with scipy, a sparse direct solve, back
substitution (oracle.stc_bwd) and element_error (oracle.element_error).  Test
infrastructure only."""
import numpy as np
import scipy.sparse as sp
import scipy.sparse.linalg as spla

import oracle

# local node tables of the hexahedron: edges and faces by their vertices
_EDGES = [(0, 1), (1, 2), (3, 2), (0, 3), (4, 5), (5, 6), (7, 6), (4, 7), (0, 4), (1, 5), (2, 6),
          (3, 7)]
_FACES = [(0, 1, 2, 3), (4, 5, 6, 7), (0, 1, 5, 4), (1, 2, 6, 5), (3, 2, 6, 7), (0, 3, 7, 4)]


class Numbering:
    """Global numbers of the interface dofs, NFIRST_DOF style
    (mumps_sc.F90:150-158): the element loop in mesh order visits each
    element's vertices, edges and faces; a node gets its dofs' numbers on its
    first visit.  Nodes are keyed by their global vertex sets; Dirichlet
    nodes (on a face of one element) get -1.  gid[e][k] follows the element's
    interface dof order."""

    def __init__(self, mesh, space):
        el = mesh.elems
        face_count = {}
        for e in range(mesh.nel):
            for f in _FACES:
                key = frozenset(int(el[e, v]) for v in f)
                face_count[key] = face_count.get(key, 0) + 1
        bnd = set()
        for key, c in face_count.items():
            if c == 1:
                bnd.add(("f", key))
                vs = sorted(key)
                for v in vs:
                    bnd.add(("v", frozenset([v])))
        for e in range(mesh.nel):
            for f in _FACES:
                if face_count[frozenset(int(el[e, v]) for v in f)] == 1:
                    for a, b in _EDGES:
                        if a in f and b in f:
                            bnd.add(("e", frozenset([int(el[e, a]), int(el[e, b])])))
        first, nxt = {}, 0
        self.gid, self.ni = [], []
        for e in range(mesh.nel):
            no = mesh.norder[e]
            nodes = [("v", frozenset([int(el[e, v])]), 1 if space == 0 else 0) for v in range(8)]
            for k, (a, b) in enumerate(_EDGES):
                nodes.append(("e", frozenset([int(el[e, a]), int(el[e, b])]),
                              no[k] - 1 if space == 0 else no[k]))
            for k, f in enumerate(_FACES):
                h, v = no[12 + k] // 10, no[12 + k] % 10
                nodes.append(("f", frozenset(int(el[e, q]) for q in f),
                              (h - 1) * (v - 1) if space == 0 else h * (v - 1) + (h - 1) * v))
            g = []
            for kind, key, cnt in nodes:
                fixed = (kind, key) in bnd
                if not fixed and (kind, key) not in first:
                    first[(kind, key)] = nxt
                    nxt += int(cnt)
                g.extend([-1] * int(cnt) if fixed else range(first[(kind, key)],
                                                            first[(kind, key)] + int(cnt)))
            self.gid.append(np.array(g, dtype=np.int64))
            self.ni.append(len(g))
        self.ndof = nxt


def element_data(mesh, space, load, coef=None):
    coef = mesh.coef if coef is None else coef
    ld = int(max(oracle.ndof_h1(no) if space == 0 else oracle.ndof_hcurl(no) for no in mesh.norder))
    return oracle.element_loop(space, mesh.norder, mesh.nori_e, mesh.nori_f, mesh.xnod(), coef,
                               load, ldk=ld, want_k=True, want_stc=True, herm=True), ld


def unpack(ref, e, n, nb):
    ni = n - nb
    aii = ref["aii"][e].T[:ni, :ni]
    bi = ref["bi"][e][:ni]
    asc = ref["aschur"][e].reshape(-1)[: nb * ni].reshape(ni, nb).T
    bsc = ref["bschur"][e][:nb]
    return aii, bi, asc, bsc


def assemble(mesh, space, ref, dm, ud=None):
    """Global condensed system; ud (per element, interface order) lifts the
    Dirichlet values into the loads (celem_systemI.F90:794-808)."""
    rows, cols, vals = [], [], []
    rhs = np.zeros(dm.ndof)
    cnt = oracle.ndof_h1 if space == 0 else oracle.ndof_hcurl
    bub = oracle.nbubble_h1 if space == 0 else oracle.nbubble_hcurl
    for e in range(mesh.nel):
        n, nb = cnt(mesh.norder[e]), bub(mesh.norder[e][18])
        aii, bi, _, _ = unpack(ref, e, n, nb)
        if ud is not None:
            bi = bi - aii @ ud[e][: n - nb]
        g = dm.gid[e][: n - nb]
        ok = g >= 0
        gi = g[ok]
        a = aii[np.ix_(ok, ok)]
        rows.append(np.repeat(gi, gi.size))
        cols.append(np.tile(gi, gi.size))
        vals.append(a.reshape(-1))
        np.add.at(rhs, gi, bi[ok])
    a = sp.coo_matrix((np.concatenate(vals), (np.concatenate(rows), np.concatenate(cols))),
                      shape=(dm.ndof, dm.ndof)).tocsr()
    return a, rhs


def element_vectors(mesh, space, ref, dm, x, ud=None):
    cnt = oracle.ndof_h1 if space == 0 else oracle.ndof_hcurl
    bub = oracle.nbubble_h1 if space == 0 else oracle.nbubble_hcurl
    xe = np.concatenate([x, [0.0]])
    out = []
    for e in range(mesh.nel):
        n, nb = cnt(mesh.norder[e]), bub(mesh.norder[e][18])
        _, _, asc, bsc = unpack(ref, e, n, nb)
        g = dm.gid[e][: n - nb]
        xi = xe[np.where(g >= 0, g, dm.ndof)]
        if ud is not None:
            xi = np.where(g >= 0, xi, ud[e][: n - nb])
        xb = oracle.stc_bwd(asc, bsc, xi) if nb else np.zeros(0)
        out.append(np.concatenate([xi, xb]))
    return out


def solve(mesh, space, load, exact=None, coef=None, ud=None):
    """ud: per-element Dirichlet values (interface order, 0 at free dofs) for
    non-zero boundary data, else homogeneous."""
    exact = load if exact is None else exact
    ref, _ = element_data(mesh, space, load, coef)
    dm = Numbering(mesh, space)
    a, rhs = assemble(mesh, space, ref, dm, ud)
    x = spla.spsolve(a.tocsc(), rhs)
    u = element_vectors(mesh, space, ref, dm, x, ud)
    err = nrm = 0.0
    xn = mesh.xnod()
    for e in range(mesh.nel):
        de, dn = oracle.element_error(space, mesh.norder[e], mesh.nori_e[e], mesh.nori_f[e],
                                      xn[e].T, u[e], exact)
        err += de
        nrm += dn
    return {"x": x, "u": u, "error2": err, "norm2": nrm, "a": a, "rhs": rhs, "dofmap": dm,
            "ref": ref}


def adjoint(mesh, theta, theta_t, load):
    """J = 1/2 ||u(theta) - u(theta_t)||^2_L2 and dJ/dtheta_e = -lambda_e^T K_e u_e,
    everything on the CPU from the oracle (element matrices, condensation,
    mass matrices) and a direct solve."""
    kappa, kappa_t = np.exp(theta), np.exp(theta_t)
    s = solve(mesh, 0, load, coef=kappa)
    st = solve(mesh, 0, load, coef=kappa_t)
    dm = s["dofmap"]
    xn = mesh.xnod()
    J = 0.0
    rhs = np.zeros(dm.ndof)
    ks, bsch_adj, ms = [], [], []
    for e in range(mesh.nel):
        no, ne, nf = mesh.norder[e], mesh.nori_e[e], mesh.nori_f[e]
        m = oracle.elem_mass(0, no, ne, nf, xn[e].T)
        d = s["u"][e] - st["u"][e]
        J += 0.5 * d @ m @ d
        r = m @ d
        k, _, _ = oracle.elem_h1(no, ne, nf, xn[e].T, kappa=kappa[e], load=load)
        nb = oracle.nbubble_h1(no[18])
        aii, bi, asc, bsc = oracle.stc_fwd(k, r, nb, herm=True)
        g = dm.gid[e][: k.shape[0] - nb]
        ok = g >= 0
        np.add.at(rhs, g[ok], bi[ok])
        ks.append(k)
        bsch_adj.append((asc, bsc))
    lam_i = spla.spsolve(s["a"].tocsc(), rhs)
    le = np.concatenate([lam_i, [0.0]])
    grad = np.zeros(mesh.nel)
    for e in range(mesh.nel):
        k = ks[e]
        nb = oracle.nbubble_h1(mesh.norder[e][18])
        g = dm.gid[e][: k.shape[0] - nb]
        li = le[np.where(g >= 0, g, dm.ndof)]
        asc, bsc = bsch_adj[e]
        lb = oracle.stc_bwd(asc, bsc, li) if nb else np.zeros(0)
        lam = np.concatenate([li, lb])
        grad[e] = -lam @ k @ s["u"][e]
    return J, grad
