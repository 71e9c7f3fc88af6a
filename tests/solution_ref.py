"""FP64 CPU reference of the solution path, built from the oracle: element
loop with condensation (oracle.element_loop), assembly with scipy from the
DofMap numbering, a sparse direct solve, back substitution (oracle.stc_bwd)
and element_error (oracle.element_error).  Test infrastructure only."""
import numpy as np
import scipy.sparse as sp
import scipy.sparse.linalg as spla

import oracle
from hp3d_amd.solve import DofMap


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


def assemble(mesh, space, ref, dm):
    rows, cols, vals = [], [], []
    rhs = np.zeros(dm.ndof)
    cnt = oracle.ndof_h1 if space == 0 else oracle.ndof_hcurl
    bub = oracle.nbubble_h1 if space == 0 else oracle.nbubble_hcurl
    for e in range(mesh.nel):
        n, nb = cnt(mesh.norder[e]), bub(mesh.norder[e][18])
        aii, bi, _, _ = unpack(ref, e, n, nb)
        g = dm.gid[e, : n - nb]
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


def element_vectors(mesh, space, ref, dm, x):
    cnt = oracle.ndof_h1 if space == 0 else oracle.ndof_hcurl
    bub = oracle.nbubble_h1 if space == 0 else oracle.nbubble_hcurl
    xe = np.concatenate([x, [0.0]])
    out = []
    for e in range(mesh.nel):
        n, nb = cnt(mesh.norder[e]), bub(mesh.norder[e][18])
        _, _, asc, bsc = unpack(ref, e, n, nb)
        g = dm.gid[e, : n - nb]
        xi = xe[np.where(g >= 0, g, dm.ndof)]
        xb = oracle.stc_bwd(asc, bsc, xi) if nb else np.zeros(0)
        out.append(np.concatenate([xi, xb]))
    return out


def solve(mesh, space, load, exact=None, coef=None):
    exact = load if exact is None else exact
    ref, _ = element_data(mesh, space, load, coef)
    dm = DofMap(mesh, space)
    a, rhs = assemble(mesh, space, ref, dm)
    x = spla.spsolve(a.tocsc(), rhs)
    u = element_vectors(mesh, space, ref, dm, x)
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
        g = dm.gid[e, : k.shape[0] - nb]
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
        g = dm.gid[e, : k.shape[0] - nb]
        li = le[np.where(g >= 0, g, dm.ndof)]
        asc, bsc = bsch_adj[e]
        lb = oracle.stc_bwd(asc, bsc, li) if nb else np.zeros(0)
        lam = np.concatenate([li, lb])
        grad[e] = -lam @ k @ s["u"][e]
    return J, grad
