"""Global system around the element path: numbering, device assembly of the
condensed element systems, solve, back substitution and element errors.

This is the host-side half of hp3D's mumps_sc flow (src/solver/omp_mumps/
mumps_sc.F90) for the synthetic structured meshes of ``hp3d_amd.mesh``:

* ``DofMap``: global numbering of the interface dofs.  Nodes are numbered
  vertices, then edges, then faces, each node's dofs contiguous and in node
  order (the NFIRST_DOF first-visit numbering of mumps_sc.F90:150-158 orders
  nodes by element visit instead; any numbering gives the same solution).
  Homogeneous Dirichlet data on the whole boundary removes every dof of a
  boundary node (ibc = 1 on all faces, as in test/poly_pois.F90 and the
  cfg meshes' manufactured solutions that vanish on the boundary).
* ``Assembly``: symbolic assembly once (CSR pattern and, per nonzero, the
  positions of its element contributions in the packed condensed matrices),
  then each numeric assembly is one ``hp3d_gpu_gather_sum`` launch; this
  replaces the triplet wire format of mumps_sc.F90:266-323.
* ``Solver``: the direct solve that MUMPS performs in the reference
  (mumps_sc.F90:368-426), by SuperLU (scipy) on the host, factorised once and
  reused for further right-hand sides (the adjoint solve reuses the
  factorisation); ``PCGSolver``: Jacobi-preconditioned conjugate gradients on
  the device (hp3d_gpu_csr_spmv) for systems too large for the host direct
  solve, with the assembled matrix and preconditioner reused across solves.
* ``solve_problem``: elements -> condensation -> assembly -> solve -> back
  substitution (hp3d_gpu_stc_bwd_batch) -> element_error
  (hp3d_gpu_field_batch), i.e. what test/poly_pois.F90 and conv_pois.F90 do.
"""
import ctypes

import numpy as np

from . import _lib
from ._lib import FIELD_ERROR, H1, Hp3dError, check
from .elem import ElementBatch
from .mesh import EDGE_VERTS, FACE_VERTS


def _face_edges():
    """Local edges of each local face (edges whose both vertices are on it)."""
    out = []
    for fv in FACE_VERTS:
        s = set(fv.tolist())
        out.append([e for e, (a, b) in enumerate(EDGE_VERTS) if a in s and b in s])
    return np.array(out)


FACE_EDGES = _face_edges()


class DofMap:
    """Global interface-dof numbering of a HexMesh for one energy space.

    gid[e, :ni_e] are the global numbers of element e's interface dofs in
    element order (vertices, edges, faces), -1 for a Dirichlet dof; columns
    beyond ni_e hold -2.  node_kind[e, k] (0 vertex, 1 edge, 2 face), node_id
    and node_j give the node each dof belongs to and its index among the
    node's dofs (equal across the elements sharing the node: trace
    conformity).

    numbering="blocks" numbers the free dofs node by node, vertices, then
    edges, then faces, each by global node id; numbering="first_visit" is
    hp3D's NFIRST_DOF: the nodes in the order the element loop first visits
    them (elements in mesh order, each element's vertices, edges, faces),
    mumps_sc.F90:150-158."""

    def __init__(self, mesh, space=H1, dirichlet=True, numbering="blocks"):
        self.space = space
        m = mesh
        nel = m.nel
        nv, ned, nfa = m.verts.shape[0], m.edges.shape[0], m.faces.shape[0]
        # dofs per node
        if space == H1:
            nvd = np.ones(nv, np.int64)
            ned_d = m.edge_order - 1
            nfd = (m.face_order[:, 0] - 1) * (m.face_order[:, 1] - 1)
        else:
            nvd = np.zeros(nv, np.int64)
            ned_d = m.edge_order.copy()
            h, v = m.face_order[:, 0], m.face_order[:, 1]
            nfd = h * (v - 1) + (h - 1) * v
        # boundary nodes: faces of a single element, their edges and vertices
        fcount = np.bincount(m.elem_faces.ravel(), minlength=nfa)
        bface = fcount == 1
        bedge = np.zeros(ned, bool)
        bvert = np.zeros(nv, bool)
        el, lf = np.nonzero(bface[m.elem_faces])
        if el.size:
            bedge[m.elem_edges[el[:, None], FACE_EDGES[lf]].ravel()] = True
            bvert[m.elems[el[:, None], FACE_VERTS[lf]].ravel()] = True
        self.boundary = (bvert, bedge, bface)
        if not dirichlet:
            bface, bedge, bvert = bface & False, bedge & False, bvert & False
        self.fixed = (bvert, bedge, bface)
        counts = np.concatenate([np.where(bvert, 0, nvd), np.where(bedge, 0, ned_d),
                                 np.where(bface, 0, nfd)])
        if numbering == "blocks":
            base = np.concatenate([[0], np.cumsum(counts)[:-1]])
        elif numbering == "first_visit":
            # node keys (vertices, then edges, then faces in one index space)
            # in the order the element loop visits them
            visit = np.concatenate([m.elems, nv + m.elem_edges, nv + ned + m.elem_faces],
                                   axis=1).ravel()
            _, first = np.unique(visit, return_index=True)
            order = visit[np.sort(first)]              # nodes by first visit
            base = np.zeros(counts.size, np.int64)
            base[order] = np.concatenate([[0], np.cumsum(counts[order])[:-1]])
        else:
            raise ValueError("numbering must be 'blocks' or 'first_visit'")
        self.numbering = numbering
        self.ndof = int(counts.sum())
        vbase, ebase, fbase = base[:nv], base[nv:nv + ned], base[nv + ned:]
        vfix, efix, ffix = bvert, bedge, bface
        # per-element blocks (kind, node, base, count, fixed) in element order
        blocks = []
        if space == H1:
            for lv in range(8):
                g = m.elems[:, lv]
                blocks.append((0, g, vbase[g], np.ones(nel, np.int64), vfix[g]))
        no = m.norder
        for le in range(12):
            g = m.elem_edges[:, le]
            c = no[:, le] - 1 if space == H1 else no[:, le].astype(np.int64)
            blocks.append((1, g, ebase[g], c.astype(np.int64), efix[g]))
        for lf in range(6):
            g = m.elem_faces[:, lf]
            h, v = no[:, 12 + lf] // 10, no[:, 12 + lf] % 10
            c = (h - 1) * (v - 1) if space == H1 else h * (v - 1) + (h - 1) * v
            blocks.append((2, g, fbase[g], c.astype(np.int64), ffix[g]))
        ni = np.sum([b[3] for b in blocks], axis=0)
        self.ni = ni
        width = int(ni.max()) if nel else 0
        gid = np.full((nel, width), -2, np.int64)
        kind = np.full((nel, width), -1, np.int8)
        nid = np.full((nel, width), -1, np.int64)
        nj = np.full((nel, width), -1, np.int64)
        pos = np.zeros(nel, np.int64)
        rows = np.arange(nel)
        for knd, node, b0, c, fixed in blocks:
            for k in range(int(c.max()) if nel else 0):
                sel = k < c
                r = rows[sel]
                col = pos[sel] + k
                gid[r, col] = np.where(fixed[sel], -1, b0[sel] + k)
                kind[r, col] = knd
                nid[r, col] = node[sel]
                nj[r, col] = k
            pos += c
        self.gid = gid
        self.node_kind, self.node_id, self.node_j = kind, nid, nj

    def element_system(self, e, aii, bi, ud=None):
        """Element e's condensed system in the layout celem_systemI hands to
        the solver (celem_systemI.F90:285-330, 794-843, with ISYM_FLAG = 2):
        the Dirichlet lift Bi - Aii ud, then the free dofs only, nodes in
        reverse element order (faces 6..1, edges 12..1, vertices 8..1), each
        node's dofs in increasing order (NEXTRACT).  Returns (Zastif row-major
        nrdofc x nrdofc as a flat array, Zbload, LCON) with LCON the 1-based
        global numbers (mumps_sc.F90:266-283).  aii is ni x ni, bi length ni
        (host arrays, element order)."""
        ni = int(self.ni[e])
        g = self.gid[e, :ni]
        kind, node = self.node_kind[e, :ni], self.node_id[e, :ni]
        b = np.asarray(bi, dtype=np.float64)[:ni].copy()
        a = np.asarray(aii, dtype=np.float64)[:ni, :ni]
        if ud is not None:
            b -= a @ np.asarray(ud, dtype=np.float64)[:ni]
        # node blocks of the element in element order, then reversed
        starts = np.nonzero(np.r_[True, (kind[1:] != kind[:-1]) | (node[1:] != node[:-1])])[0]
        ends = np.r_[starts[1:], ni]
        nextract = np.concatenate([np.arange(s0, s1) for s0, s1 in zip(starts[::-1], ends[::-1])])
        nextract = nextract[g[nextract] >= 0]
        zastif = a[np.ix_(nextract, nextract)].reshape(-1)
        return zastif, b[nextract], g[nextract] + 1


class Assembly:
    """CSR pattern of the global condensed system and the contribution lists
    that the gather-sum kernel reduces (symbolic phase on the device)."""

    def __init__(self, batch: ElementBatch, dofmap: DofMap):
        import torch

        self.batch = batch
        self.dofmap = dofmap
        dev = batch.device
        nel = batch.nel
        ni = batch._ni                                   # internal order
        if not np.array_equal(dofmap.ni[batch.perm], ni):
            raise ValueError("dof map and element batch disagree on interface sizes")
        gid = torch.as_tensor(dofmap.gid[batch.perm], device=dev)   # internal order
        ni_t = torch.as_tensor(ni, device=dev)
        nd = dofmap.ndof
        self.ndof = nd
        # matrix contributions: element e, local (i, j), column-major in Aii
        cnt = ni_t * ni_t
        e_idx = torch.repeat_interleave(torch.arange(nel, device=dev), cnt)
        start = torch.cumsum(cnt, 0) - cnt
        local = torch.arange(int(cnt.sum()), device=dev) - start[e_idx]
        nie = ni_t[e_idx]
        i, j = local % nie, local // nie
        gi, gj = gid[e_idx, i], gid[e_idx, j]
        keep = (gi >= 0) & (gj >= 0)
        src = torch.as_tensor(batch.off_c_h, device=dev)[e_idx] + local
        key = (gi * nd + gj)[keep]
        src = src[keep]
        key, order = torch.sort(key, stable=True)
        self.a_src = src[order].contiguous()
        ukey, counts = torch.unique_consecutive(key, return_counts=True)
        self.a_ptr = torch.cat([torch.zeros(1, dtype=torch.int64, device=dev),
                                torch.cumsum(counts, 0)]).contiguous()
        rows, cols = ukey // nd, ukey % nd
        self.indptr = torch.cat([torch.zeros(1, dtype=torch.int64, device=dev),
                                 torch.cumsum(torch.bincount(rows, minlength=nd), 0)])
        self.indices = cols.contiguous()
        self.nnz = int(ukey.numel())
        # right-hand-side contributions: element e, local i at off_v[e] + i
        e_v = torch.repeat_interleave(torch.arange(nel, device=dev), ni_t)
        start_v = torch.cumsum(ni_t, 0) - ni_t
        li = torch.arange(int(ni_t.sum()), device=dev) - start_v[e_v]
        g = gid[e_v, li]
        keep = g >= 0
        srcv = (torch.as_tensor(batch.off_v_h, device=dev)[e_v] + li)[keep]
        g = g[keep]
        g, order = torch.sort(g, stable=True)
        self.b_src = srcv[order].contiguous()
        counts = torch.bincount(g, minlength=nd)
        self.b_ptr = torch.cat([torch.zeros(1, dtype=torch.int64, device=dev),
                                torch.cumsum(counts, 0)]).contiguous()
        # scatter map global -> element slots (interface part, internal order)
        self.slot_pos = (torch.as_tensor(batch.off_v_h, device=dev)[e_v] + li).contiguous()
        self.slot_gid = torch.where(gid[e_v, li] >= 0, gid[e_v, li],
                                    torch.full_like(li, nd)).contiguous()

    def _gather(self, ptr, src, packed, nout):
        import torch

        lib = _lib.load()
        out = torch.empty(nout, dtype=torch.float64, device=self.batch.device)
        st = torch.cuda.current_stream().cuda_stream
        rc = lib.hp3d_gpu_gather_sum(nout, ctypes.c_void_p(ptr.data_ptr()),
                                     ctypes.c_void_p(src.data_ptr()),
                                     ctypes.c_void_p(packed.data_ptr()),
                                     ctypes.c_void_p(out.data_ptr()), ctypes.c_void_p(st))
        check(rc, "hp3d_gpu_gather_sum")
        return out

    def matrix_values(self, aii):
        """CSR values of the global condensed matrix from packed Aii (device)."""
        return self._gather(self.a_ptr, self.a_src, aii, self.nnz)

    def rhs(self, bi):
        """Global condensed right-hand side from packed Bi slots (device)."""
        return self._gather(self.b_ptr, self.b_src, bi, self.ndof)

    def scatter(self, x):
        """Element dof slots (size_v, internal order) holding the interface
        values of the global vector x (0 at Dirichlet dofs; bubbles left 0)."""
        import torch

        xe = torch.cat([x.to(torch.float64), torch.zeros(1, dtype=torch.float64, device=x.device)])
        u = torch.zeros(self.batch.size_v, dtype=torch.float64, device=x.device)
        u[self.slot_pos] = xe[self.slot_gid]
        return u

    def csr_host(self, vals):
        import scipy.sparse as sp

        return sp.csr_matrix((vals.cpu().numpy(), self.indices.cpu().numpy(),
                              self.indptr.cpu().numpy()), shape=(self.ndof, self.ndof))


class Solver:
    """Sparse direct solve of the condensed global system (the reference
    hands it to MUMPS): SuperLU factorisation, reused for every solve."""

    def __init__(self, a_csr):
        import scipy.sparse.linalg as spla

        self.lu = spla.splu(a_csr.tocsc(), permc_spec="MMD_AT_PLUS_A")

    def __call__(self, b):
        import torch

        x = self.lu.solve(b.cpu().numpy())
        return torch.as_tensor(x, device=b.device)


class PCGSolver:
    """Conjugate gradients with the inverse diagonal as preconditioner, all
    on the device: one hp3d_gpu_csr_spmv launch and a few vector kernels per
    iteration, scalars kept on the device (a host synchronisation only every
    `check_every` iterations for the stopping test)."""

    def __init__(self, asm, vals, rtol=1e-12, maxiter=20000, check_every=25):
        import torch

        if asm.ndof >= 2 ** 31 - 1:
            raise Hp3dError("PCG: %d unknowns exceed the int32 column indices" % asm.ndof)
        self.n = asm.ndof
        self.indptr = asm.indptr.contiguous()
        self.indices = asm.indices.to(torch.int32).contiguous()
        self.vals = vals.contiguous()
        rows = torch.repeat_interleave(torch.arange(self.n, device=vals.device),
                                       self.indptr[1:] - self.indptr[:-1])
        diag = torch.zeros(self.n, dtype=torch.float64, device=vals.device)
        on = rows == asm.indices
        diag[rows[on]] = vals[on]
        self.dinv = 1.0 / diag
        self.rtol, self.maxiter, self.check_every = rtol, maxiter, check_every
        self.iterations = 0
        self.relres = None

    def matvec(self, x, y):
        import torch

        lib = _lib.load()
        st = torch.cuda.current_stream().cuda_stream
        p = lambda t: ctypes.c_void_p(t.data_ptr())
        rc = lib.hp3d_gpu_csr_spmv(self.n, p(self.indptr), p(self.indices), p(self.vals), p(x),
                                   p(y), ctypes.c_void_p(st))
        check(rc, "hp3d_gpu_csr_spmv")
        return y

    def __call__(self, b, x0=None, rtol=None):
        import torch

        rtol = self.rtol if rtol is None else rtol
        b = b.to(torch.float64)
        x = torch.zeros_like(b) if x0 is None else x0.clone()
        ap = torch.empty_like(b)
        r = b - self.matvec(x, ap) if x0 is not None else b.clone()
        z = self.dinv * r
        p = z.clone()
        rz = torch.dot(r, z)
        bnorm = float(torch.linalg.vector_norm(b))
        if bnorm == 0.0:
            return torch.zeros_like(b)
        it = 0
        while True:
            self.matvec(p, ap)
            alpha = rz / torch.dot(p, ap)
            x.add_(alpha * p)
            r.sub_(alpha * ap)
            it += 1
            if it % self.check_every == 0 or it >= self.maxiter:
                self.relres = float(torch.linalg.vector_norm(r)) / bnorm
                if self.relres <= rtol:
                    break
                if it >= self.maxiter:
                    raise Hp3dError("PCG did not reach rtol %.1e in %d iterations (residual %.2e)"
                                    % (rtol, it, self.relres))
            torch.mul(self.dinv, r, out=z)
            rz_new = torch.dot(r, z)
            p.mul_(rz_new / rz).add_(z)
            rz = rz_new
        self.iterations = it
        return x


def back_substitute(batch, out, u):
    """Bubble dofs of every element from its interface dofs (in place)."""
    import torch

    lib = _lib.load()
    st = torch.cuda.current_stream().cuda_stream
    for lo, cnt, _mnd, mnb in batch.bins:
        if mnb == 0:
            continue
        at = batch._ptrs(lo)
        rc = lib.hp3d_gpu_stc_bwd_batch(batch.space, cnt, at(batch.norder, 19),
                                        at(batch.off_s, 1), ctypes.c_void_p(out["aschur"].data_ptr()),
                                        at(batch.off_v, 1), ctypes.c_void_p(out["bschur"].data_ptr()),
                                        ctypes.c_void_p(u.data_ptr()), ctypes.c_void_p(st))
        check(rc, "hp3d_gpu_stc_bwd_batch")
    return u


def field(batch, mode, u, v=None, exact=0, with_r=False):
    """hp3d_gpu_field_batch over the batch (internal order).  Returns
    (out, r): out (nel, 2) for FIELD_ERROR else (nel,), r the element
    vectors (FIELD_L2) or None.  The kernel reports into batch.status
    (negative Jacobian / invalid orders), replacing the element kernels'
    codes."""
    import torch

    lib = _lib.load()
    dev = batch.device
    width = 2 if mode == FIELD_ERROR else 1
    out = torch.empty(batch.nel * width, dtype=torch.float64, device=dev)
    r = torch.zeros(batch.size_v, dtype=torch.float64, device=dev) if with_r else None
    p = lambda t: ctypes.c_void_p(t.data_ptr()) if t is not None else None
    st = torch.cuda.current_stream().cuda_stream
    rc = lib.hp3d_gpu_field_batch(batch.space, mode, batch.nel, p(batch.norder), p(batch.nori_e),
                                  p(batch.nori_f), p(batch.xnod), batch.ngeo, p(batch.coef),
                                  p(batch.off_v), p(u), p(v), int(exact), p(out), p(r),
                                  p(batch.status), ctypes.c_void_p(st))
    check(rc, "hp3d_gpu_field_batch")
    return (out.view(batch.nel, 2) if width == 2 else out), r


def dirichlet_lift(batch, out, ud):
    """Condensed element loads minus Aii ud on the device
    (hp3d_gpu_dirichlet_lift; ud packed like out['bi'], internal order)."""
    import torch

    lib = _lib.load()
    st = torch.cuda.current_stream().cuda_stream
    p = lambda t: ctypes.c_void_p(t.data_ptr())
    rc = lib.hp3d_gpu_dirichlet_lift(batch.space, batch.nel, p(batch.norder), p(batch.off_c),
                                     p(out["aii"]), p(batch.off_v), p(ud), p(out["bi"]),
                                     ctypes.c_void_p(st))
    check(rc, "hp3d_gpu_dirichlet_lift")


def pack_element_vectors(batch, vec):
    """(nel, width) per-element interface vectors (caller order) -> a packed
    device vector of batch.size_v slots (internal order, 0 beyond ni)."""
    import torch

    v = np.zeros(batch.size_v)
    ni = batch._ni
    e_v = np.repeat(np.arange(batch.nel), ni)
    li = np.arange(int(ni.sum())) - np.repeat(np.cumsum(ni) - ni, ni)
    v[batch.off_v_h[e_v] + li] = np.asarray(vec)[batch.perm[e_v], li]
    return torch.as_tensor(v, device=batch.device)


def solve_problem(mesh, space=H1, load=None, exact=None, precision=0, dirichlet=True,
                  solver="direct", rtol=1e-12, boundary=None, numbering="blocks"):
    """Element matrices, condensation, assembly, solve, back substitution and
    element errors for a mesh with Dirichlet data on the whole boundary:
    homogeneous, or (boundary = a selector of hp3d_amd.dirichlet.exact_h1, or
    a callable x -> (u, grad u)) the projection-based interpolant of the
    exact solution on the boundary nodes (update_Ddof), lifted into the
    condensed loads as celem_systemI does (celem_systemI.F90:794-808).

    Returns a dict with the global solution x, the element dof vectors u
    (internal order), the squared error and exact norm sums (errorH, rnormH
    as element_error accumulates them), and the objects built on the way."""
    import torch

    from ._lib import LOAD_CURL_SIN, LOAD_SINSINSIN

    if load is None:
        load = LOAD_SINSINSIN if space == H1 else LOAD_CURL_SIN
    exact = load if exact is None else exact
    batch = ElementBatch.from_mesh(space, mesh, load=load)
    out = batch.run(precision)
    batch.check_status()
    dm = DofMap(mesh, space, dirichlet, numbering=numbering)
    ud = None
    if boundary is not None and dirichlet:
        from .dirichlet import dirichlet_values

        ud = pack_element_vectors(batch, dirichlet_values(mesh, dm, boundary))
        dirichlet_lift(batch, out, ud)
    asm = Assembly(batch, dm)
    vals = asm.matrix_values(out["aii"])
    b = asm.rhs(out["bi"])
    solver = Solver(asm.csr_host(vals)) if solver == "direct" else PCGSolver(asm, vals, rtol=rtol)
    x = solver(b)
    ue = asm.scatter(x)
    if ud is not None:
        ue = ue + ud
    u = back_substitute(batch, out, ue)
    errs, _ = field(batch, FIELD_ERROR, u, exact=exact)
    torch.cuda.synchronize()
    e = errs.cpu().numpy()
    return {"x": x, "u": u, "error2": float(e[:, 0].sum()), "norm2": float(e[:, 1].sum()),
            "batch": batch, "out": out, "dofmap": dm, "assembly": asm, "solver": solver,
            "values": vals, "rhs": b}
