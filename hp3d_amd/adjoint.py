"""Per-element coefficient gradients by the adjoint method.

Problem: -div(exp(theta_e) grad u) = f on the mesh, u = 0 on the boundary,
objective J(theta) = 1/2 ||u_h(theta) - u_t||^2_L2 with u_t a fixed discrete
field (the solution for a second theta).  With K(theta) u = F:

    K lambda = M (u - u_t)                 (adjoint solve, same operator)
    dJ/dtheta_e = -lambda_e^T (dK_e/dtheta_e) u_e = -lambda_e^T K_e u_e
                = -exp(theta_e) int_e grad lambda_h . grad u_h

The last integral is the hand-written HIP field kernel
(hp3d_gpu_field_batch, mode HP3D_FIELD_ADJOINT): it evaluates grad u_h and
grad lambda_h at the element's quadrature points from the element dof
vectors and never forms dK_e.  Every step runs on the device through the C
ABI: element kernels with condensation, the gather-sum assembly, the solve,
hp3d_gpu_stc_bwd_batch, and the field kernel for J, M (u - u_t) and the
gradient.  The adjoint right-hand side is condensed with the same element
matrices (hp3d_gpu_stc_batch on the stored K), and the adjoint solve reuses
the forward operator (factorisation or preconditioner).

hp3D itself has no adjoint; the element-level pieces it would use are
elem_opt (POISSON/GALERKIN/elem_opt.F90:22-141) and stc (modules/stc.F90).
"""
import numpy as np

from ._lib import FIELD_ADJOINT, FIELD_L2, FP64, H1, LOAD_SINSINSIN
from .elem import ElementBatch
from .solve import Assembly, DofMap, PCGSolver, Solver, back_substitute, field


class AdjointProblem:
    """Holds the mesh-dependent state (element batch, numbering, assembly
    pattern) and the target field; `objective` and `gradient` do the work.

    The element matrices, Schur factors and coefficients of the latest
    forward solve live in buffers shared by every call (`out`, `batch.coef`).
    A state returned by `objective` records its theta and the generation of
    those buffers; `gradient` on a state whose buffers a later forward has
    overwritten first recomputes them from the state's theta (the kernels
    are deterministic, so the restored K and Schur factors are the same
    bits), so two forwards before a backward still give each its own
    gradient."""

    def __init__(self, mesh, theta_target, load=LOAD_SINSINSIN, precision=FP64,
                 solver="direct", rtol=1e-13):
        self.mesh = mesh
        self.precision = precision
        self.solver_kind = solver
        self.rtol = rtol
        self.batch = ElementBatch.from_mesh(H1, mesh, load=load)
        self.dofmap = DofMap(mesh, H1)
        self.asm = Assembly(self.batch, self.dofmap)
        self.out = self.batch.alloc_outputs(precision, want_k=True, want_stc=True)
        self._gen = 0
        self.u_target = self._solve_state(np.asarray(theta_target, dtype=np.float64))[0].clone()

    def _set_theta(self, theta):
        import torch

        th = torch.as_tensor(np.asarray(theta, dtype=np.float64)[self.batch.perm],
                             device=self.batch.device)
        self.batch.coef.copy_(torch.exp(th))

    def _element_pass(self, theta):
        """Element matrices, condensation and coefficients of theta into the
        shared buffers (a new generation of them)."""
        self._set_theta(theta)
        self.batch.run(self.precision, want_k=True, want_stc=True, out=self.out)
        self.batch.check_status()
        self._gen += 1
        return self._gen

    def _solve_state(self, theta, x0=None):
        """u(theta): element dof vectors (internal order), plus the solver."""
        self._element_pass(theta)
        vals = self.asm.matrix_values(self.out["aii"])
        b = self.asm.rhs(self.out["bi"])
        if self.solver_kind == "direct":
            solver = Solver(self.asm.csr_host(vals))
            x = solver(b)
        else:
            solver = PCGSolver(self.asm, vals, rtol=self.rtol)
            x = solver(b, x0=x0)
        u = back_substitute(self.batch, self.out, self.asm.scatter(x))
        return u, solver, x

    def objective(self, theta, x0=None):
        """J(theta) and the state needed for its gradient."""
        theta = np.array(theta, dtype=np.float64)   # the state's own copy
        u, solver, x = self._solve_state(theta, x0)
        d = u - self.u_target
        jl, r = field(self.batch, FIELD_L2, d, with_r=True)
        J = 0.5 * float(jl.sum())
        return J, {"u": u, "r": r, "solver": solver, "x": x, "theta": theta, "gen": self._gen}

    def gradient(self, state):
        """dJ/dtheta_e for every element, caller order (numpy)."""
        import torch

        b = self.batch
        if state["gen"] != self._gen:
            # a later forward overwrote K, the Schur factors and coef: restore
            # them for this state (u, r and the solver are the state's own)
            state["gen"] = self._element_pass(state["theta"])
        # condense the adjoint right-hand side M (u - u_t) with the stored K
        adj = {"k": self.out["k"], "b": state["r"],
               "aii": torch.empty_like(self.out["aii"]), "bi": torch.empty_like(self.out["bi"]),
               "aschur": torch.empty_like(self.out["aschur"]),
               "bschur": torch.empty_like(self.out["bschur"])}
        b.condense(adj, self.precision)
        rhs = self.asm.rhs(adj["bi"])
        lam_i = state["solver"](rhs)
        lam = back_substitute(b, {"aschur": self.out["aschur"], "bschur": adj["bschur"]},
                              self.asm.scatter(lam_i))
        g, _ = field(b, FIELD_ADJOINT, state["u"], v=lam)
        torch.cuda.synchronize()
        return g.cpu().numpy()[b.iperm], lam


class _Objective:
    """torch.autograd.Function wrapper built lazily (torch imported on use)."""
    fn = None

    @classmethod
    def get(cls):
        if cls.fn is None:
            import torch

            class Hp3dObjective(torch.autograd.Function):
                @staticmethod
                def forward(ctx, theta, problem):
                    J, state = problem.objective(theta.detach().cpu().numpy())
                    ctx.problem = problem
                    ctx.state = state
                    return torch.tensor(J, dtype=torch.float64, device=theta.device)

                @staticmethod
                def backward(ctx, grad_out):
                    g, _ = ctx.problem.gradient(ctx.state)
                    g = torch.as_tensor(g, dtype=grad_out.dtype, device=grad_out.device)
                    return grad_out * g, None

            cls.fn = Hp3dObjective
        return cls.fn


def objective(theta, problem):
    """J(theta) as a differentiable torch scalar (backward runs the adjoint)."""
    return _Objective.get().apply(theta, problem)
