"""torch.autograd.Function over the C-ABI (test harness for the adjoint).

forward:  x = A(vals)^{-1} b with the context's Krylov method and
          preconditioner (xs_mat_update_values + xs_solve_device).
backward: the forward's values are put back into the context
          (xs_mat_update_values: a later forward may have replaced them), then
          A^T lam = dL/dx by xs_solve_transpose_device (transposed SpMV,
          transposed ILU triangular solves), dL/db = lam and
          dL/dA_ij = -lam_i x_j on the sparsity pattern (xs_grad_values_device).
No reference counterpart.
"""
import torch

from . import XSolver, XSolverError


class SparseSolve(torch.autograd.Function):
    @staticmethod
    def forward(ctx, vals, b, solver):
        vals = vals.detach().contiguous()
        b = b.detach().contiguous()
        torch.cuda.synchronize()
        solver_lib_update(solver, vals)
        x = torch.zeros_like(b)
        st = solver.solve_device(b, x)
        if st != 0:
            raise XSolverError(st, f"forward solve did not converge (status {st})")
        ctx.solver = solver
        ctx.save_for_backward(x, vals)
        return x

    @staticmethod
    def backward(ctx, gx):
        # the values of this forward's A: the shared context may hold those of
        # a later forward by now
        x, vals = ctx.saved_tensors
        solver = ctx.solver
        gx = gx.detach().contiguous()
        lam = torch.zeros_like(gx)
        torch.cuda.synchronize()
        solver_lib_update(solver, vals)
        st = solver.solve_transpose_device(gx, lam)
        if st != 0:
            raise XSolverError(st, f"adjoint solve did not converge (status {st})")
        gvals = torch.empty(solver.nnz, dtype=torch.float64, device=x.device)
        solver.grad_values(lam, x, gvals, torch.cuda.current_stream().cuda_stream)
        return gvals, lam, None


def solver_lib_update(solver, vals):
    from . import _check, _ptr, lib
    _check(lib().xs_mat_update_values(solver.h, _ptr(vals), 1), "xs_mat_update_values")
    return 0


def make_solver(rowptr, colidx, vals, ksp="cg", pc="ilu0", tol=1e-12, maxiter=5000, **opts):
    """A solver context holding the pattern (values are replaced per forward)."""
    S = XSolver()
    S.mat_setup(rowptr, colidx, vals)
    S.ksp_settype(ksp)
    S.ksp_setoption(tol, maxiter)
    S.pc_settype(pc)
    for k, v in opts.items():
        S.set_option(k, v)
    S.nnz = len(vals)
    return S


def sparse_solve(vals, b, solver):
    return SparseSolve.apply(vals, b, solver)
