"""Differentiable IFGF Helmholtz single-layer matvec (torch.autograd.Function).

    u = helmholtz_apply(sigma, srcs, tgts=None, kappa=40.0)   # u_j = sum_i G(y_j - x_i) sigma_i

with G(d) = e^{i kappa |d|} / (4 pi |d|) and G(0) = 0 (helmholtz_ifgf.hpp:37-41).  sigma is a
complex (N,) tensor, srcs (N, 3) and tgts (M, 3) float64 tensors (tgts=None: the targets are
the sources).  The reference has no derivatives; this wrapper follows torch's convention for
complex outputs (grad_output g = dL/dRe u + i dL/dIm u):

    dL/dsigma = A^H g                                    (ifgf_mult_adjoint)
    dL/dy_k   = Re(conj(g_k) [D sigma]_k)                 (target moves)
    dL/dx_k   = Re(sigma_k [D' conj(g)]_k)                (source moves)

where D = dG/dy (target derivative, the gradient kernel of grad_helmholtz_ifgf.hpp) from the
sources to the targets and D' the same operator from the targets to the sources (the
gradient of G is odd in y - x).  With tgts=None both terms add up on the one point set and
one plan serves both products.  Every backward recomputes the interpolants (nothing from the
forward pass is stored but the operator), and every forward builds the plan for the current
points: the tree and cone sets change with the coordinates.
"""
import numpy as np
import torch

from . import pyifgf


def _np_points(p):
    return np.ascontiguousarray(p.detach().to(torch.float64).cpu().numpy())


class _HelmholtzApply(torch.autograd.Function):
    @staticmethod
    def forward(ctx, sigma, srcs, tgts, kappa, leaf, order, n_elem, tol, precision):
        same = tgts is None
        x = _np_points(srcs)
        y = x if same else _np_points(tgts)
        op = pyifgf.HelmholtzIfgfOperator(kappa, leaf, order, n_elem, tol, precision=precision)
        op.init(x.T, y.T)
        s = sigma.detach().to(torch.complex128)
        u = op.mult(s) if s.is_cuda else torch.from_numpy(op.mult(s.numpy()))
        ctx.save_for_backward(s)
        ctx.op = op
        ctx.same = same
        ctx.xy = (x, y)
        ctx.args = (kappa, leaf, order, n_elem, tol, precision)
        return u.to(sigma.dtype if sigma.is_complex() else torch.complex128)

    @staticmethod
    def backward(ctx, g):
        (s,) = ctx.saved_tensors
        dev = s.device
        g = g.detach().to(torch.complex128).to(dev)
        kappa, leaf, order, n_elem, tol, precision = ctx.args
        x, y = ctx.xy
        grad_sigma = grad_x = grad_y = None

        def run(op, w):
            return op.mult(w) if w.is_cuda else torch.from_numpy(op.mult(w.numpy()))

        if ctx.needs_input_grad[0]:
            gs = ctx.op.mult_adjoint(g) if g.is_cuda else torch.from_numpy(
                ctx.op.mult_adjoint(g.numpy()))
            grad_sigma = gs.to(dev)
        want_x = ctx.needs_input_grad[1]
        want_y = (not ctx.same) and ctx.needs_input_grad[2]
        if want_x or want_y or (ctx.same and ctx.needs_input_grad[1]):
            # the gradient kernel: e^{-k d} with k = -i kappa
            def grad_op(a, b):
                op = pyifgf.GradHelmholtzIfgfOperator(-1j * kappa, leaf, order, n_elem, tol,
                                                      precision=precision)
                op.init(a.T, b.T)
                return op
            fwd = grad_op(x, y)
            bwd = fwd if ctx.same else (grad_op(y, x) if want_x else None)
            gc = torch.conj(g)
            tx = torch.zeros((x.shape[0], 3), dtype=torch.float64, device=dev)
            ty = torch.zeros((y.shape[0], 3), dtype=torch.float64, device=dev)
            for d in range(3):
                fwd.setDx(d)
                if ctx.same:
                    r = run(fwd, torch.stack([s, gc], dim=1)).to(dev)
                    tx[:, d] = (torch.conj(g) * r[:, 0]).real + (s * r[:, 1]).real
                    continue
                if want_y:
                    ty[:, d] = (torch.conj(g) * run(fwd, s).to(dev)).real
                if want_x:
                    bwd.setDx(d)
                    tx[:, d] = (s * run(bwd, gc).to(dev)).real
            grad_x = tx if want_x or ctx.same else None
            grad_y = ty if want_y else None
        return grad_sigma, grad_x, grad_y, None, None, None, None, None, None


def helmholtz_apply(sigma, srcs, tgts=None, kappa=1.0, leaf=64, order=6, n_elem=1, tol=-1,
                    precision="mixed"):
    """u = A sigma with A the IFGF Helmholtz single-layer operator on (srcs -> tgts);
    differentiable in sigma, srcs and tgts.  Point tensors are (N, 3) float64."""
    return _HelmholtzApply.apply(sigma, srcs, tgts, float(kappa), int(leaf), int(order),
                                 int(n_elem), float(tol), precision)
