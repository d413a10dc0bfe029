"""Adjoint products and the autograd wrapper on the GPU: the adjoint against the oracle (the
same products through the CPU restatement), gradients against an FP64 direct-sum reference
(torch autograd through the dense sum) and central finite differences."""
import numpy as np
import pytest
import torch

import ifgf_amd
from ifgf_amd.autograd import helmholtz_apply
from conftest import cdens, perturbed_sphere, sphere
from oracle import pyoracle as po

pytestmark = pytest.mark.gpu


def rel(a, b):
    return float(np.linalg.norm(a - b) / np.linalg.norm(b))


@pytest.mark.parametrize("prec,tol", [("fp64", 1e-11), ("mixed", 2e-5)])
def test_adjoint_matches_oracle(gpu, prec, tol):
    """A^H g = conj(B conj g) (B: targets -> sources) for the potentials, minus that for the
    gradient kernel; same points reuse the plan, other targets build the transposed plan."""
    x = sphere(5000, 61)
    y = sphere(3000, 62) * 1.7
    g_x = cdens(5000, 63)
    g_y = cdens(3000, 64)
    cases = [(po.HELMHOLTZ, 9.0, x, x, g_x), (po.HELMHOLTZ, 9.0, x, y, g_y),
             (po.MODHELMHOLTZ, 0.5 - 7.0j, x, y, g_y), (po.LAPLACE, 0.0, x, y, g_y)]
    for kind, k, src, tgt, g in cases:
        print(f"adjoint kind={kind} k={k} N={len(src)} M={len(tgt)} prec={prec}", flush=True)
        order = {po.MODHELMHOLTZ: 8, po.LAPLACE: 7}.get(kind, 6)  # fine order = order - 3
        orc = po.Oracle(kind, k, 64, order, 1, -1)
        orc.init(tgt, src)
        ref = np.conj(orc.mult(np.conj(g)))
        if kind == po.LAPLACE:
            op = ifgf_amd.LaplaceIfgfOperator(64, order, 1, -1, precision=prec)
            g = g.real
            ref = ref.real
        elif kind == po.MODHELMHOLTZ:
            op = ifgf_amd.MofifiedHelmholtzIfgfOperator(k, 64, order, 1, -1, precision=prec)
        else:
            op = ifgf_amd.HelmholtzIfgfOperator(k, 64, order, 1, -1, precision=prec)
        op.init(src.T, tgt.T)
        out = op.mult_adjoint(g)
        assert out.shape == (len(src),)
        assert rel(out, ref) < tol, (kind, rel(out, ref))
    # gradient kernel: odd in x - y
    for dx in (0, 2):
        orc = po.Oracle(po.GRAD, 1.0 + 8.0j, 64, 6, 1, -1, dx=dx)
        orc.init(y, x)
        ref = -np.conj(orc.mult(np.conj(g_y)))
        op = ifgf_amd.GradHelmholtzIfgfOperator(1.0 + 8.0j, 64, 6, 1, -1, precision=prec)
        op.setDx(dx)
        op.init(x.T, y.T)
        assert rel(op.mult_adjoint(g_y), ref) < tol
    # no adjoint for the kernels with source normals
    n = x / np.linalg.norm(x, axis=1)[:, None]
    dl = ifgf_amd.DoubleLayerHelmholtzIfgfOperator(0.5 + 9.0j, 40, 8, 1, -1, precision=prec)
    dl.init(x.T, x.T, n.T)
    with pytest.raises(ifgf_amd.IfgfError):
        dl.mult_adjoint(g_x)


def direct_loss(sig, pts, b, kappa):
    """||A sigma - b||^2 with the dense FP64 single-layer matrix (G(0) = 0)."""
    diff = pts[:, None, :] - pts[None, :, :]
    d2 = (diff * diff).sum(-1)
    eye = torch.eye(len(pts), dtype=torch.bool, device=pts.device)
    dd = torch.sqrt(torch.where(eye, torch.ones_like(d2), d2))  # no sqrt'(0) on the diagonal
    G = torch.where(eye, torch.zeros((), dtype=torch.complex128, device=pts.device),
                    torch.exp(1j * kappa * dd) / (4 * np.pi * dd))
    r = G @ sig - b
    return (r.conj() * r).real.sum()


@pytest.mark.parametrize("prec", ["fp64", "mixed"])
def test_gradients_vs_direct_sum_and_fd(gpu, prec):
    """Small perturbed sphere: IFGF gradients (sigma and coordinates) against the FP64 dense
    reference gradient, and directional derivatives against central differences (h = 1e-6)."""
    dev = torch.device("cuda:0")
    N, kappa = 3000, 8.0
    x0 = torch.as_tensor(perturbed_sphere(N, 71), device=dev)
    sig0 = torch.as_tensor(cdens(N, 72), device=dev)
    b = torch.as_tensor(cdens(N, 73), device=dev)
    order = 8
    # IFGF
    sig = sig0.clone().requires_grad_(True)
    x = x0.clone().requires_grad_(True)
    u = helmholtz_apply(sig, x, kappa=kappa, order=order, precision=prec)
    L = ((u - b).conj() * (u - b)).real.sum()
    L.backward()
    # dense FP64 reference
    sr = sig0.clone().requires_grad_(True)
    xr = x0.clone().requires_grad_(True)
    Lr = direct_loss(sr, xr, b, kappa)
    Lr.backward()
    es = rel(sig.grad.cpu().numpy(), sr.grad.cpu().numpy())
    ex = rel(x.grad.cpu().numpy(), xr.grad.cpu().numpy())
    el = abs(L.item() - Lr.item()) / Lr.item()
    print(f"{prec}: loss {el:.2e} dL/dsigma {es:.2e} dL/dx {ex:.2e}", flush=True)
    assert el < 1e-4 and es < 1e-4 and ex < 1e-3
    # central differences of the dense loss along random directions (h = 1e-6, FP64)
    rng = np.random.default_rng(74)
    h = 1e-6
    for i in range(8):
        dx = torch.as_tensor(rng.standard_normal((N, 3)), device=dev)
        ds = torch.as_tensor(cdens(N, 75 + i), device=dev)
        with torch.no_grad():
            fd_x = (direct_loss(sig0, x0 + h * dx, b, kappa) -
                    direct_loss(sig0, x0 - h * dx, b, kappa)).item() / (2 * h)
            fd_s = (direct_loss(sig0 + h * ds, x0, b, kappa) -
                    direct_loss(sig0 - h * ds, x0, b, kappa)).item() / (2 * h)
        an_x = float((x.grad * dx).sum())
        an_s = float((sig.grad.conj() * ds).real.sum())
        assert abs(an_x - fd_x) <= 2e-3 * abs(fd_x) + 1e-6 * float(x.grad.norm() * dx.norm())
        assert abs(an_s - fd_s) <= 2e-4 * abs(fd_s) + 1e-6 * float(sig.grad.norm() * ds.norm())


def test_gradients_targets_other_than_sources(gpu):
    """Separate target points: dL/dy from D sigma, dL/dx from the transposed gradient operator."""
    dev = torch.device("cuda:0")
    xs = torch.as_tensor(sphere(2500, 81), device=dev).requires_grad_(True)
    ys = torch.as_tensor(sphere(1500, 82) * 1.5, device=dev).requires_grad_(True)
    sig = torch.as_tensor(cdens(2500, 83), device=dev).requires_grad_(True)
    b = torch.as_tensor(cdens(1500, 84), device=dev)
    kappa = 6.0
    u = helmholtz_apply(sig, xs, ys, kappa=kappa, order=8, precision="fp64")
    L = ((u - b).conj() * (u - b)).real.sum()
    L.backward()
    xr = xs.detach().clone().requires_grad_(True)
    yr = ys.detach().clone().requires_grad_(True)
    sr = sig.detach().clone().requires_grad_(True)
    diff = yr[:, None, :] - xr[None, :, :]
    d = torch.sqrt((diff * diff).sum(-1))
    G = torch.exp(1j * kappa * d) / (4 * np.pi * d)
    r = G @ sr - b
    Lr = (r.conj() * r).real.sum()
    Lr.backward()
    # every target lies in the far field of most sources here, so the IFGF interpolation error
    # (order 8) shows in all three gradients at the level of the forward error
    eu = rel(u.detach().cpu().numpy(), (G @ sr).detach().cpu().numpy())
    errs = [rel(a.cpu().numpy(), bb.cpu().numpy())
            for a, bb in ((sig.grad, sr.grad), (xs.grad, xr.grad), (ys.grad, yr.grad))]
    print(f"other targets: forward {eu:.2e} dL/dsigma {errs[0]:.2e} dL/dx {errs[1]:.2e} "
          f"dL/dy {errs[2]:.2e}", flush=True)
    assert eu < 2e-3 and max(errs) < 5e-3


def test_configuration4_full_size(gpu):
    """N=131,072 perturbed sphere, kappa=40 (mixed form): the backward's products against FP64
    direct sums on 256 sampled points, and 64 central-difference directions in sigma."""
    dev = torch.device("cuda:0")
    N, kappa = 131072, 40.0
    pts = perturbed_sphere(N, 1004)
    x0 = torch.as_tensor(pts, device=dev)
    sig0 = torch.as_tensor(cdens(N, 2004), device=dev)
    b = torch.as_tensor(cdens(N, 2104), device=dev)
    sig = sig0.clone().requires_grad_(True)
    x = x0.clone().requires_grad_(True)
    print("configuration 4: forward", flush=True)
    u = helmholtz_apply(sig, x, kappa=kappa, order=6, precision="mixed")
    r = u - b
    L = (r.conj() * r).real.sum()
    print("configuration 4: backward", flush=True)
    L.backward()
    g = (2 * r).detach()
    idx = torch.as_tensor(np.random.default_rng(3004).choice(N, 256, replace=False), device=dev)
    # direct sums at the sampled points: u, A^H g, D sigma, D conj(g)
    y = x0[idx]
    diff = y[:, None, :] - x0[None, :, :]
    d = torch.sqrt((diff * diff).sum(-1))
    self_ = d == 0
    dd = torch.where(self_, torch.ones_like(d), d)
    e = torch.exp(1j * kappa * dd) / (4 * np.pi * dd)
    G = torch.where(self_, torch.zeros_like(e), e)
    u_ref = G @ sig0
    ah_ref = torch.conj(G) @ g  # G symmetric: (A^H g)_k = sum_j conj(G_kj) g_j
    dG = G * (1j * kappa - 1.0 / dd)  # dG/dd
    gx = torch.zeros((256, 3), dtype=torch.float64, device=dev)
    for c in range(3):
        Dc = dG * diff[:, :, c] / dd  # dG(y - x)/dy_c
        gx[:, c] = (torch.conj(g[idx]) * (Dc @ sig0)).real + \
            (sig0[idx] * (Dc @ torch.conj(g))).real
    eu = rel(u.detach()[idx].cpu().numpy(), u_ref.cpu().numpy())
    es = rel(sig.grad[idx].cpu().numpy(), ah_ref.cpu().numpy())
    ex = rel(x.grad[idx].cpu().numpy(), gx.cpu().numpy())
    print(f"configuration 4: rel. error forward {eu:.2e} dL/dsigma {es:.2e} dL/dx {ex:.2e}",
          flush=True)
    assert eu < 3e-3 and es < 3e-3 and ex < 1e-2
    # 64 central differences along random sigma directions of the loss with the FP64 IFGF
    # operator: the loss is quadratic in sigma, so the quotient is exact up to rounding; the
    # backward's IFGF adjoint equals the adjoint of the IFGF forward up to the IFGF error
    op = ifgf_amd.HelmholtzIfgfOperator(kappa, 64, 6, 1, -1, precision="fp64")
    op.init(pts.T, pts.T)
    h = 1e-6
    rng = np.random.default_rng(3104)
    fds, ans = [], []
    for i in range(64):
        ds = torch.as_tensor(rng.standard_normal(N) + 1j * rng.standard_normal(N), device=dev)
        lp = op.mult(sig0 + h * ds) - b
        lm = op.mult(sig0 - h * ds) - b
        fds.append(((lp.conj() * lp).real.sum() - (lm.conj() * lm).real.sum()).item() / (2 * h))
        ans.append(float((sig.grad.conj() * ds).real.sum()))
    # l2 over the 64 directional derivatives (a single random projection can be near zero)
    e_dir = rel(np.array(ans), np.array(fds))
    print(f"configuration 4: 64 directional derivatives vs central differences {e_dir:.2e}",
          flush=True)
    assert e_dir < 5e-3


def direct_matvec(X, sig, kappa, chunk=256):
    """u_j = sum_{i != j} e^{i kappa |x_j - x_i|} / (4 pi |x_j - x_i|) sigma_i, FP64 with exact
    coordinate differences (torch on the GPU, chunks of targets)."""
    n = X.shape[0]
    sr, si = sig.real.contiguous(), sig.imag.contiguous()
    ur = torch.empty(n, dtype=torch.float64, device=X.device)
    ui = torch.empty_like(ur)
    for j0 in range(0, n, chunk):
        diff = X[j0:j0 + chunk, None, :] - X[None, :, :]
        d = torch.sqrt((diff * diff).sum(-1))
        del diff
        inv = torch.where(d > 0, 1.0 / (4 * np.pi * torch.where(d > 0, d, 1.0)), 0.0)
        c = torch.cos(kappa * d) * inv
        s = torch.sin(kappa * d) * inv
        ur[j0:j0 + chunk] = c @ sr - s @ si
        ui[j0:j0 + chunk] = c @ si + s @ sr
    return torch.complex(ur, ui)


def kernel_block(Y, X, kappa):
    """e^{i kappa |y - x|} / (4 pi |y - x|) for every (y, x) pair, 0 where y = x (FP64)."""
    diff = Y[:, None, :] - X[None, :, :]
    d = torch.sqrt((diff * diff).sum(-1))
    dd = torch.where(d > 0, d, 1.0)
    e = torch.exp(1j * kappa * dd) / (4 * np.pi * dd)
    return torch.where(d > 0, e, torch.zeros((), dtype=e.dtype, device=e.device))


def test_configuration4_coordinate_directions(gpu):
    """N=131,072 perturbed sphere, kappa=40: the backward's gradient with respect to the point
    coordinates (mixed form, the gradient-kernel IFGF operators) along 64 random unit
    directions v in R^{3N}, each moving 512 random points, against central differences
    (h = 1e-6, FP64) of the exact loss L(x) = ||A(x) sigma - b||^2.  A(x +- h v) sigma differs
    from the FP64 direct sum A(x) sigma only in the rows and columns of the moved points, so
    L(x + h v) - L(x - h v) = sum_j 2 Re(conj(r_j) (D+_j - D-_j)) + |D+_j|^2 - |D-_j|^2 with
    D+-  those exact changes (FP64 direct sums) and r = A(x) sigma - b."""
    dev = torch.device("cuda:0")
    N, kappa, h, M = 131072, 40.0, 1e-6, 512
    pts = perturbed_sphere(N, 1004)
    x0 = torch.as_tensor(pts, device=dev)
    sig0 = torch.as_tensor(cdens(N, 2004), device=dev)
    b = torch.as_tensor(cdens(N, 2104), device=dev)
    x = x0.clone().requires_grad_(True)
    print("configuration 4 (coordinates): forward and backward", flush=True)
    u = helmholtz_apply(sig0, x, kappa=kappa, order=6, precision="mixed")
    r = u - b
    ((r.conj() * r).real.sum()).backward()
    gx = x.grad.detach()
    r0 = direct_matvec(x0, sig0, kappa) - b
    print("configuration 4 (coordinates): direct sum", flush=True)

    def change(S, XS):
        """A(x') sigma - A(x) sigma with the points S moved to XS (all other points fixed)."""
        xp = x0.clone()
        xp[S] = XS
        D = torch.zeros(N, dtype=torch.complex128, device=dev)
        for j0 in range(0, N, 16384):  # every target, the moved sources
            K1 = kernel_block(xp[j0:j0 + 16384], XS, kappa)
            K0 = kernel_block(x0[j0:j0 + 16384], x0[S], kappa)
            D[j0:j0 + 16384] = (K1 - K0) @ sig0[S]
        # the moved targets' rows against the sources that did not move
        rest = torch.ones(N, dtype=torch.bool, device=dev)
        rest[S] = False
        K1 = kernel_block(XS, x0[rest], kappa)
        K0 = kernel_block(x0[S], x0[rest], kappa)
        D[S] += (K1 - K0) @ sig0[rest]
        return D

    rng = np.random.default_rng(3204)
    fds, ans = [], []
    for i in range(64):
        S = torch.as_tensor(np.sort(rng.choice(N, M, replace=False)), device=dev)
        v = rng.standard_normal((M, 3))
        v = torch.as_tensor(v / np.linalg.norm(v), device=dev)
        dp = change(S, x0[S] + h * v)
        dm = change(S, x0[S] - h * v)
        diff = (2 * (r0.conj() * (dp - dm)).real + (dp.conj() * dp).real
                - (dm.conj() * dm).real).sum()
        fds.append(float(diff) / (2 * h))
        ans.append(float((gx[S] * v).sum()))
        if i % 16 == 15:
            print(f"configuration 4 (coordinates): {i + 1} directions", flush=True)
    e_dir = rel(np.array(ans), np.array(fds))
    print(f"configuration 4: 64 coordinate directions vs central differences {e_dir:.2e}",
          flush=True)
    assert e_dir < 1e-3
