"""The headline size (N = 1,048,576 on the unit sphere, kappa = 100, order 6, leaf 64) through
size-independent properties: linearity, agreement of the mixed and FP64 forms, and the
error against an FP64 direct sum on sampled targets."""
import numpy as np
import pytest

import ifgf_amd
from conftest import cdens, sphere

pytestmark = pytest.mark.gpu


def test_full_size_properties(gpu):
    import torch

    from bench import direct_sum_torch

    N = 1 << 20
    x = sphere(N, 1002)
    a, b = cdens(N, 2002), cdens(N, 2003)
    dev = torch.device("cuda:0")
    op = ifgf_amd.HelmholtzIfgfOperator(100.0, 64, 6, 1, -1, precision="mixed")
    op.init(x.T, x.T)
    print("full size plan:", op.plan_counts(), flush=True)
    ta, tb = (torch.as_tensor(v, device=dev) for v in (a, b))
    ua, ub = op.mult(ta), op.mult(tb)
    uab = op.mult(2 * ta - 1j * tb)
    lin = torch.linalg.norm(uab - (2 * ua - 1j * ub)) / torch.linalg.norm(uab)
    assert lin.item() < 1e-5, lin.item()

    op64 = ifgf_amd.HelmholtzIfgfOperator(100.0, 64, 6, 1, -1, precision="fp64")
    op64.init(x.T, x.T)
    u64 = op64.mult(ta)
    diff = (torch.linalg.norm(ua - u64) / torch.linalg.norm(u64)).item()
    assert diff < 1e-5, diff

    idx = np.random.default_rng(3002).choice(N, 256, replace=False)
    ref = direct_sum_torch(x, a, idx, 100.0, dev)
    err = np.linalg.norm(ua.cpu().numpy()[idx] - ref) / np.linalg.norm(ref)
    err64 = np.linalg.norm(u64.cpu().numpy()[idx] - ref) / np.linalg.norm(ref)
    print(f"full size: rel. error vs direct sum mixed {err:.3e} fp64 {err64:.3e}; "
          f"linearity {lin.item():.2e}; mixed vs fp64 {diff:.2e}", flush=True)
    # order 6 leaves the method's own interpolation error at ~1e-3 (measured with the oracle
    # at the same points per wavelength); FP32 arithmetic adds ~1e-6 on top
    assert err < 3e-3 and abs(err - err64) < 1e-4


def test_headline_tolerance_path(gpu):
    """The benchmarked workload: N = 1,048,576 on the unit sphere, kappa = 100, order 6,
    leaf 64, tolerance 1e-5 (the order estimate raises the orders to (6,8,7), fine (3,5,4)),
    mixed form: relative error <= 1e-4 against an FP64 direct sum on 1,000 sampled targets."""
    import torch

    from bench import direct_sum_torch

    N = 1 << 20
    x = sphere(N, 1002)
    w = cdens(N, 2002)
    dev = torch.device("cuda:0")
    print("headline: N=1048576 kappa=100 order=6 tol=1e-5 mixed", flush=True)
    op = ifgf_amd.HelmholtzIfgfOperator(100.0, 64, 6, 1, 1e-5, precision="mixed")
    op.init(x.T, x.T)
    o = op.plan_orders()
    assert o["base"].tolist() == [6, 8, 7] and o["high"].tolist() == [6, 8, 7], o
    assert o["low"].tolist() == [3, 5, 4], o
    u = op.mult(torch.as_tensor(w, device=dev)).cpu().numpy()
    idx = np.random.default_rng(3002).choice(N, 1000, replace=False)
    ref = direct_sum_torch(x, w, idx, 100.0, dev)
    err = np.linalg.norm(u[idx] - ref) / np.linalg.norm(ref)
    print(f"headline: rel. error vs direct sum {err:.3e}", flush=True)
    assert err <= 1e-4, err


def test_combined_field_torus_mixed_vs_fp64(gpu):
    """Combined field on a torus (R=1, r=0.4, analytic normals), order 8, tolerance 1e-5 (the
    estimateRefinement path), N = 262,144, kappa = 64: mixed vs FP64 form and both vs an FP64
    direct sum on 200 sampled targets."""
    from conftest import torus
    from oracle import pyoracle as po
    N, kappa = 262144, 64.0
    x, n = torus(N, 1015)
    w = cdens(N, 2015)
    u = {}
    for prec in ("mixed", "fp64"):
        op = ifgf_amd.CombinedFieldHelmholtzIfgfOperator(kappa, 64, 8, 1, 1e-5, precision=prec)
        op.init(x.T, x.T, n.T)
        print(f"torus {prec}: {op.plan_orders()} {op.plan_counts()}", flush=True)
        u[prec] = op.mult(w)
        del op
    idx = np.random.default_rng(3015).choice(N, 200, replace=False)
    ref = po.Oracle(po.COMBINED, kappa, 64, 8, 1, 1e-5).direct_sum(x, x, w, idx, normals=n)
    e_mix = np.linalg.norm(u["mixed"][idx] - ref) / np.linalg.norm(ref)
    e_64 = np.linalg.norm(u["fp64"][idx] - ref) / np.linalg.norm(ref)
    d = np.linalg.norm(u["mixed"] - u["fp64"]) / np.linalg.norm(u["fp64"])
    print(f"torus: vs direct mixed {e_mix:.2e} fp64 {e_64:.2e}; mixed vs fp64 {d:.2e}", flush=True)
    assert d < 1e-5 and e_64 < 1e-4 and e_mix < 1e-4


@pytest.mark.timeout(1200)
def test_configuration5_torus_4M(gpu):
    """N = 4,194,304 on the torus, combined field, kappa = 200, order 8, tolerance 1e-5 (the
    orders the tolerance path picks, (8,8,8)/(5,5,5), give 12 M coarse and 169 M fine cones):
    the mixed form with every level resident (218 GB of values and coefficients), then the FP64
    form, whose 436 GB of values and coefficients do not fit one GPU, with level storage (the
    values of one level and the fine coefficients of two levels, about 260 GB), chosen by init;
    both against an FP64 direct sum on 200 sampled targets, and against each other."""
    import time

    import torch
    from conftest import torus
    from oracle import pyoracle as po
    N, kappa = 4194304, 200.0
    x, n = torus(N, 1005)
    w = cdens(N, 2005)
    wd = torch.as_tensor(w, device="cuda:0")
    u = {}
    for prec in ("mixed", "fp64"):
        print(f"configuration 5 {prec}: plan", flush=True)
        op = ifgf_amd.CombinedFieldHelmholtzIfgfOperator(kappa, 64, 8, 1, 1e-5, precision=prec)
        op.init(x.T, x.T, n.T)
        c = op.plan_counts()
        print(f"configuration 5 {prec}: {op.plan_orders()} {c}", flush=True)
        assert c["level_ring"] == (1 if prec == "fp64" else 0)
        op.mult(wd)
        torch.cuda.synchronize()
        t0 = time.time()
        u[prec] = op.mult(wd)
        torch.cuda.synchronize()
        print(f"configuration 5 {prec}: mult {1e3 * (time.time() - t0):.0f} ms", flush=True)
        u[prec] = u[prec].cpu().numpy()
        del op
        torch.cuda.empty_cache()
    idx = np.random.default_rng(3005).choice(N, 200, replace=False)
    ref = po.Oracle(po.COMBINED, kappa, 64, 8, 1, 1e-5).direct_sum(x, x, w, idx, normals=n)
    err = {p: np.linalg.norm(v[idx] - ref) / np.linalg.norm(ref) for p, v in u.items()}
    d = np.linalg.norm(u["mixed"] - u["fp64"]) / np.linalg.norm(u["fp64"])
    print(f"configuration 5: rel. error vs direct sum mixed {err['mixed']:.2e} fp64 "
          f"{err['fp64']:.2e}; mixed vs fp64 {d:.2e}", flush=True)
    assert err["mixed"] < 1e-4 and err["fp64"] < 1e-4 and d < 1e-5
