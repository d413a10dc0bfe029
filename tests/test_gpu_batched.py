"""Batched mult (several density columns per pass, one geometry): against the single-column
mult column by column, against the oracle, and the batched configuration at full size."""
import os

import numpy as np
import pytest

import ifgf_amd
from conftest import cdens, sphere, spheroid
from oracle import pyoracle as po

pytestmark = pytest.mark.gpu

# batched vs single-column mult: the same arithmetic per column, other summation orders
TOL_SAME = {"fp64": 1e-13, "mixed": 2e-6}
TOL_ORACLE = {"fp64": 1e-11, "mixed": 2e-5}


def rel(a, b):
    return float(np.linalg.norm(a - b) / np.linalg.norm(b))


def single_columns(op, W):
    os.environ["IFGF_NO_BATCH"] = "1"
    try:
        return op.mult(W)
    finally:
        del os.environ["IFGF_NO_BATCH"]


@pytest.mark.parametrize("prec", ["fp64", "mixed"])
@pytest.mark.parametrize("nrhs", [3, 20, 40])
def test_batched_matches_single_columns(gpu, prec, nrhs):
    x = sphere(6000, 91)
    W = np.stack([cdens(6000, 92 + i) for i in range(nrhs)], axis=1)
    print(f"batched helmholtz N=6000 nrhs={nrhs} prec={prec}", flush=True)
    op = ifgf_amd.HelmholtzIfgfOperator(8.0, 64, 6, 1, -1, precision=prec)
    op.init(x.T, x.T)
    U = op.mult(W)
    assert U.shape == (6000, nrhs)
    ref = single_columns(op, W)
    for i in range(nrhs):
        assert rel(U[:, i], ref[:, i]) < TOL_SAME[prec], (i, rel(U[:, i], ref[:, i]))
    orc = po.Oracle(po.HELMHOLTZ, 8.0, 64, 6, 1, -1).init(x, x)
    for i in (0, nrhs - 1):
        assert rel(U[:, i], orc.mult(W[:, i])) < TOL_ORACLE[prec]


@pytest.mark.parametrize("prec", ["fp64", "mixed"])
def test_batched_kernel_families_and_order8(gpu, prec):
    """Every kernel family in the batched kernels (order 8), targets other than the sources,
    the adjoint through the batched path, and device tensors."""
    import torch
    x = spheroid(5000, 93)
    y = sphere(3000, 94) * 1.4
    n = x * np.array([1.0, 1.0, 4.0])
    n /= np.linalg.norm(n, axis=1)[:, None]
    W = np.stack([cdens(5000, 95 + i) for i in range(5)], axis=1)
    cases = [
        (ifgf_amd.MofifiedHelmholtzIfgfOperator(0.5 - 9.0j, 64, 8, 1, -1, precision=prec), None),
        (ifgf_amd.HelmholtzIfgfOperator(9.0, 64, 8, 1, -1, precision=prec), None),
        (ifgf_amd.LaplaceIfgfOperator(64, 8, 1, -1, precision=prec), None),
        (ifgf_amd.DoubleLayerHelmholtzIfgfOperator(0.5 + 9.0j, 40, 8, 1, -1, precision=prec), n),
        (ifgf_amd.CombinedFieldHelmholtzIfgfOperator(9.0, 40, 8, 1, -1, precision=prec), n),
    ]
    for op, nrm in cases:
        print(f"batched {type(op).__name__} prec={prec}", flush=True)
        if nrm is None:
            op.init(x.T, y.T)
        else:
            op.init(x.T, y.T, nrm.T)
        Win = W.real if isinstance(op, ifgf_amd.LaplaceIfgfOperator) else W
        U = op.mult(Win)
        ref = single_columns(op, Win)
        assert rel(U, ref) < TOL_SAME[prec]
    g = ifgf_amd.GradHelmholtzIfgfOperator(1.0 + 8.0j, 64, 8, 1, -1, precision=prec)
    g.setDx(1)
    g.init(x.T, y.T)
    G = np.stack([cdens(3000, 96 + i) for i in range(4)], axis=1)
    A = g.mult_adjoint(G)
    os.environ["IFGF_NO_BATCH"] = "1"
    try:
        Aref = g.mult_adjoint(G)
    finally:
        del os.environ["IFGF_NO_BATCH"]
    assert A.shape == (5000, 4) and rel(A, Aref) < TOL_SAME[prec]
    Ut = g.mult(torch.as_tensor(W, device="cuda:0"))
    torch.cuda.synchronize()
    assert rel(Ut.cpu().numpy(), single_columns(g, W)) < TOL_SAME[prec]


@pytest.mark.parametrize("nrhs", [5, 40])
def test_mfma_transform_matches_separable_transform(gpu, nrhs):
    """The batched transform on MFMA tiles (orders (3,6,6)/(2,3,3), mixed form, 8- and
    32-column passes) against the separable VALU transform, the single-column mult and the
    oracle; the gradient kernel shares the transform."""
    x = spheroid(8000, 97)
    W = np.stack([cdens(8000, 98 + i) for i in range(nrhs)], axis=1)
    ops = []
    for mfma in (False, True):
        if mfma:
            os.environ["IFGF_TRANSFORM_MFMA"] = "1"
        try:
            print(f"batched transform mfma={mfma} nrhs={nrhs}", flush=True)
            op = ifgf_amd.HelmholtzIfgfOperator(12.0, 64, 6, 1, -1)
            op.init(x.T, x.T)
            ops.append((op, op.mult(W)))
        finally:
            os.environ.pop("IFGF_TRANSFORM_MFMA", None)
    (op, U), (_, Um) = ops
    o = op.plan_orders()
    assert list(o["high"]) == [3, 6, 6] and list(o["low"]) == [2, 3, 3]
    assert rel(Um, U) < TOL_SAME["mixed"]
    ref = single_columns(op, W)
    assert rel(Um, ref) < TOL_SAME["mixed"]
    orc = po.Oracle(po.HELMHOLTZ, 12.0, 64, 6, 1, -1).init(x, x)
    assert rel(Um[:, 0], orc.mult(W[:, 0])) < TOL_ORACLE["mixed"]
    res = []
    for mfma in (False, True):
        if mfma:
            os.environ["IFGF_TRANSFORM_MFMA"] = "1"
        try:
            g = ifgf_amd.GradHelmholtzIfgfOperator(12.0, 64, 6, 1, -1)
            g.setDx(1)
            g.init(x.T, x.T)
            res.append(g.mult(W))
        finally:
            os.environ.pop("IFGF_TRANSFORM_MFMA", None)
    assert rel(res[1], res[0]) < TOL_SAME["mixed"]


def test_transform_forms_configuration3_spheroid(gpu):
    """Configuration 3's spheroid (N=262,144, kappa=64, 32 columns, one stream) with the
    separable VALU transform and with the MFMA transform: the same results to rounding; the
    times per mult are printed (the kernel trace of this test compares the two kernels)."""
    import torch
    N = 262144
    x = spheroid(N, 1003)
    rng = np.random.default_rng(2003)
    Wd = torch.as_tensor(rng.standard_normal((N, 32)) + 1j * rng.standard_normal((N, 32)),
                         device="cuda:0")
    out = {}
    for mfma in (False, True):
        if mfma:
            os.environ["IFGF_TRANSFORM_MFMA"] = "1"
        try:
            op = ifgf_amd.HelmholtzIfgfOperator(64.0, 64, 6, 1, -1)
            op.init(x.T, x.T)
            op.launch_shape = dict(overlap=0)
            op.mult(Wd)
            s, e = torch.cuda.Event(enable_timing=True), torch.cuda.Event(enable_timing=True)
            s.record()
            for _ in range(3):
                U = op.mult(Wd)
            e.record()
            torch.cuda.synchronize()
            out[mfma] = U.cpu().numpy()
            print(f"transform on {'MFMA' if mfma else 'VALU'}: {s.elapsed_time(e) / 3:.2f} ms "
                  f"per 32-column mult", flush=True)
        finally:
            os.environ.pop("IFGF_TRANSFORM_MFMA", None)
        del op
    assert rel(out[True], out[False]) < TOL_SAME["mixed"]


def test_configuration3_spheroid_32_columns(gpu):
    """N=262,144 oblate spheroid (1,1,0.5), kappa=64, 32 complex N(0,1) density columns in one
    batched pass (mixed form): columns against the single-column mult and against an FP64
    direct sum on 500 sampled targets."""
    import time

    import torch
    N, kappa = 262144, 64.0
    x = spheroid(N, 1003)
    rng = np.random.default_rng(2003)
    W = rng.standard_normal((N, 32)) + 1j * rng.standard_normal((N, 32))
    op = ifgf_amd.HelmholtzIfgfOperator(kappa, 64, 6, 1, -1)
    op.init(x.T, x.T)
    Wd = torch.as_tensor(W, device="cuda:0")
    op.mult(Wd)
    torch.cuda.synchronize()
    t0 = time.time()
    U = op.mult(Wd)
    torch.cuda.synchronize()
    tb = time.time() - t0
    os.environ["IFGF_NO_BATCH"] = "1"
    try:
        t0 = time.time()
        Us = op.mult(Wd)
        torch.cuda.synchronize()
        ts = time.time() - t0
    finally:
        del os.environ["IFGF_NO_BATCH"]
    U = U.cpu().numpy()
    Us = Us.cpu().numpy()
    print(f"configuration 3: 32 columns batched {tb * 1e3:.1f} ms, one by one "
          f"{ts * 1e3:.1f} ms", flush=True)
    assert rel(U, Us) < TOL_SAME["mixed"]
    idx = np.random.default_rng(3003).choice(N, 500, replace=False)
    X = torch.as_tensor(x, device="cuda:0")
    Y = X[torch.as_tensor(idx, device="cuda:0")]
    diff = Y[:, None, :] - X[None, :, :]
    d = torch.sqrt((diff * diff).sum(-1))
    dd = torch.where(d > 0, d, torch.ones_like(d))
    G = torch.where(d > 0, torch.exp(1j * kappa * dd) / (4 * np.pi * dd),
                    torch.zeros_like(dd, dtype=torch.complex128))
    for c in (0, 31):
        ref = (G @ Wd[:, c]).cpu().numpy()
        err = rel(U[idx, c], ref)
        print(f"configuration 3: column {c} rel. error vs direct sum {err:.2e}", flush=True)
        assert err < 3e-3


@pytest.mark.parametrize("prec", ["fp64", "mixed"])
def test_merged_geometries_match_separate_operators(gpu, prec):
    """Three geometries with their own wavenumbers merged into one operator (one launch per
    level): each part's result against its own operator, the adjoint, and a direct sum."""
    geos = [(sphere(6000, 111), 8.0), (spheroid(4000, 112) * 1.3, 15.0), (sphere(1500, 113), 3.0)]
    ops, ws = [], []
    for i, (x, k) in enumerate(geos):
        op = ifgf_amd.HelmholtzIfgfOperator(k, 64, 6, 1, -1, precision=prec)
        op.init(x.T, x.T)
        ops.append(op)
        ws.append(cdens(len(x), 114 + i))
    m = ifgf_amd.merge(ops)
    us = m.mult(ws)
    for op, w, u in zip(ops, ws, us):
        assert rel(u, op.mult(w)) < TOL_SAME[prec]
    # the adjoint of the merged operator (targets = sources in every part)
    gs = [cdens(len(x), 120 + i) for i, (x, _) in enumerate(geos)]
    a = m.mult_adjoint(np.concatenate(gs))
    at = 0
    for op, g in zip(ops, gs):
        assert rel(a[at:at + len(g)], op.mult_adjoint(g)) < TOL_SAME[prec]
        at += len(g)
    x, k = geos[2]
    idx = np.arange(0, len(x), 7)
    d = po.Oracle(po.HELMHOLTZ, k, 64, 6, 1, -1).direct_sum(x, x, ws[2], idx)
    assert rel(us[2][idx], d) < 2e-3


def test_merged_tolerance_order_class(gpu):
    """Two geometries (one the other scaled by 1.5, kappa by 1/1.5) whose tolerance estimates
    give the orders (6,8,7)/(3,5,4), merged into one operator: the LDS-staged upward pass reads
    each part's lists of child cones (mixed form), each part against its own operator."""
    x = sphere(12000, 100 + 12000 % 97)
    geos = [(x, 16.0), (x * 1.5, 16.0 / 1.5)]
    ops, ws = [], []
    for i, (p, k) in enumerate(geos):
        op = ifgf_amd.HelmholtzIfgfOperator(k, 64, 6, 1, 1e-5)
        op.init(p.T, p.T)
        o = op.plan_orders()
        assert o["high"].tolist() == [6, 8, 7] and o["low"].tolist() == [3, 5, 4], o
        ops.append(op)
        ws.append(cdens(len(p), 130 + i))
    m = ifgf_amd.merge(ops)
    us = m.mult(ws)
    for op, w, u in zip(ops, ws, us):
        assert rel(u, op.mult(w)) < TOL_SAME["mixed"]


def test_configuration3_sixteen_spheres_one_launch_per_level(gpu):
    """16 spheres, N log-uniform in [4096, 262144], kappa uniform in [8, 64], merged into one
    operator (mixed form): every part against its own operator; timings of one merged mult
    against the 16 separate mults."""
    import time

    import torch
    rng = np.random.default_rng(1013)
    Ns = np.floor(4096 * 64.0 ** rng.random(16)).astype(int)
    ks = rng.uniform(8, 64, 16)
    ops, ws = [], []
    for i, (n, k) in enumerate(zip(Ns, ks)):
        x = sphere(int(n), 1100 + i)
        op = ifgf_amd.HelmholtzIfgfOperator(float(k), 64, 6, 1, -1)
        op.init(x.T, x.T)
        ops.append(op)
        ws.append(cdens(int(n), 2100 + i))
    m = ifgf_amd.merge(ops)
    print(f"configuration 3: 16 spheres, N = {Ns.tolist()}, total {int(Ns.sum())}", flush=True)
    wd = torch.as_tensor(np.concatenate(ws), device="cuda:0")
    wds = [torch.as_tensor(w, device="cuda:0") for w in ws]
    m.mult(wd)
    for op, w in zip(ops, wds):
        op.mult(w)
    torch.cuda.synchronize()
    t0 = time.time()
    for _ in range(3):
        u = m.mult(wd)
    torch.cuda.synchronize()
    tm = (time.time() - t0) / 3
    t0 = time.time()
    for _ in range(3):
        us = [op.mult(w) for op, w in zip(ops, wds)]
    torch.cuda.synchronize()
    ts = (time.time() - t0) / 3
    print(f"configuration 3: merged mult {tm * 1e3:.1f} ms, 16 separate mults {ts * 1e3:.1f} ms",
          flush=True)
    u = u.cpu().numpy()
    at = 0
    for n, ui in zip(Ns, us):
        assert rel(u[at:at + n], ui.cpu().numpy()) < TOL_SAME["mixed"]
        at += n


@pytest.mark.parametrize("prec", ["fp64", "mixed"])
def test_batched_and_merged_graph_replays_match_eager(gpu, prec):
    """Batched mults (20 columns: one pass of 32) and mults of a merged operator (one column and
    8 columns) replayed as captured HIP graphs: every replay, with eager work of another operator
    and of torch and synchronisations in between, equals the eager mult bit for bit."""
    import torch
    x = sphere(7000, 160)
    op = ifgf_amd.HelmholtzIfgfOperator(9.0, 64, 6, 1, -1, precision=prec)
    op.init(x.T, x.T)
    other = ifgf_amd.HelmholtzIfgfOperator(5.0, 32, 6, 1, -1, precision=prec)
    other.init(x.T, x.T)
    geos = [(sphere(5000, 161), 8.0), (spheroid(3000, 162) * 1.3, 14.0)]
    parts = []
    for i, (p, k) in enumerate(geos):
        o = ifgf_amd.HelmholtzIfgfOperator(k, 64, 6, 1, -1, precision=prec)
        o.init(p.T, p.T)
        parts.append(o)
    m = ifgf_amd.merge(parts)
    dev = torch.device("cuda:0")
    cases = [
        (op, torch.as_tensor(np.stack([cdens(7000, 163 + i) for i in range(20)], 1), device=dev)),
        (m, torch.as_tensor(cdens(8000, 190), device=dev)),
        (m, torch.as_tensor(np.stack([cdens(8000, 191 + i) for i in range(8)], 1), device=dev)),
    ]
    for case, (o, w) in enumerate(cases):
        print(f"graph replay case {case} prec={prec} shape={tuple(w.shape)}", flush=True)
        o.use_graph = False
        eager = o.mult(w).cpu()
        o.use_graph = True
        for rep in range(3):
            u = o.mult(w)
            other.mult(w[:, 0] if w.dim() == 2 and w.shape[0] == 7000 else
                       torch.as_tensor(cdens(7000, 199), device=dev))
            (w * 2).sum()
            torch.cuda.synchronize()
            assert torch.equal(u.cpu(), eager), (case, rep)
        o.use_graph = False


@pytest.mark.parametrize("prec", ["fp64", "mixed"])
def test_batched_tolerance_order_class(gpu, prec):
    """The batched kernels at the orders (6,8,7)/(3,5,4) that the tolerance estimate picks for
    order 6 at tol 1e-5 (N = 12,000, kappa = 16): 10 columns in one pass against the
    single-column mult column by column and two columns against the oracle."""
    x = sphere(12000, 100 + 12000 % 97)
    W = np.stack([cdens(12000, 170 + i) for i in range(10)], axis=1)
    print(f"batched tolerance class prec={prec}", flush=True)
    op = ifgf_amd.HelmholtzIfgfOperator(16.0, 64, 6, 1, 1e-5, precision=prec)
    op.init(x.T, x.T)
    o = op.plan_orders()
    assert o["high"].tolist() == [6, 8, 7] and o["low"].tolist() == [3, 5, 4], o
    U = op.mult(W)
    ref = single_columns(op, W)
    for i in range(10):
        assert rel(U[:, i], ref[:, i]) < TOL_SAME[prec], (i, rel(U[:, i], ref[:, i]))
    orc = po.Oracle(po.HELMHOLTZ, 16.0, 64, 6, 1, 1e-5).init(x, x)
    for i in (0, 9):
        assert rel(U[:, i], orc.mult(W[:, i])) < TOL_ORACLE[prec]


def test_configuration3_one_operator(gpu):
    """Configuration 3 as one operator: the oblate spheroid (N = 262,144, kappa = 64) and 16
    spheres (N log-uniform in [4096, 262144], kappa uniform in [8, 64]) merged, 32 density
    columns per part, one batched pass (one launch per level for every geometry and column),
    FP64 form: every part against its own operator (same columns) to 1e-12, and the three
    smallest spheres against the oracle on two columns to 1e-11; the mixed form against the
    FP64 one; timings of the merged pass against the 17 separate batched mults."""
    import time

    import torch
    dev = torch.device("cuda:0")
    rng = np.random.default_rng(1013)
    Ns = np.floor(4096 * 64.0 ** rng.random(16)).astype(int)
    ks = rng.uniform(8, 64, 16)
    geos = [(spheroid(262144, 1003), 64.0)] + [(sphere(int(n), 1100 + i), float(k))
                                               for i, (n, k) in enumerate(zip(Ns, ks))]
    print(f"configuration 3 one operator: N = {[len(g[0]) for g in geos]}", flush=True)
    res = {}
    for prec in ("fp64", "mixed"):
        ops, Ws = [], []
        for i, (x, k) in enumerate(geos):
            op = ifgf_amd.HelmholtzIfgfOperator(k, 64, 6, 1, -1, precision=prec)
            op.init(x.T, x.T)
            ops.append(op)
            r = np.random.default_rng(2200 + i)
            Ws.append(r.standard_normal((len(x), 32)) + 1j * r.standard_normal((len(x), 32)))
        m = ifgf_amd.merge(ops)
        Wd = torch.as_tensor(np.concatenate(Ws), device=dev)
        m.mult(Wd)
        torch.cuda.synchronize()
        t0 = time.time()
        U = m.mult(Wd)
        torch.cuda.synchronize()
        tm = time.time() - t0
        wds = [torch.as_tensor(W, device=dev) for W in Ws]
        for op, W in zip(ops, wds):
            op.mult(W)
        torch.cuda.synchronize()
        t0 = time.time()
        us = [op.mult(W) for op, W in zip(ops, wds)]
        torch.cuda.synchronize()
        ts = time.time() - t0
        print(f"configuration 3 one operator {prec}: merged 32 columns {tm * 1e3:.1f} ms, 17 "
              f"separate batched mults {ts * 1e3:.1f} ms", flush=True)
        U = U.cpu().numpy()
        res[prec] = U
        at = 0
        for (x, k), ui in zip(geos, us):
            part = U[at:at + len(x)]
            at += len(x)
            tol = 1e-12 if prec == "fp64" else TOL_SAME["mixed"]
            assert rel(part, ui.cpu().numpy()) < tol
        if prec == "fp64":
            order = np.argsort([len(g[0]) for g in geos[1:]])[:3] + 1
            offs = np.cumsum([0] + [len(g[0]) for g in geos])
            for j in order:
                x, k = geos[j]
                orc = po.Oracle(po.HELMHOLTZ, k, 64, 6, 1, -1).init(x, x)
                for c in (0, 31):
                    assert rel(U[offs[j]:offs[j + 1], c], orc.mult(Ws[j][:, c])) < 1e-11
        del m, ops
        torch.cuda.empty_cache()
    d = rel(res["mixed"], res["fp64"])
    print(f"configuration 3 one operator: mixed vs fp64 {d:.2e}", flush=True)
    assert d < 2e-5
