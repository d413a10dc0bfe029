"""Launch shapes and the autotuner: every shape gives the same results as the default one
(shapes change the mapping of work to workgroups and XCDs, never the arithmetic), the tuner
returns a valid shape and leaves it applied, and captured graphs follow a shape change."""
import itertools

import numpy as np
import pytest

import ifgf_amd
from conftest import cdens, sphere, spheroid

pytestmark = pytest.mark.gpu

CPB = {("mixed", 6): [4, 2, 6, 8], ("fp64", 6): [3, 1, 2, 4], ("mixed", 8): [2, 1, 3],
       ("fp64", 8): [1]}


@pytest.mark.parametrize("prec,order", [("mixed", 6), ("fp64", 6), ("mixed", 8), ("fp64", 8)])
def test_every_shape_gives_the_same_result(gpu, prec, order):
    x = sphere(12000, 601)
    w = cdens(12000, 602)
    op = ifgf_amd.HelmholtzIfgfOperator(12.0, 64, order, 1, -1, precision=prec)
    op.init(x.T, x.T)
    ref = op.mult(w)
    waves = [1, 2, 4, 8] if prec == "mixed" else [4]
    for lg, cpb, k, fw, ov in itertools.product([0, 2, 6], CPB[(prec, order)], [2, 4, 16], waves,
                                                [0, 1, 2]):
        op.launch_shape = dict(xcd_run_log2=lg, transform_cones=cpb, up_cones=k, far_waves=fw,
                               overlap=ov)
        print(f"shape prec={prec} order={order} xcd_log2={lg} cpb={cpb} up_k={k} far_waves={fw} "
              f"overlap={ov}", flush=True)
        assert np.array_equal(op.mult(w), ref), (lg, cpb, k, fw, ov)


def test_tolerance_orders_every_shape_gives_the_same_result(gpu):
    """The orders (6,8,7)/(3,5,4) (the tolerance estimate at tol 1e-5 on this geometry): the
    FP32 upward pass stages the child coefficients in LDS per parent cone, so the cones per
    item change the staging order but not the sums."""
    x = sphere(12000, 100 + 12000 % 97)
    w = cdens(12000, 607)
    op = ifgf_amd.HelmholtzIfgfOperator(16.0, 64, 6, 1, 1e-5)
    op.init(x.T, x.T)
    o = op.plan_orders()
    assert o["high"].tolist() == [6, 8, 7] and o["low"].tolist() == [3, 5, 4], o
    ref = op.mult(w)
    for lg, cpb, k, fw, ov in itertools.product([0, 4], [2, 1, 3], [2, 4, 8, 16], [1, 4], [0, 1]):
        op.launch_shape = dict(xcd_run_log2=lg, transform_cones=cpb, up_cones=k, far_waves=fw,
                               overlap=ov)
        print(f"shape (6,8,7) xcd_log2={lg} cpb={cpb} up_k={k} far_waves={fw} overlap={ov}",
              flush=True)
        assert np.array_equal(op.mult(w), ref), (lg, cpb, k, fw, ov)


def test_batched_shapes_and_autotune(gpu):
    import torch
    x = spheroid(20000, 603)
    W = np.stack([cdens(20000, 604 + i) for i in range(20)], axis=1)
    op = ifgf_amd.HelmholtzIfgfOperator(16.0, 64, 6, 1, -1)
    op.init(x.T, x.T)
    ref = op.mult(W)
    for rb, k in itertools.product([8, 32], [2, 4, 16]):
        op.launch_shape = dict(batch_columns=rb, up_cones=k)
        print(f"batched shape rb={rb} up_k={k}", flush=True)
        U = op.mult(W)
        assert np.array_equal(U, ref), (rb, k)
    best = op.autotune(nrhs=20, reps=2)
    print(f"batched autotune: {best}", flush=True)
    assert best == op.launch_shape
    assert best["batch_columns"] in (8, 32) and best["up_cones"] in (2, 4, 8, 16)
    Wd = torch.as_tensor(W, device="cuda:0")
    assert np.array_equal(op.mult(Wd).cpu().numpy(), ref)


def test_autotune_single_column_and_graph_recapture(gpu):
    import torch
    x = sphere(30000, 605)
    w = cdens(30000, 606)
    op = ifgf_amd.HelmholtzIfgfOperator(20.0, 64, 6, 1, -1)
    op.init(x.T, x.T)
    ref = op.mult(w)
    wd = torch.as_tensor(w, device="cuda:0")
    op.use_graph = True
    g1 = op.mult(wd).cpu().numpy()
    best = op.autotune(nrhs=1, reps=2)
    print(f"single-column autotune: {best}", flush=True)
    assert best == op.launch_shape
    assert 0 <= best["xcd_run_log2"] <= 6 and best["transform_cones"] in CPB[("mixed", 6)]
    # a different shape after the capture: the graph is captured again with the new kernels
    op.launch_shape = dict(up_cones=4 if best["up_cones"] != 4 else 16, transform_cones=2)
    g2 = op.mult(wd).cpu().numpy()
    assert np.array_equal(g1, ref) and np.array_equal(g2, ref)
    with pytest.raises(ifgf_amd.IfgfError):
        op.launch_shape = dict(up_cones=5)
    with pytest.raises(ifgf_amd.IfgfError):
        op.launch_shape = dict(transform_cones=7)
    with pytest.raises(ifgf_amd.IfgfError):
        op.launch_shape = dict(far_waves=3)
    # the overlapped near field inside a captured graph: replays against the eager mult
    op.launch_shape = dict(overlap=1)
    for _ in range(3):
        g3 = op.mult(wd).cpu().numpy()
        assert np.array_equal(g3, ref)
