"""Launch state that travels with each operator: operators with different launch shapes
enqueued back to back on one non-blocking stream (their far groups on their own side streams)
give the results each gives alone, bit for bit; a captured graph of the gradient operator is
dropped when setDx changes the direction."""
import numpy as np
import pytest

import ifgf_amd
from conftest import cdens, sphere

pytestmark = pytest.mark.gpu


def test_operators_with_different_xcd_runs_back_to_back(gpu):
    import torch
    x = sphere(60000, 140)
    y = sphere(50000, 141)
    ops = []
    for pts, lg in ((x, 0), (y, 6)):
        op = ifgf_amd.HelmholtzIfgfOperator(40.0, 64, 6, 1, -1)
        op.init(pts.T, pts.T)
        op.launch_shape = {"xcd_run_log2": lg, "overlap": 1}
        ops.append(op)
    ws = [torch.as_tensor(cdens(60000, 142), device="cuda:0"),
          torch.as_tensor(cdens(50000, 143), device="cuda:0")]
    alone = []
    for op, w in zip(ops, ws):
        u = op.mult(w)
        torch.cuda.synchronize()
        alone.append(u.cpu())
    s = torch.cuda.Stream()  # non-blocking, default priority
    print("back to back: 2 operators x 4 rounds on one stream", flush=True)
    with torch.cuda.stream(s):
        outs = []
        for _ in range(4):
            for op, w in zip(ops, ws):
                outs.append(op.mult(w))
    s.synchronize()
    for i, u in enumerate(outs):
        assert torch.equal(u.cpu(), alone[i % 2]), i


def test_set_dx_recaptures_graph(gpu):
    import torch
    x = sphere(8000, 144)
    w = torch.as_tensor(cdens(8000, 145), device="cuda:0")
    eager = {}
    ref = ifgf_amd.GradHelmholtzIfgfOperator(1.0 + 12.0j, 64, 6, 1, -1)
    ref.init(x.T, x.T)
    for dx in (0, 1):
        ref.setDx(dx)
        eager[dx] = ref.mult(w).cpu()
    assert not torch.equal(eager[0], eager[1])
    op = ifgf_amd.GradHelmholtzIfgfOperator(1.0 + 12.0j, 64, 6, 1, -1)
    op.init(x.T, x.T)
    op.use_graph = True
    op.setDx(0)
    assert torch.equal(op.mult(w).cpu(), eager[0])
    op.setDx(1)
    assert torch.equal(op.mult(w).cpu(), eager[1])
    op.setDx(0)
    assert torch.equal(op.mult(w).cpu(), eager[0])
