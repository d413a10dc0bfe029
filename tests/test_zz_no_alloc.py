"""The allocation-free mult (ifgf_mult_noalloc) through the C ABI: its results equal those of
ifgf_mult with device pointers bit for bit, on two calls in a row of which the second has the
smaller operator (the fixed scratch keeps the first call's contents), in both precisions; an
operator larger than the scratch is refused with an error and nothing is computed."""
import ctypes

import numpy as np
import pytest

import ifgf_amd
from ifgf_amd import _lib
from conftest import cdens, sphere

pytestmark = pytest.mark.gpu


def noalloc(op, w):
    """One ifgf_mult_noalloc call on torch's current stream (the caller owns w and out)."""
    import torch
    out = torch.empty(op._m, dtype=torch.complex128, device=w.device)
    stream = torch.cuda.current_stream(w.device).cuda_stream
    rc = _lib.lib().ifgf_mult_noalloc(op._h, w.data_ptr(), out.data_ptr(), stream)
    return rc, out


@pytest.mark.parametrize("prec", ["mixed", "fp64"])
def test_noalloc_matches_mult_two_calls(gpu, prec):
    import torch
    print(f"noalloc {prec}: operators of 24000 and 7000 points", flush=True)
    ops, ws, refs = [], [], []
    for n, seed, kappa in ((24000, 150, 30.0), (7000, 151, 12.0)):
        x = sphere(n, seed)
        op = ifgf_amd.HelmholtzIfgfOperator(kappa, 64, 6, 1, -1, precision=prec)
        op.init(x.T, x.T)
        w = torch.as_tensor(cdens(n, seed + 10), device="cuda:0")
        refs.append(op.mult(w))  # ifgf_mult, IFGF_DEVICE_POINTERS
        ops.append(op)
        ws.append(w)
    torch.cuda.synchronize()
    for rnd in range(2):
        for i, (op, w) in enumerate(zip(ops, ws)):
            rc, out = noalloc(op, w)
            assert rc == _lib.IFGF_OK, _lib.lib().ifgf_last_error(op._h)
            torch.cuda.synchronize()
            assert torch.equal(out, refs[i]), (prec, rnd, i)


def test_noalloc_gradient_and_targets(gpu):
    """Targets other than the sources and the gradient kernel: the same bit for bit."""
    import torch
    x = sphere(9000, 152)
    y = sphere(5000, 153) * 1.3
    op = ifgf_amd.GradHelmholtzIfgfOperator(2.0 + 14.0j, 64, 6, 1, -1)
    op.init(x.T, y.T)
    op.setDx(1)
    w = torch.as_tensor(cdens(9000, 154), device="cuda:0")
    ref = op.mult(w)
    rc, out = noalloc(op, w)
    assert rc == _lib.IFGF_OK, _lib.lib().ifgf_last_error(op._h)
    torch.cuda.synchronize()
    assert torch.equal(out, ref)


def test_noalloc_refuses_operators_beyond_the_scratch(gpu):
    """FP64: 16 bytes per point, so 1,000 sources and 4,194,304 targets exceed 64 MiB."""
    import torch
    rng = np.random.default_rng(155)
    x = sphere(1000, 156)
    y = rng.uniform(-2.0, 2.0, (4194304, 3))
    op = ifgf_amd.HelmholtzIfgfOperator(1.0, 64, 6, 1, -1, precision="fp64")
    op.init(x.T, y.T)
    w = torch.zeros(1000, dtype=torch.complex128, device="cuda:0")
    rc, _ = noalloc(op, w)
    assert rc == _lib.IFGF_EINVAL
    assert b"scratch" in _lib.lib().ifgf_last_error(op._h)
