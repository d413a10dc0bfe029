"""Plan files on the device: an operator loaded from a file gives the results of the
operator that wrote it, bit for bit (mixed and FP64 forms, the saved launch shape, level
storage, the gradient direction of the loading operator, adjoints, a cutOff plan)."""
import time

import numpy as np
import pytest

import ifgf_amd
from conftest import cdens, sphere, spheroid
from ifgf_amd._lib import IfgfError

pytestmark = pytest.mark.gpu


@pytest.mark.parametrize("prec", ["mixed", "fp64"])
def test_loaded_plan_gives_identical_results(gpu, tmp_path, prec):
    N = 60000
    x = sphere(N, 51)
    w = cdens(N, 52)
    print(f"plan file helmholtz N={N} tol=1e-5 prec={prec}", flush=True)
    op = ifgf_amd.HelmholtzIfgfOperator(40.0, 64, 6, 1, 1e-5, precision=prec)
    t0 = time.time()
    op.init(x.T, x.T)
    t_init = time.time() - t0
    op.launch_shape = dict(xcd_run_log2=2, up_cones=4, overlap=2)
    u = op.mult(w)
    op.save_plan(tmp_path / "p.plan")
    op2 = ifgf_amd.HelmholtzIfgfOperator(40.0, 64, 6, 1, 1e-5, precision=prec)
    t0 = time.time()
    op2.load_plan(tmp_path / "p.plan")
    t_load = time.time() - t0
    print(f"init {t_init * 1e3:.0f} ms, load {t_load * 1e3:.0f} ms", flush=True)
    assert op2.launch_shape == op.launch_shape
    assert np.array_equal(op2.mult(w), u)
    # level storage chosen at load time: the resident plan's results
    op3 = ifgf_amd.HelmholtzIfgfOperator(40.0, 64, 6, 1, 1e-5, precision=prec)
    op3.load_plan(tmp_path / "p.plan", level_ring=True)
    assert op3.plan_counts()["level_ring"] == 1
    assert np.array_equal(op3.mult(w), u)
    # the adjoint (targets = sources) and several columns
    g = cdens(N, 53)
    assert np.array_equal(op2.mult_adjoint(g), op.mult_adjoint(g))
    W = np.stack([w, g], axis=1)
    assert np.array_equal(op2.mult(W), op.mult(W))


def test_gradient_direction_and_adjoint_of_other_targets(gpu, tmp_path):
    x = spheroid(20000, 54)
    y = sphere(8000, 55) * 1.5
    w = cdens(20000, 56)
    g = cdens(8000, 57)
    print("plan file gradient kernel, other targets", flush=True)
    a = ifgf_amd.GradHelmholtzIfgfOperator(1.0 + 12.0j, 64, 8, 1, -1)
    a.setDx(2)
    a.init(x.T, y.T)
    a.save_plan(tmp_path / "g.plan")
    b = ifgf_amd.GradHelmholtzIfgfOperator(1.0 + 12.0j, 64, 8, 1, -1)
    b.setDx(0)
    b.load_plan(tmp_path / "g.plan")
    a.setDx(0)
    assert np.array_equal(b.mult(w), a.mult(w))
    assert np.array_equal(b.mult_adjoint(g), a.mult_adjoint(g))


def test_cutoff_plan(gpu, tmp_path):
    x = sphere(20000, 58)
    y = sphere(6000, 59) * 1.2
    w = cdens(20000, 60)
    print("plan file with a cutOff callback", flush=True)
    a = ifgf_amd.HelmholtzIfgfOperator(20.0, 64, 6, 1, -1)
    a.init(x.T, y.T, cutOff=lambda d: d > 0.5)
    a.save_plan(tmp_path / "c.plan")
    b = ifgf_amd.HelmholtzIfgfOperator(20.0, 64, 6, 1, -1)
    b.load_plan(tmp_path / "c.plan")
    assert np.array_equal(b.mult(w), a.mult(w))
    with pytest.raises(IfgfError, match="cutOff"):
        b.mult_adjoint(cdens(6000, 61))


def test_merged_plan_and_allocation_free_mult(gpu, tmp_path):
    """A merged operator's plan file loads into an operator of its first part's parameters and
    gives the merged results (host and device densities); a loaded plan serves the
    allocation-free mult with the same bits as its mult."""
    import ctypes

    import torch

    from ifgf_amd import _lib
    a = ifgf_amd.HelmholtzIfgfOperator(10.0, 64, 6, 1, -1)
    b = ifgf_amd.HelmholtzIfgfOperator(16.0, 64, 6, 1, -1)
    xa, xb = sphere(12000, 62), spheroid(9000, 63)
    a.init(xa.T, xa.T)
    b.init(xb.T, xb.T)
    m = ifgf_amd.merge([a, b])
    w = cdens(21000, 64)
    u = m.mult(w)
    m.save_plan(tmp_path / "m.plan")
    print("merged plan file: 12,000 + 9,000 points", flush=True)
    op = ifgf_amd.HelmholtzIfgfOperator(10.0, 64, 6, 1, -1)
    op.load_plan(tmp_path / "m.plan")
    assert np.array_equal(op.mult(w), u)
    wd = torch.as_tensor(w, device="cuda:0")
    ud = op.mult(wd)
    torch.cuda.synchronize()
    assert np.array_equal(ud.cpu().numpy(), u)
    out = torch.empty_like(wd)
    stream = torch.cuda.current_stream(wd.device).cuda_stream
    rc = _lib.lib().ifgf_mult_noalloc(op._h, ctypes.c_void_p(wd.data_ptr()),
                                      ctypes.c_void_p(out.data_ptr()), ctypes.c_void_p(stream))
    assert rc == _lib.IFGF_OK
    torch.cuda.synchronize()
    assert np.array_equal(out.cpu().numpy(), ud.cpu().numpy())
