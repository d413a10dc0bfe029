"""Octrees built on the GPU (Morton keys, radix sort, exact tie order, greedy child scans)
against the host builder: the same permutation, boxes, point ranges and bounding boxes bit for
bit, and therefore the same interaction lists and cones, on ordinary geometries and on inputs
that exercise every tie rule of the Morton order (zorder_less.hpp:212-248) and the closed-box
child scan (octree.hpp:1106-1199)."""
import time

import numpy as np
import pytest

import ifgf_amd
from conftest import cdens, sphere, spheroid, torus

pytestmark = pytest.mark.gpu


def _bits(a):
    return np.ascontiguousarray(a).view(np.int64)


def _same_tree(a, b):
    for k in ("level_off", "begin", "end", "parent", "leaf", "perm"):
        assert np.array_equal(a[k], b[k]), k
    assert np.array_equal(_bits(a["bbox"]), _bits(b["bbox"])), "bbox"


def _plans_equal(dev, host):
    cd, ch = dev.plan_counts(), host.plan_counts()
    device_only = ("device_bytes", "transform_cones")
    assert {k: v for k, v in cd.items() if k not in device_only} == \
        {k: v for k, v in ch.items() if k not in device_only}
    for t in (0, 1):
        _same_tree(dev.plan_tree(t), host.plan_tree(t))
    for far in (True, False):
        for x, y in zip(dev.plan_pairs(far), host.plan_pairs(far)):
            assert np.array_equal(x, y)
    cd, ch = dev.plan_cones(), host.plan_cones()
    for k in cd:
        assert np.array_equal(cd[k], ch[k]), k


def lattice_with_signed_zeros():
    """Dyadic lattice points symmetric about the origin (the padded box is centred at 0, so
    normalised coordinates hit +0, -0 and child-box faces exactly), with -0.0 copies."""
    g = np.arange(-8, 9) / 8.0
    x = np.stack(np.meshgrid(g, g, g, indexing="ij"), -1).reshape(-1, 3)
    z = x.copy()
    z[z == 0.0] = -0.0
    return np.concatenate([x, z[::3]])


def near_ties(seed):
    """Points that agree with another point to below 2^-40 of the box (equal keys, ordered by
    the exact comparison) and exact duplicates (equal under the comparison: index order)."""
    x = sphere(4000, seed)
    rng = np.random.default_rng(seed + 1)
    pick = rng.choice(4000, 600, replace=False)
    close = x[pick] * (1.0 + rng.choice([-1.0, 1.0], (600, 1)) * 3e-14)
    dup = x[pick[:300]].copy()
    y = np.concatenate([x, close, dup])
    return y[rng.permutation(len(y))]


def clusters(seed):
    """Dense clusters of very different sizes (deep, uneven trees)."""
    rng = np.random.default_rng(seed)
    parts = [rng.standard_normal((n, 3)) * s + c for n, s, c in
             [(6000, 1e-3, (0.3, 0.1, -0.2)), (3000, 1e-6, (-0.5, 0.4, 0.1)),
              (4000, 0.5, (0.0, 0.0, 0.0)), (500, 1e-9, (0.1, -0.7, 0.3))]]
    return np.concatenate(parts)


def long_run(seed):
    """1,500 copies of one point (one run of equal keys longer than a device thread sorts)."""
    x = sphere(3000, seed)
    return np.concatenate([x[:1500], np.repeat(x[1500:1501], 1500, axis=0), x[1500:]])


CASES = {
    "sphere": (lambda: sphere(20000, 501), 64),
    "spheroid": (lambda: spheroid(15000, 502), 40),
    "torus": (lambda: torus(12000, 503)[0], 64),
    "lattice_signed_zeros": (lattice_with_signed_zeros, 16),
    "near_ties_and_duplicates": (lambda: near_ties(504), 32),
    "clusters": (lambda: clusters(505), 64),
    "tiny": (lambda: sphere(7, 506), 64),
    "long_run_of_equal_keys": (lambda: long_run(507), 2048),
}


@pytest.mark.parametrize("case", list(CASES))
def test_device_tree_matches_host_tree(gpu, case):
    make, leaf = CASES[case]
    x = make()
    y = sphere(3000, 510) * 1.7 if case in ("sphere", "clusters") else x
    print(f"device tree: case={case} N={len(x)} M={len(y)} leaf={leaf}", flush=True)
    dev = ifgf_amd.HelmholtzIfgfOperator(6.0, leaf, 6, 1, -1)
    dev.init(x.T, y.T)
    host = ifgf_amd.HelmholtzIfgfOperator(6.0, leaf, 6, 1, -1)
    host.init(x.T, y.T, host_plan=True)
    _plans_equal(dev, host)
    w = cdens(len(x), 511)
    a, b = dev.mult(w), host.mult(w)
    assert np.array_equal(a, b)


def test_device_tree_full_size_sphere(gpu):
    """N = 1,048,576 on the unit sphere, leaf 64: the device-built trees against the host's,
    with both init times."""
    x = sphere(1 << 20, 1002)
    print("device tree: N=1048576 sphere leaf=64", flush=True)
    host = ifgf_amd.HelmholtzIfgfOperator(100.0, 64, 6, 1, -1)
    t0 = time.time()
    host.init(x.T, x.T, plan_only=True)
    th = time.time() - t0
    dev = ifgf_amd.HelmholtzIfgfOperator(100.0, 64, 6, 1, -1)
    t0 = time.time()
    dev.init(x.T, x.T)
    td = time.time() - t0
    print(f"device tree: init with host trees (plan only) {th:.2f} s, with device trees and "
          f"upload {td:.2f} s", flush=True)
    for t in (0, 1):
        _same_tree(dev.plan_tree(t), host.plan_tree(t))
    cd, ch = dev.plan_counts(), host.plan_counts()
    for k in ("far_pairs", "near_pairs", "coarse_cones", "fine_cones"):
        assert cd[k] == ch[k], k


def test_device_plan_matches_oracle_on_face_points(gpu):
    """Points placed exactly on the child-box faces of the first three levels (the closed-box
    scan, octree.hpp:1106-1199, and the Morton order decide every tie): the device-built plan
    against the oracle directly (trees, interaction lists, cones), and its mult against the
    oracle's."""
    from oracle import pyoracle as po
    from test_oracle_pins import faces_input
    from test_plan import assert_same_plan
    x, _ = faces_input()
    print(f"device plan vs oracle: face points N={len(x)} leaf=16", flush=True)
    op = ifgf_amd.HelmholtzIfgfOperator(4.0, 16, 6, 1, -1, precision="fp64")
    op.init(x.T, x.T)
    op._pts = (x, x)
    orc = po.Oracle(po.HELMHOLTZ, 4.0, 16, 6).init(x, x)
    assert_same_plan(op, orc)
    w = cdens(len(x), 512)
    u, ref = op.mult(w), orc.mult(w)
    assert np.linalg.norm(u - ref) / np.linalg.norm(ref) < 1e-11
