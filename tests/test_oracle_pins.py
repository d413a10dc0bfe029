"""Pins the parts of the CPU oracle that the end-to-end direct-sum checks leave open:

* the double-layer and combined-field mults against FP64 direct sums written here in numpy
  from the kernels' closed forms (double_layer_helmholtz_ifgf.hpp:58-66,
  combined_field_helmholtz_ifgf.hpp:39-50), so a wrong sign or a dropped term of the
  factored form, the centre factor or the transfer factor fails;
* estimateRefinement (ifgfoperator.hpp:108-244, util.hpp:214-234): the slice norm on arrays
  whose slice energies are known in closed form, the orders at the two extreme tolerances
  (nothing refined at tol = 1; every axis refined to the iteration cap at tol = 1e-300), and
  one sampling step against node positions, kernel sums and a Chebyshev solve done in numpy;
* the closed-box greedy child scan of the octree (octree.hpp:1106-1199) on points placed
  exactly on child-box faces, against a brute-force scan written here; a half-open scan
  gives another tree on the same input, so the check separates the two rules.

The library's host plan is compared with the oracle on the same inputs (orders and trees)."""
import numpy as np
import pytest
from numpy.polynomial import chebyshev as C

import ifgf_amd
from conftest import cdens, sphere, torus
from oracle import pyoracle as po


def _G(k, d):
    return np.exp(-k * d) / (4 * np.pi * d)


def direct_normals(kind, k, x, n, w, y, idx):
    """FP64 direct sums of the double-layer and combined-field kernels at y[idx] over the
    sources x with normals n (self terms skipped), from the closed forms:
    double layer  -dG/dn_src = (k + 1/d) G (x - y).n / d  with G = e^{-k d}/(4 pi d);
    combined      dG/dn_src - i kappa G  with k = -i kappa."""
    out = np.zeros(len(idx), np.complex128)
    for a, j in enumerate(idx):
        r = x - y[j]
        d = np.sqrt((r * r).sum(axis=1))
        m = d > 0
        rn = (r[m] * n[m]).sum(axis=1)
        dm = d[m]
        if kind == po.DOUBLELAYER:
            g = (k + 1.0 / dm) * _G(k, dm) * rn / dm
        else:
            kap = float(np.real(k))
            kk = -1j * kap
            g = -(kk + 1.0 / dm) * _G(kk, dm) * rn / dm - 1j * kap * _G(kk, dm)
        out[a] = (g * w[m]).sum()
    return out


def test_normals_direct_sums_are_the_kernel_derivatives():
    """The numpy direct sums above against central differences of G along the source
    normal, on one point pair per kernel (the closed forms themselves are pinned here)."""
    rng = np.random.default_rng(40)
    x, y = rng.standard_normal((1, 3)), rng.standard_normal((1, 3))
    n = rng.standard_normal((1, 3))
    n /= np.linalg.norm(n)
    h = 1e-6
    w = np.ones(1, np.complex128)
    k = 0.5 + 9.0j
    fd = (_G(k, np.linalg.norm(x + h * n - y)) - _G(k, np.linalg.norm(x - h * n - y))) / (2 * h)
    assert np.isclose(direct_normals(po.DOUBLELAYER, k, x, n, w, y, [0])[0], -fd, rtol=1e-6)
    kap = 9.0
    kk = -1j * kap
    fd = (_G(kk, np.linalg.norm(x + h * n - y)) - _G(kk, np.linalg.norm(x - h * n - y))) / (2 * h)
    want = fd - 1j * kap * _G(kk, np.linalg.norm(x - y))
    assert np.isclose(direct_normals(po.COMBINED, kap, x, n, w, y, [0])[0], want, rtol=1e-6)


@pytest.mark.parametrize("kind,k", [(po.DOUBLELAYER, 0.5 + 9.0j), (po.DOUBLELAYER, 2.35 + 12.0j),
                                    (po.COMBINED, 9.0)])
@pytest.mark.parametrize("targets", ["same", "other"])
def test_normals_mult_matches_direct_sum(kind, k, targets):
    """Oracle double-layer / combined-field mults at order 8 against the numpy direct sums:
    the IFGF error is ~1e-3 of the result; a sign error in any factor is O(1)."""
    N = 4000
    x, n = torus(N, 41)
    y = x if targets == "same" else sphere(2500, 42) * 2.0
    w = cdens(N, 43)
    o = po.Oracle(kind, k, 64, 8, 1, -1).init(x, y, normals=n)
    u = o.mult(w)
    idx = np.arange(0, len(y), 11)
    d = direct_normals(kind, k, x, n, w, y, idx)
    err = np.linalg.norm(u[idx] - d) / np.linalg.norm(d)
    assert err < 3e-3, err
    # the oracle's own direct sum (used by the GPU tests) agrees with the closed form
    assert np.allclose(o.direct(w, idx), d, rtol=1e-10, atol=1e-12 * np.abs(d).max())


# --- estimateRefinement -------------------------------------------------------------------
def test_slice_norm_closed_forms():
    """util.hpp:214-234 with layers = 3: |slice ns[axis] - 3| / max(1, |all|), flat index
    i0 + n0 (i1 + n1 i2) (axis 0 fastest)."""
    ns = (4, 6, 5)
    ones = np.ones(120, np.complex128)
    # all ones: |all|^2 = 120; a slice normal to axis a holds 120 / ns[a] entries
    for a in range(3):
        assert np.isclose(po.slice_norm(ones, ns, a, 3), np.sqrt(1.0 / ns[a]), rtol=1e-15)
    # one entry 0.5 + 0i at (i0, i1, i2) = (1, 3, 2): |all| = 0.5 < 1, so no division;
    # slice index ns - 3 = (1, 3, 2) on every axis, so every axis sees it
    v = np.zeros(120, np.complex128)
    v[1 + 4 * (3 + 6 * 2)] = 0.5
    for a in range(3):
        assert po.slice_norm(v, ns, a, 3) == 0.5
    # moved to i0 = 3 (the last layer, not ns - 3): axis 0 no longer sees it
    v = np.zeros(120, np.complex128)
    v[3 + 4 * (3 + 6 * 2)] = 3 + 4j
    assert po.slice_norm(v, ns, 0, 3) == 0.0
    assert po.slice_norm(v, ns, 1, 3) == 1.0  # |5| / max(1, 5)
    assert po.slice_norm(v, ns, 2, 3) == 1.0
    # a geometric decay along axis 2: the slice i2 = 2 against the whole array
    c = np.zeros((5, 6, 4), np.complex128)  # [i2, i1, i0]
    for i2 in range(5):
        c[i2] = 10.0 ** (-i2)
    flat = c.ravel()
    want = np.sqrt(24 * 1e-4) / np.sqrt(24 * sum(10.0 ** (-2 * i) for i in range(5)))
    assert np.isclose(po.slice_norm(flat, ns, 2, 3), want, rtol=1e-14)


def _library_orders(k, order, tol, x):
    op = ifgf_amd.HelmholtzIfgfOperator(k, 32, order, 1, tol)
    op.init(x.T, x.T, plan_only=True)
    return op.plan_orders()


@pytest.mark.parametrize("order", [4, 6])
def test_estimate_extreme_tolerances(order):
    """tol = 1: every slice norm is <= 1, so the first pass refines nothing and the base
    orders stay (p, p, p) (no radial reduction on the tolerance path, ifgfoperator.hpp:78-87).
    tol = 1e-300: every axis is refined each pass until the cap maxR = 10 stops the loop; the
    result is the last pass's orders, p + 9 per axis, not the bumped p + 10."""
    x = sphere(1200, 44)
    for tol, want in ((1.0, order), (1e-300, order + 9)):
        o = po.Oracle(po.HELMHOLTZ, 6.0, 32, order, 1, tol).init(x, x)
        assert o.base_order().tolist() == [want] * 3, (tol, o.base_order())
        lib = _library_orders(6.0, order, tol, x)
        assert lib["base"].tolist() == [want] * 3, (tol, lib)


def test_estimate_sample_matches_numpy():
    """One sampling step (ifgfoperator.hpp:181-217) on two boxes of the deepest level: the
    nodes are element 0 of the coarse grid (elements (1,2,4) ceil(max(kappa H / 4, 1)) on
    [1e-4, sqrt(3)/3] x [0, pi] x [-pi, pi]) mapped to Cartesian coordinates, the samples are
    sum_i e^{i kappa (d - dc)} dc / d over the box's sources, and the coefficients solve the
    tensor Chebyshev Vandermonde system on the nodes (numpy's chebvander3d)."""
    kappa = 24.0
    x = sphere(3000, 45)
    o = po.Oracle(po.HELMHOLTZ, kappa, 32, 6, 1, 1e-4).init(x, x)
    lv = o.levels(0) - 1
    t = o.tree_level(0, lv)
    pts_sorted = x[o.perm(0)]
    nb = len(t["begin"])
    order = (5, 7, 6)
    for box in (0, (nb // 10) * 3):
        y, data, trafo, info = o.estimate_sample(lv, box, order)
        bb = t["bbox"][box]
        H = (bb[3:] - bb[:3]).max()
        xc = (bb[:3] + bb[3:]) / 2
        assert np.isclose(info["H"], H, rtol=1e-15) and np.allclose(info["centre"], xc)
        els = np.array([1, 2, 4]) * int(np.ceil(max(kappa * H / 4, 1.0)))
        assert info["nels"].tolist() == els.tolist() and info["smin"] == 1e-4
        lo = np.array([1e-4, 0.0, -np.pi])
        hi = np.array([np.sqrt(3.0) / 3, np.pi, np.pi])
        h = (hi - lo) / els
        xi = [np.cos((2 * np.arange(n) + 1) * np.pi / (2 * n)) for n in order]
        X2, X1, X0 = np.meshgrid(xi[2], xi[1], xi[0], indexing="ij")
        xi3 = np.stack([X0.ravel(), X1.ravel(), X2.ravel()], axis=1)  # axis 0 fastest
        sp = xi3 * h / 2 + lo + h / 2
        s, th, ph = sp[:, 0], sp[:, 1], sp[:, 2]
        want_y = xc + (H / s)[:, None] * np.stack(
            [np.cos(ph) * np.sin(th), np.sin(ph) * np.sin(th), np.cos(th)], axis=1)
        assert np.allclose(y, want_y, rtol=1e-13, atol=1e-13)
        src = pts_sorted[t["begin"][box]:t["end"][box]]
        d = np.linalg.norm(want_y[:, None, :] - src[None, :, :], axis=2)
        dc = np.linalg.norm(want_y - xc, axis=1)[:, None]
        want_data = (np.exp(1j * kappa * (d - dc)) * dc / d).sum(axis=1)
        assert np.allclose(data, want_data, rtol=1e-11)
        V = C.chebvander3d(xi3[:, 0], xi3[:, 1], xi3[:, 2], [n - 1 for n in order])
        cv = np.linalg.solve(V, want_data)  # index (i0 n1 + i1) n2 + i2
        want_trafo = cv.reshape(order).transpose(2, 1, 0).ravel()  # i0 + n0 (i1 + n1 i2)
        assert np.allclose(trafo, want_trafo, rtol=1e-9, atol=1e-9 * np.abs(want_trafo).max())


# --- closed-box child scan (octree.hpp:1106-1199) ------------------------------------------
def _scan_tree(pts, root, leaf, closed):
    """Brute-force octree of the sorted points: per box, children j = 0..7 in order (bit 0 =
    x), each the greedy run of points from the current position inside child box j (closed:
    lo <= p <= hi; otherwise lo <= p < hi); empty children are dropped.  Returns per level a
    list of (begin, end, lo, hi)."""
    levels = []
    cur = [(0, len(pts), root[:3].copy(), root[3:].copy())]
    while cur:
        levels.append(cur)
        nxt = []
        for b, e, lo, hi in cur:
            if e - b <= leaf:
                continue
            size = 0.5 * (hi - lo)
            pos = b
            for j in range(8):
                t = np.array([(j >> d) & 1 for d in range(3)], dtype=np.float64)
                clo = lo + size * t
                chi = clo + size
                end = pos
                while end < e:
                    p = pts[end]
                    inside = np.all(clo <= p) and (np.all(p <= chi) if closed else np.all(p < chi))
                    if not inside:
                        break
                    end += 1
                if end > pos:
                    nxt.append((pos, end, clo, chi))
                pos = end
        cur = nxt
    return levels


def faces_input():
    """Random points plus points with coordinates set exactly to the child-box faces of the
    first three levels (computed with the tree's own arithmetic from the padded root box)."""
    rng = np.random.default_rng(46)
    base = rng.uniform(-1, 1, (1500, 3))
    base[0] = (-1, -1, -1)
    base[1] = (1, 1, 1)
    orc = po.Oracle(po.HELMHOLTZ, 4.0, 16, 6).init(base, base)
    root = orc.root_box(0)
    faces = [[] for _ in range(3)]
    boxes = [(root[:3], root[3:])]
    for _ in range(3):
        nxt = []
        for lo, hi in boxes:
            size = 0.5 * (hi - lo)
            for d in range(3):
                faces[d].append((lo + size)[d])
            for j in range(8):
                t = np.array([(j >> d) & 1 for d in range(3)], dtype=np.float64)
                nxt.append((lo + size * t, lo + size * t + size))
        boxes = nxt
    extra = rng.uniform(-1, 1, (1500, 3))
    for i in range(len(extra)):
        for d in range(3):
            if rng.random() < 0.6:
                extra[i, d] = faces[d][rng.integers(len(faces[d]))]
    return np.concatenate([base, extra]), root


def test_child_scan_uses_closed_boxes():
    x, root = faces_input()
    leaf = 16
    o = po.Oracle(po.HELMHOLTZ, 4.0, leaf, 6).init(x, x)
    assert np.array_equal(o.root_box(0), root)  # the face points kept the root box
    pts = x[o.perm(0)]
    want = _scan_tree(pts, root, leaf, closed=True)
    assert len(want) == o.levels(0)
    for l, boxes in enumerate(want):
        t = o.tree_level(0, l)
        assert t["begin"].tolist() == [b for b, _, _, _ in boxes], l
        assert t["end"].tolist() == [e for _, e, _, _ in boxes], l
        assert np.array_equal(t["bbox"], np.array([np.concatenate([lo, hi]) for _, _, lo, hi in boxes]))
    # the input separates the rules: a half-open scan builds another tree
    other = _scan_tree(pts, root, leaf, closed=False)
    assert [[(b, e) for b, e, _, _ in lv] for lv in other] != \
        [[(b, e) for b, e, _, _ in lv] for lv in want]
    # the library's host build is the oracle's
    op = ifgf_amd.HelmholtzIfgfOperator(4.0, leaf, 6, 1, -1)
    op.init(x.T, x.T, plan_only=True)
    tr = op.plan_tree(0)
    # the same point sequence (identical points may come in either order, util.hpp:14-22)
    assert np.array_equal(x[tr["perm"]], pts)
    for l in range(o.levels(0)):
        sl = slice(tr["level_off"][l], tr["level_off"][l + 1])
        ref = o.tree_level(0, l)
        assert np.array_equal(tr["begin"][sl], ref["begin"])
        assert np.array_equal(tr["end"][sl], ref["end"])
