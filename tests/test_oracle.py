"""Pins the CPU oracle (oracle/ifgf_oracle.cpp) with checks that do not restate its code:
closed forms of the Chebyshev machinery, numpy's own Chebyshev routines, an exact big-integer
z-order, brute-force admissibility and exact-once coverage of the interaction lists, finite
differences of the kernel families, and the FP64 direct sum."""
import math

import numpy as np
import pytest
from numpy.polynomial import chebyshev as C

from conftest import cdens, sphere
from oracle import pyoracle as po


# --- Chebyshev (chebinterp.hpp:27-346) -------------------------------------------------------
@pytest.mark.parametrize("n", [1, 2, 3, 6, 9])
def test_chebnodes_closed_form(n):
    x = po.chebnodes1d(n)
    # roots of T_n, descending (chebinterp.hpp:33-36)
    assert np.allclose(C.chebval(x, [0] * n + [1]), 0, atol=1e-13)
    assert np.all(np.diff(x) < 0)


def _nodes3(ns):
    r, t, p = (po.chebnodes1d(k) for k in ns)
    P, T, R = np.meshgrid(p, t, r, indexing="ij")  # axis 0 fastest in the flat index
    return np.stack([R.ravel(), T.ravel(), P.ravel()], axis=1)


@pytest.mark.parametrize("ns", [(3, 6, 6), (2, 3, 3), (5, 8, 8), (1, 2, 4)])
def test_chebtransform_of_basis_is_unit_vector(ns):
    pts = _nodes3(ns)
    rng = np.random.default_rng(0)
    for _ in range(4):
        a, b, c = (int(rng.integers(0, k)) for k in ns)
        vals = (C.chebval(pts[:, 0], np.eye(ns[0])[a]) * C.chebval(pts[:, 1], np.eye(ns[1])[b]) *
                C.chebval(pts[:, 2], np.eye(ns[2])[c])).astype(complex)
        coef = po.chebtransform(ns, vals)
        e = np.zeros(ns[0] * ns[1] * ns[2])
        e[a + ns[0] * (b + ns[1] * c)] = 1
        assert np.allclose(coef, e, atol=1e-13)


@pytest.mark.parametrize("ns", [(3, 6, 6), (2, 5, 5), (4, 4, 1)])
def test_clenshaw_matches_numpy_chebval3d(ns):
    rng = np.random.default_rng(1)
    coef = rng.standard_normal(np.prod(ns)) + 1j * rng.standard_normal(np.prod(ns))
    pts = rng.uniform(-1, 1, (50, 3))
    got = po.clenshaw(ns, coef, pts)
    cube = coef.reshape(ns[2], ns[1], ns[0]).transpose(2, 1, 0)  # c[i0, i1, i2]
    want = C.chebval3d(pts[:, 0], pts[:, 1], pts[:, 2], cube)
    assert np.allclose(got, want, rtol=1e-12, atol=1e-12)


def test_interpolation_exact_on_polynomials():
    ns = (3, 6, 6)
    pts = _nodes3(ns)
    rng = np.random.default_rng(2)
    cube = rng.standard_normal(ns)  # degree < n per axis
    f = lambda p: C.chebval3d(p[:, 0], p[:, 1], p[:, 2], cube)
    coef = po.chebtransform(ns, f(pts).astype(complex))
    q = rng.uniform(-1, 1, (40, 3))
    assert np.allclose(po.clenshaw(ns, coef, q), f(q), atol=1e-12)


def test_chebinterp_known_answer():
    """test/test_chebinterp.cpp:23-65: exp(|x|^2 + z) at order 2 on [-1,1]^3; the reference
    prints the RMS error, here the oracle's interpolant must equal numpy's interpolant."""
    ns = (2, 2, 2)
    pts = _nodes3(ns)
    f = lambda p: np.exp((p ** 2).sum(axis=1) + p[:, 2])
    coef = po.chebtransform(ns, f(pts).astype(complex))
    q = np.random.default_rng(3).uniform(-1, 1, (1000, 3))
    got = po.clenshaw(ns, coef, q).real
    # numpy: least squares on the 8 nodes with the tensor Chebyshev basis = interpolation
    V = C.chebvander3d(pts[:, 0], pts[:, 1], pts[:, 2], [1, 1, 1])
    c = np.linalg.solve(V, f(pts))
    want = C.chebvander3d(q[:, 0], q[:, 1], q[:, 2], [1, 1, 1]) @ c
    assert np.allclose(got, want, atol=1e-12)
    rms = np.sqrt(np.mean((got - f(q)) ** 2))
    assert 0.1 < rms < 10  # order 2 is a coarse fit of this function (the reference's demo)


# --- cone coordinates (util.hpp:96-171) ------------------------------------------------------
def test_cone_coordinate_round_trip():
    rng = np.random.default_rng(4)
    xc = rng.standard_normal(3)
    for _ in range(50):
        p = np.array([rng.uniform(1e-3, 0.577), rng.uniform(0, np.pi), rng.uniform(-np.pi, np.pi)])
        x = po.interp_to_cart(p, xc, 0.3)
        assert np.isclose(np.linalg.norm(x - xc), 0.3 / p[0], rtol=1e-14)
        assert np.allclose(po.cart_to_interp(x, xc, 0.3), p, rtol=1e-12, atol=1e-12)


# --- Morton order (zorder_less.hpp:212-248) --------------------------------------------------
def _exact_key(v):
    """z-order word of one normalised coordinate: non-zero flag at bit position 0 (the
    reference's FloatExp(0) == 0), then the exact binary expansion of |v| below the point;
    complemented for negative values (signed comparison on the deciding axis)."""
    BITS = 1100
    neg = v < 0
    m, e = math.frexp(abs(v))
    F = int(m * (1 << 53)) << (e - 53 + BITS) if v != 0 else 0
    w = ((1 if v != 0 else 0) << BITS) | F
    if neg:
        w = ~w & ((1 << (BITS + 1)) - 1)
    return neg, format(w, f"0{BITS + 1}b")


def _zorder_sort(pts, box):
    lo, hi = box[:3], box[3:]
    ctr = (hi + lo) / 2.0
    diam = hi - lo
    keys = []
    for p in pts:
        q = (p - ctr) / diam
        (nx, wx), (ny, wy), (nz, wz) = (_exact_key(v) for v in q)
        bits = ["0" if nz else "1", "0" if ny else "1", "0" if nx else "1"]
        bits += [b for t in zip(wz, wy, wx) for b in t]
        keys.append(int("".join(bits), 2))
    return np.argsort(np.array(keys, dtype=object), kind="stable")


def _morton_inputs():
    rng = np.random.default_rng(5)
    p = rng.uniform(-1, 1, (600, 3))
    p[:40] = np.round(p[:40] * 8) / 8             # dyadic points on box boundaries
    p[40:60] = p[60:80]                           # duplicates
    p[80:100, 0] = 0.0                            # coordinate zero (after centring below)
    p[100:110, 1] = -0.0
    p[110:120] = p[120:130] * (1 + 1e-15)         # agree to ~50 bits
    return p


def test_morton_sort_is_exact_zorder():
    p = _morton_inputs()
    perm, box = po.morton_sort(p)
    ref = _zorder_sort(p, box)
    assert np.array_equal(p[perm], p[ref])  # equal as point sequences (duplicates may swap)


def test_morton_zero_coordinate_rule():
    """Hand-worked from zorder_less.hpp:180-248: with normalised x = 0 vs 0.01 and y, z
    differing at bit -2, FloatExp(0) == 0 makes the x axis decide (p < q), where the plain
    z-order of magnitudes would let z decide (q < p)."""
    box = np.array([-0.5, -0.5, -0.5, 0.5, 0.5, 0.5])
    p = np.array([0.0, 0.3, 0.3])
    q = np.array([0.01, 0.2, 0.2])
    assert po.morton_less(box, p, q)
    assert not po.morton_less(box, q, p)
    # signs first: z, then y, then x (zorder_less.hpp:233-237)
    assert po.morton_less(box, np.array([0.4, 0.4, -0.1]), np.array([-0.4, -0.4, 0.1]))


# --- octree and interaction lists (octree.hpp:202-349, 1069-1103) ---------------------------
def _tree_invariants(o, pts_sorted, leaf):
    n = pts_sorted.shape[0]
    covered = np.zeros(n, int)  # points counted by a leaf on an earlier level
    for l in range(o.levels(0)):
        t = o.tree_level(0, l)
        cnt = covered.copy()
        for b, e, bb, lf in zip(t["begin"], t["end"], t["bbox"], t["leaf"]):
            assert e > b
            seg = pts_sorted[b:e]
            assert np.all(seg >= bb[:3]) and np.all(seg <= bb[3:])
            cnt[b:e] += 1
            if lf:
                assert e - b <= leaf
                covered[b:e] += 1
            else:
                assert e - b > leaf
        assert np.all(cnt == 1), "sanitize: each point in exactly one node per level"


def test_octree_sanitize_and_boxes():
    x = sphere(5000, 11)
    o = po.Oracle(po.HELMHOLTZ, 5.0, 32, 6).init(x, x)
    pts = x[o.perm(0)]
    _tree_invariants(o, pts, 32)
    rb = o.root_box(0)
    lo, hi = x.min(0), x.max(0)
    assert np.all(rb[:3] == lo - 1e-8 * np.linalg.norm(lo))
    assert np.all(rb[3:] == hi + 1e-8 * np.linalg.norm(hi))


def _coverage(o, ns, nt):
    """CHECK_CONNECTIVITY (ifgfoperator.hpp:853-860): every (source, target) pair exactly once."""
    cover = np.zeros((ns, nt), np.int32)
    for l in range(o.levels(0)):
        t = o.tree_level(0, l)
        for far in (True, False):
            for b, tb, te in o.interactions(l, far):
                cover[t["begin"][b]:t["end"][b], tb:te] += 1
    return cover


@pytest.mark.parametrize("targets", ["same", "other"])
def test_interactions_cover_every_pair_once(targets):
    x = sphere(1500, 12)
    y = x if targets == "same" else sphere(1200, 13) * 1.5
    o = po.Oracle(po.HELMHOLTZ, 5.0, 20, 6).init(x, y)
    assert np.all(_coverage(o, len(x), len(y)) == 1)


def test_far_pairs_are_admissible_brute_force():
    x = sphere(2000, 14)
    o = po.Oracle(po.HELMHOLTZ, 5.0, 20, 6).init(x, x)
    tl = {l: o.tree_level(1, l) for l in range(o.levels(1))}
    ranges = {}
    for l, t in tl.items():
        for b, e, bb in zip(t["begin"], t["end"], t["bbox"]):
            ranges.setdefault((b, e), []).append(bb)
    for l in range(o.levels(0)):
        s = o.tree_level(0, l)
        for b, tb, te in o.interactions(l, True):
            bb = s["bbox"][b]
            c = (bb[:3] + bb[3:]) / 2
            H = (bb[3:] - bb[:3]).max()
            ok = False
            for tbb in ranges[(tb, te)]:
                d = np.maximum(np.maximum(tbb[:3] - c, c - tbb[3:]), 0)
                ok |= np.sqrt((d * d).sum()) >= np.sqrt(3.0) * H
            assert ok
        for b, tb, te in o.interactions(l, False):
            bb = s["bbox"][b]
            c = (bb[:3] + bb[3:]) / 2
            H = (bb[3:] - bb[:3]).max()
            for tbb in ranges[(tb, te)]:
                d = np.maximum(np.maximum(tbb[:3] - c, c - tbb[3:]), 0)
                assert np.sqrt((d * d).sum()) < np.sqrt(3.0) * H


# --- kernel families (helmholtz_ifgf.hpp ... grad_helmholtz_ifgf.hpp) -----------------------
def _G(k, d):
    return np.exp(-k * d) / (4 * np.pi * d)


def test_kernel_formulas_and_derivatives():
    rng = np.random.default_rng(6)
    xs, y, n = rng.standard_normal(3), rng.standard_normal(3), rng.standard_normal(3)
    n /= np.linalg.norm(n)
    d = np.linalg.norm(xs - y)
    kap, k = 7.5, 2.35 + 4.4j
    h = 1e-6
    pts = np.stack([xs, y])
    o = po.Oracle(po.HELMHOLTZ, kap, 64, 6).init(pts, pts)
    assert np.isclose(o.kernel(xs, y), np.exp(1j * kap * d) / (4 * np.pi * d), rtol=1e-14)
    assert o.kernel(xs, xs) == 0  # G(0) = 0 (helmholtz_ifgf.hpp:39-40)
    o = po.Oracle(po.MODHELMHOLTZ, k, 64, 8).init(pts, pts)
    assert np.isclose(o.kernel(xs, y), _G(k, d), rtol=1e-14)
    o = po.Oracle(po.LAPLACE, 0, 64, 8).init(pts, pts)
    assert np.isclose(o.kernel(xs, y).real, 1 / (4 * np.pi * d), rtol=1e-14)
    # double layer = -d/dt G(x_src + t n - y)   (double_layer_helmholtz_ifgf.hpp:58-66)
    o = po.Oracle(po.DOUBLELAYER, k, 64, 8).init(pts, pts, normals=np.stack([n, n]))
    fd = (_G(k, np.linalg.norm(xs + h * n - y)) - _G(k, np.linalg.norm(xs - h * n - y))) / (2 * h)
    assert np.isclose(o.kernel(xs, y, 0), -fd, rtol=1e-6)
    # combined field = d/dn_src G - i kappa G with G = e^{i kappa d}/(4 pi d)
    o = po.Oracle(po.COMBINED, kap, 64, 8).init(pts, pts, normals=np.stack([n, n]))
    kk = -1j * kap
    fd = (_G(kk, np.linalg.norm(xs + h * n - y)) - _G(kk, np.linalg.norm(xs - h * n - y))) / (2 * h)
    assert np.isclose(o.kernel(xs, y, 0), fd - 1j * kap * _G(kk, d), rtol=1e-6)
    # gradient: d/dy_dx G (grad_helmholtz_ifgf.hpp:44-58)
    for dx in range(3):
        o = po.Oracle(po.GRAD, k, 64, 6, dx=dx).init(pts, pts)
        e = np.eye(3)[dx]
        fd = (_G(k, np.linalg.norm(xs - y - h * e)) - _G(k, np.linalg.norm(xs - y + h * e))) / (2 * h)
        assert np.isclose(o.kernel(xs, y), fd, rtol=1e-6)


# --- end to end: mult against the FP64 direct sum (test/test_ifgf.cpp:143-165) --------------
@pytest.mark.parametrize("kind,k,order,tol,bound", [
    (po.HELMHOLTZ, 8.0, 6, -1, 2e-3),
    (po.HELMHOLTZ, 8.0, 6, 1e-6, 5e-5),
    (po.MODHELMHOLTZ, 2.35 + 10.0j, 8, -1, 2e-3),
    (po.LAPLACE, 0.0, 8, -1, 1e-3),
])
def test_mult_matches_direct_sum(kind, k, order, tol, bound):
    N = 3000
    x = sphere(N, 15)
    w = cdens(N, 16)
    if kind == po.LAPLACE:
        w = w.real.astype(complex)
    o = po.Oracle(kind, k, 64, order, 1, tol).init(x, x)
    u = o.mult(w)
    idx = np.arange(0, N, 7)
    d = o.direct(w, idx)
    err = np.linalg.norm(u[idx] - d) / np.linalg.norm(d)
    assert err < bound, err


def test_mult_is_linear_and_permutation_covariant():
    N = 2500
    x = sphere(N, 17)
    o = po.Oracle(po.HELMHOLTZ, 6.0, 40, 6).init(x, x)
    a, b = cdens(N, 18), cdens(N, 19)
    ua, ub = o.mult(a), o.mult(b)
    assert np.allclose(o.mult(2 * a - 1j * b), 2 * ua - 1j * ub, rtol=1e-12, atol=1e-12)
    # relabelling the points relabels the result (the octree sorts them anyway)
    p = np.random.default_rng(20).permutation(N)
    o2 = po.Oracle(po.HELMHOLTZ, 6.0, 40, 6).init(x[p], x[p])
    assert np.allclose(o2.mult(a[p]), ua[p], rtol=1e-10, atol=1e-12)
