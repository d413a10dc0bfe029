"""The library's plan (Morton sort, octree, interaction lists, cone activity) against the
oracle, integer for integer.  The plan is built on the host (IFGF_INIT_PLAN_ONLY), so these
run without a GPU."""
import numpy as np
import pytest

import ifgf_amd
from conftest import sphere, spheroid
from oracle import pyoracle as po

KINDS = {
    "helmholtz": (ifgf_amd.HelmholtzIfgfOperator, po.HELMHOLTZ),
    "modhelmholtz": (ifgf_amd.MofifiedHelmholtzIfgfOperator, po.MODHELMHOLTZ),
    "laplace": (ifgf_amd.LaplaceIfgfOperator, po.LAPLACE),
    "doublelayer": (ifgf_amd.DoubleLayerHelmholtzIfgfOperator, po.DOUBLELAYER),
    "combined": (ifgf_amd.CombinedFieldHelmholtzIfgfOperator, po.COMBINED),
    "grad": (ifgf_amd.GradHelmholtzIfgfOperator, po.GRAD),
}


def make_pair(kind, k, leaf, order, n_elem, tol, x, y, normals=None, dx=-1):
    cls, ok = KINDS[kind]
    if kind == "laplace":
        op = cls(leaf, order, n_elem, tol)
    else:
        op = cls(k, leaf, order, n_elem, tol)
    if kind == "grad":
        op.setDx(dx)
    if normals is not None:
        op.init(x.T, y.T, normals.T, plan_only=True)
    else:
        op.init(x.T, y.T, plan_only=True)
    orc = po.Oracle(ok, k, leaf, order, n_elem, tol, dx=dx).init(x, y, normals=normals)
    op._pts = (x, y)
    return op, orc


def assert_same_plan(op, orc):
    o = op.plan_orders()
    assert np.array_equal(o["base"], orc.base_order())
    for tree in (0, 1):
        tr = op.plan_tree(tree)
        pts = op._pts[tree]
        # the same sorted point sequence; indices agree except among identical points, whose
        # order std::sort leaves unspecified (util.hpp:14-22)
        assert np.array_equal(pts[tr["perm"]], pts[orc.perm(tree)])
        if len(np.unique(pts, axis=0)) == len(pts):
            assert np.array_equal(tr["perm"], orc.perm(tree))
        lo = tr["level_off"]
        assert len(lo) - 1 == orc.levels(tree)
        for l in range(orc.levels(tree)):
            ref = orc.tree_level(tree, l)
            sl = slice(lo[l], lo[l + 1])
            assert np.array_equal(tr["begin"][sl], ref["begin"])
            assert np.array_equal(tr["end"][sl], ref["end"])
            assert np.array_equal(tr["leaf"][sl], ref["leaf"])
            assert np.array_equal(tr["bbox"][sl], ref["bbox"])  # bit-exact boxes
            par = tr["parent"][sl]
            if l > 0:
                assert np.array_equal(par - lo[l - 1], ref["parent"])
    src = op.plan_tree(0)
    tgt = op.plan_tree(1)
    lo = src["level_off"]
    for far in (True, False):
        s, t = op.plan_pairs(far)
        lvl = np.searchsorted(lo, s, side="right") - 1
        mine = sorted(zip(lvl.tolist(), (s - lo[lvl]).tolist(), tgt["begin"][t].tolist(),
                          tgt["end"][t].tolist()))
        ref = sorted((l, *map(int, r)) for l in range(orc.levels(0))
                     for r in orc.interactions(l, far))
        assert mine == ref, "far" if far else "near"
    cn = op.plan_cones()
    for l in range(orc.levels(0)):
        b0, b1 = lo[l], lo[l + 1]
        for step in (0, 1):
            nel, nact, ids = orc.cones(l, step)
            assert np.array_equal(cn["nf" if step == 0 else "nc"][b0:b1], nel)
            if step == 1:
                off, cells = cn["cone_off"], cn["cell"]
            else:
                off, cells = cn["fine_off"], cn["fine_ids"]
            assert np.array_equal(np.diff(off[b0:b1 + 1]), nact), (l, step)
            assert np.array_equal(cells[off[b0]:off[b1]], ids), (l, step)
    assert op.plan_counts()["fine_uncovered"] == 0


@pytest.mark.parametrize("kind,k,leaf,order", [
    ("helmholtz", 12.0, 64, 6),
    ("helmholtz", 30.0, 32, 8),
    ("modhelmholtz", 2.35 + 43.6j, 100, 8),
    ("laplace", 0.0, 50, 7),
    ("grad", 1.0 + 12.0j, 64, 6),
])
def test_plan_matches_oracle_sphere(kind, k, leaf, order):
    x = sphere(12000, 21)
    op, orc = make_pair(kind, k, leaf, order, 1, -1, x, x)
    assert_same_plan(op, orc)


def test_plan_matches_oracle_other_targets_and_normals():
    x = spheroid(6000, 22)
    y = sphere(4000, 23) * 0.7 + np.array([0.1, -0.2, 0.05])
    n = x / np.linalg.norm(x, axis=1)[:, None]
    for kind, k in (("doublelayer", 0.5 + 9.0j), ("combined", 9.0)):
        op, orc = make_pair(kind, k, 40, 8, 1, -1, x, y, normals=n)
        assert_same_plan(op, orc)


def test_plan_matches_oracle_tolerance_path_and_n_elem():
    x = sphere(8000, 24)
    op, orc = make_pair("helmholtz", 10.0, 64, 6, 1, 1e-6, x, x)
    assert_same_plan(op, orc)
    op, orc = make_pair("helmholtz", 10.0, 64, 6, 2, -1, x, x)
    assert_same_plan(op, orc)


def test_plan_edge_cases():
    rng = np.random.default_rng(25)
    # duplicates, points on dyadic boundaries, a cube cloud with many near pairs
    x = rng.uniform(-1, 1, (3000, 3))
    x[:300] = np.round(x[:300] * 4) / 4
    x[300:400] = x[400:500]
    op, orc = make_pair("helmholtz", 4.0, 16, 6, 1, -1, x, x)
    assert_same_plan(op, orc)
    # fewer points than a leaf: a single box, near field only
    x = sphere(40, 26)
    op, orc = make_pair("helmholtz", 4.0, 64, 6, 1, -1, x, x)
    assert_same_plan(op, orc)
    assert op.plan_counts()["far_pairs"] == 0


def test_cutoff_drops_admissible_pairs():
    x = sphere(5000, 27)
    op = ifgf_amd.HelmholtzIfgfOperator(10.0, 32, 6, 1, -1)
    op.init(x.T, x.T, cutOff=lambda d: d > 1.0, plan_only=True)
    c = op.plan_counts()
    full = ifgf_amd.HelmholtzIfgfOperator(10.0, 32, 6, 1, -1)
    full.init(x.T, x.T, plan_only=True)
    assert c["dropped"] > 0
    assert c["far_pairs"] + c["dropped"] == full.plan_counts()["far_pairs"]


def test_plan_errors():
    op = ifgf_amd.MofifiedHelmholtzIfgfOperator(1 + 10j, 64, 3, 1, -1)
    with pytest.raises(ifgf_amd._lib.IfgfError):  # fine order = 3 - 3 = 0 (no clamp)
        op.init(sphere(500, 1).T, sphere(500, 1).T, plan_only=True)
    # Laplace at order 6: radial base order 3, fine order 0 (laplace_ifgf.hpp:133-145); the
    # library and the oracle both refuse it
    x = sphere(500, 1)
    with pytest.raises(ifgf_amd._lib.IfgfError):
        ifgf_amd.LaplaceIfgfOperator(64, 6, 1, -1).init(x.T, x.T, plan_only=True)
    with pytest.raises(RuntimeError, match="order below 1"):
        po.Oracle(po.LAPLACE, 0.0, 64, 6, 1, -1).init(x, x)
    ifgf_amd.LaplaceIfgfOperator(64, 7, 1, -1).init(x.T, x.T, plan_only=True)
    op = ifgf_amd.HelmholtzIfgfOperator(1.0, 64, 6, 1, -1)
    with pytest.raises(ValueError):
        op.init(np.zeros((2, 5)), np.zeros((2, 5)), plan_only=True)
    with pytest.raises(ifgf_amd._lib.IfgfError):
        op.mult(np.ones(5))
    with pytest.raises(ValueError):
        ifgf_amd.DoubleLayerHelmholtzIfgfOperator(1j, 64, 8, 1, -1).init(
            sphere(100, 2).T, sphere(100, 2).T, plan_only=True)


def test_merged_plan_is_the_parts_side_by_side():
    """ifgf_merge (host plan): boxes level-major with the parts in order, pairs, cones and
    point ranges of every part carried over with the id offsets recomputed here."""
    geos = [(sphere(3000, 71), 8.0), (spheroid(2500, 72) * 1.3, 15.0), (sphere(900, 73), 3.0)]
    ops = []
    for x, k in geos:
        op = ifgf_amd.HelmholtzIfgfOperator(k, 32, 6, 1, -1)
        op.init(x.T, x.T, plan_only=True)
        ops.append(op)
    m = ifgf_amd.merge(ops, plan_only=True)
    assert (m._n, m._m) == (sum(len(x) for x, _ in geos),) * 2
    cm = m.plan_counts()
    cs = [o.plan_counts() for o in ops]
    for key in ("src_boxes", "far_pairs", "near_pairs", "coarse_cones", "fine_cones"):
        assert cm[key] == sum(c[key] for c in cs), key
    assert cm["src_levels"] == max(c["src_levels"] for c in cs)
    # id maps: level-major, parts in order within a level
    trees = [o.plan_tree(0) for o in ops]
    tm = m.plan_tree(0)
    L = cm["src_levels"]
    maps, nxt, poff = [np.full(c["src_boxes"], -1) for c in cs], 0, [0]
    for x, _ in geos:
        poff.append(poff[-1] + len(x))
    for lv in range(L):
        assert tm["level_off"][lv] == nxt
        for g, t in enumerate(trees):
            if lv < len(t["level_off"]) - 1:
                a, b = t["level_off"][lv], t["level_off"][lv + 1]
                maps[g][a:b] = np.arange(nxt, nxt + (b - a))
                nxt += b - a
    for g, t in enumerate(trees):
        mp = maps[g]
        assert np.array_equal(tm["begin"][mp], t["begin"] + poff[g])
        assert np.array_equal(tm["end"][mp], t["end"] + poff[g])
        assert np.array_equal(tm["leaf"][mp], t["leaf"])
        assert np.array_equal(tm["bbox"][mp], t["bbox"])
        par = np.where(t["parent"] >= 0, mp[np.maximum(t["parent"], 0)], -1)
        assert np.array_equal(tm["parent"][mp], par)
        assert np.array_equal(tm["perm"][poff[g]:poff[g + 1]], t["perm"] + poff[g])
    for far in (True, False):
        s_m, t_m = m.plan_pairs(far)
        got = set(zip(s_m.tolist(), t_m.tolist()))
        want = set()
        for g, o in enumerate(ops):
            s_, t_ = o.plan_pairs(far)
            want |= set(zip(maps[g][s_].tolist(), maps[g][t_].tolist()))
        assert got == want
    cm_ = m.plan_cones()
    for g, o in enumerate(ops):
        c = o.plan_cones()
        for b in range(len(c["cone_off"]) - 1):
            nb = maps[g][b]
            cells = c["cell"][c["cone_off"][b]:c["cone_off"][b + 1]]
            masks = c["mask"][c["cone_off"][b]:c["cone_off"][b + 1]]
            assert np.array_equal(cm_["cell"][cm_["cone_off"][nb]:cm_["cone_off"][nb + 1]], cells)
            assert np.array_equal(cm_["mask"][cm_["cone_off"][nb]:cm_["cone_off"][nb + 1]], masks)
            assert np.array_equal(cm_["nc"][nb], c["nc"][b])
    # different interpolation orders do not merge
    other = ifgf_amd.HelmholtzIfgfOperator(8.0, 32, 8, 1, -1)
    other.init(geos[0][0].T, geos[0][0].T, plan_only=True)
    with pytest.raises(ifgf_amd.IfgfError):
        ifgf_amd.merge([ops[0], other], plan_only=True)


def test_plan_matches_oracle_on_face_points():
    """Points exactly on the child-box faces of the first three levels (the closed-box scan
    and the Morton order decide every tie)."""
    from test_oracle_pins import faces_input
    x, _ = faces_input()
    op, orc = make_pair("helmholtz", 4.0, 16, 6, 1, -1, x, x)
    assert_same_plan(op, orc)
