"""Plan files (ifgf_plan_save / ifgf_init_from_plan) on the host: a saved and reloaded plan
is the same plan (every accessor, and the file written again byte for byte), for several
kernel families, targets other than the sources, the tolerance estimate and merged
geometries; damaged, foreign and mismatched files are refused and leave the operator as it
was; a file whose checksum is made to match after a change in its arrays is refused by the
array checks."""
import struct

import numpy as np
import pytest

import ifgf_amd
from conftest import sphere, spheroid
from ifgf_amd._lib import IfgfError

FNV_PRIME = 0x100000001B3
MASK = (1 << 64) - 1


def checksum(payload):
    """The file's checksum: FNV-1a steps over the payload's 8-byte little-endian words (the
    last one zero-padded), then over the byte count."""
    pad = (-len(payload)) % 8
    words = np.frombuffer(payload + b"\0" * pad, dtype="<u8")
    h = 0xCBF29CE484222325
    for w in words.tolist():
        h = ((h ^ w) * FNV_PRIME) & MASK
    return ((h ^ len(payload)) * FNV_PRIME) & MASK


def plan_state(op):
    st = {"counts": op.plan_counts(), "orders": op.plan_orders(), "work": op.plan_work(),
          "far": op.plan_pairs(True), "near": op.plan_pairs(False), "cones": op.plan_cones()}
    for t in (0, 1):
        st[f"tree{t}"] = op.plan_tree(t)
    return st


def assert_same(a, b, path=""):
    if isinstance(a, dict):
        assert a.keys() == b.keys(), path
        for k in a:
            assert_same(a[k], b[k], f"{path}.{k}")
    elif isinstance(a, (tuple, list)):
        assert len(a) == len(b), path
        for i, (x, y) in enumerate(zip(a, b)):
            assert_same(x, y, f"{path}[{i}]")
    else:
        assert np.array_equal(np.asarray(a), np.asarray(b)), path


def cases():
    x = spheroid(3000, 41)
    y = sphere(1500, 42) * 1.3
    n = x * np.array([1.0, 1.0, 4.0])
    n /= np.linalg.norm(n, axis=1)[:, None]
    g = lambda: ifgf_amd.GradHelmholtzIfgfOperator(1.0 + 8.0j, 40, 8, 1, -1)
    return [
        ("helmholtz_tol", lambda: ifgf_amd.HelmholtzIfgfOperator(20.0, 64, 6, 1, 1e-5), x, x,
         None),
        ("helmholtz_targets", lambda: ifgf_amd.HelmholtzIfgfOperator(9.0, 32, 8, 1, -1), x, y,
         None),
        ("laplace", lambda: ifgf_amd.LaplaceIfgfOperator(40, 7, 1, -1), x, x, None),
        ("double_layer", lambda: ifgf_amd.DoubleLayerHelmholtzIfgfOperator(
            0.5 + 9.0j, 40, 8, 1, -1), x, y, n),
        ("grad", g, x, y, None),
    ]


@pytest.mark.parametrize("name", [c[0] for c in cases()])
def test_round_trip_is_the_same_plan(tmp_path, name):
    _, make, x, y, n = next(c for c in cases() if c[0] == name)
    op = make()
    if n is None:
        op.init(x.T, y.T, plan_only=True)
    else:
        op.init(x.T, y.T, n.T, plan_only=True)
    p1, p2 = tmp_path / "a.plan", tmp_path / "b.plan"
    op.save_plan(p1)
    op2 = make()
    op2.load_plan(p1, plan_only=True)
    assert_same(plan_state(op), plan_state(op2))
    op2.save_plan(p2)
    assert p1.read_bytes() == p2.read_bytes()


def test_merged_plan_round_trip(tmp_path):
    a = ifgf_amd.HelmholtzIfgfOperator(8.0, 64, 6, 1, -1)
    b = ifgf_amd.HelmholtzIfgfOperator(12.0, 64, 6, 1, -1)
    a.init(sphere(2500, 43).T, sphere(2500, 43).T, plan_only=True)
    b.init(spheroid(2000, 44).T, spheroid(2000, 44).T, plan_only=True)
    m = ifgf_amd.merge([a, b], plan_only=True)
    m.save_plan(tmp_path / "m.plan")
    op = ifgf_amd.HelmholtzIfgfOperator(8.0, 64, 6, 1, -1)
    op.load_plan(tmp_path / "m.plan", plan_only=True)
    assert_same(plan_state(m), plan_state(op))
    assert (op._n, op._m) == (4500, 4500)


def test_refusals_leave_the_operator_as_it_was(tmp_path):
    x = sphere(2000, 45)
    op = ifgf_amd.HelmholtzIfgfOperator(10.0, 64, 6, 1, -1)
    op.init(x.T, x.T, plan_only=True)
    good = tmp_path / "p.plan"
    op.save_plan(good)
    raw = good.read_bytes()
    before = plan_state(op)

    def refused(data, what, make=lambda: op):
        f = tmp_path / "bad.plan"
        f.write_bytes(data)
        target = make()
        with pytest.raises(IfgfError, match=what):
            target.load_plan(f, plan_only=True)

    other = sphere(1000, 46)
    op_other = ifgf_amd.HelmholtzIfgfOperator(10.0, 64, 6, 1, -1)
    op_other.init(other.T, other.T, plan_only=True)
    state_other = plan_state(op_other)
    refused(raw[:-1], "truncated", lambda: op_other)
    assert_same(state_other, plan_state(op_other))
    flipped = bytearray(raw)
    flipped[len(raw) // 2] ^= 0x10
    refused(bytes(flipped), "checksum", lambda: op_other)
    assert_same(state_other, plan_state(op_other))
    refused(b"#!/bin/sh\n" + raw[10:], "not a plan file")
    refused(raw[:8] + struct.pack("<I", 99) + raw[12:], "format version")
    refused(raw[:12] + struct.pack("<I", 0x04030201) + raw[16:], "byte order")
    with pytest.raises(IfgfError, match="cannot open"):
        op.load_plan(tmp_path / "missing.plan", plan_only=True)
    for make in (lambda: ifgf_amd.HelmholtzIfgfOperator(11.0, 64, 6, 1, -1),
                 lambda: ifgf_amd.HelmholtzIfgfOperator(10.0, 32, 6, 1, -1),
                 lambda: ifgf_amd.HelmholtzIfgfOperator(10.0, 64, 8, 1, -1),
                 lambda: ifgf_amd.HelmholtzIfgfOperator(10.0, 64, 6, 1, 1e-4),
                 lambda: ifgf_amd.MofifiedHelmholtzIfgfOperator(-10.0j, 64, 6, 1, -1)):
        with pytest.raises(IfgfError, match="another kernel"):
            make().load_plan(good, plan_only=True)
    assert_same(before, plan_state(op))
    with pytest.raises(IfgfError, match="before init"):
        ifgf_amd.HelmholtzIfgfOperator(10.0, 64, 6, 1, -1).save_plan(tmp_path / "none.plan")


def test_array_checks_behind_a_matching_checksum(tmp_path):
    """Change the source permutation (a duplicate index) and a parent index (out of range)
    and recompute the checksum: the array checks refuse both."""
    x = sphere(1200, 47)
    op = ifgf_amd.HelmholtzIfgfOperator(6.0, 64, 6, 1, -1)
    op.init(x.T, x.T, plan_only=True)
    f = tmp_path / "p.plan"
    op.save_plan(f)
    raw = f.read_bytes()
    head, payload = raw[:16], bytearray(raw[16:-16])
    assert struct.unpack("<QQ", raw[-16:]) == (len(payload), checksum(bytes(payload)))
    # payload: 48-byte record, kernel spec (4 + 8 + 8 + 8 + 4 + 8), orders and elements
    # (24 + 4 * 12), then the source tree: n, pts (count, 3n doubles), perm (count, n)
    at = 48 + 40 + 72
    n = struct.unpack_from("<q", payload, at)[0]
    assert n == 1200
    at += 8
    assert struct.unpack_from("<q", payload, at)[0] == 3 * n
    at += 8 + 24 * n
    assert struct.unpack_from("<q", payload, at)[0] == n
    perm_at = at + 8

    def forged(data, what):
        g = tmp_path / "forged.plan"
        g.write_bytes(head + bytes(data) + struct.pack("<QQ", len(data), checksum(bytes(data))))
        with pytest.raises(IfgfError, match=what):
            ifgf_amd.HelmholtzIfgfOperator(6.0, 64, 6, 1, -1).load_plan(g, plan_only=True)

    dup = bytearray(payload)
    dup[perm_at + 8:perm_at + 16] = dup[perm_at:perm_at + 8]
    forged(dup, "permutation")
    # after perm: lo, hi (48 bytes), levels (4), level_off, begin, end, then parent
    at = perm_at + 8 * n + 48
    levels = struct.unpack_from("<i", payload, at)[0]
    at += 4
    assert struct.unpack_from("<q", payload, at)[0] == levels + 1
    nb = struct.unpack_from("<q", payload, at + 8 + 8 * levels)[0]
    at += 8 + 8 * (levels + 1)
    for _ in range(2):  # begin, end
        assert struct.unpack_from("<q", payload, at)[0] == nb
        at += 8 + 8 * nb
    assert struct.unpack_from("<q", payload, at)[0] == nb
    bad_parent = bytearray(payload)
    struct.pack_into("<i", bad_parent, at + 8 + 4 * (nb - 1), nb + 5)
    forged(bad_parent, "box ranges")
