"""ctypes wrapper of the CPU oracle (oracle/ifgf_oracle.cpp).

TEST INFRASTRUCTURE ONLY: a from-scratch FP64 restatement of TimHRO/ifgf's
IfgfOperator init/mult path used as the checker for the HIP library.  Only
tests/, __graft_entry__.smoke() and bench.py's cpu_baseline leg import it.
"""
import ctypes
import os
import subprocess

import numpy as np

_HERE = os.path.dirname(os.path.abspath(__file__))
_LIB = os.path.join(_HERE, "liboracle.so")

HELMHOLTZ, MODHELMHOLTZ, LAPLACE, DOUBLELAYER, COMBINED, GRAD = range(6)

_libs = {}
_lib = None  # the IEEE checker, once loaded

# builds of the same source (oracle/Makefile): "" the IEEE checker of the tests, "_o3" and
# "_fast" (-O3 for x86-64-v3, IEEE / -ffast-math) for bench.py's CPU baseline
VARIANTS = ("", "_o3", "_fast")


def build():
    subprocess.check_call(["make", "-s", "-C", _HERE])


def lib(variant=""):
    global _lib
    if variant not in _libs:
        path = os.path.join(_HERE, f"liboracle{variant}.so")
        if not os.path.exists(path):
            build()
        L = ctypes.CDLL(path)
        P, D, I, I64 = ctypes.c_void_p, ctypes.c_double, ctypes.c_int, ctypes.c_int64
        L.orc_create.restype = P
        L.orc_create.argtypes = [I, D, D, I64, I, I, D, I]
        L.orc_destroy.argtypes = [P]
        L.orc_last_error.restype = ctypes.c_char_p
        L.orc_last_error.argtypes = [P]
        L.orc_init.argtypes = [P, P, I64, P, I64, P, I]
        L.orc_mult.argtypes = [P, P, P, I]
        L.orc_base_order.argtypes = [P, P]
        L.orc_set_cutoff.argtypes = [P, D]
        L.orc_mult_sample.restype = D
        L.orc_mult_sample.argtypes = [P, P, I, I, P]
        L.orc_levels.argtypes = [P, I]
        L.orc_num_boxes.restype = I64
        L.orc_num_boxes.argtypes = [P, I, I]
        L.orc_perm.argtypes = [P, I, P]
        L.orc_root_box.argtypes = [P, I, P]
        L.orc_tree_level.argtypes = [P, I, I, P, P, P, P, P]
        L.orc_num_interactions.restype = I64
        L.orc_num_interactions.argtypes = [P, I, I]
        L.orc_interactions.argtypes = [P, I, I, P]
        L.orc_cone_counts.argtypes = [P, I, I, P, P]
        L.orc_cone_active.argtypes = [P, I, I, P]
        L.orc_min_rec.argtypes = [P]
        L.orc_chebnodes1d.argtypes = [I, P]
        L.orc_chebtransform.argtypes = [P, P, P]
        L.orc_clenshaw.argtypes = [P, P, P, I64, P]
        L.orc_morton_sort.argtypes = [P, I64, P, P]
        L.orc_morton_less.argtypes = [P, P, P]
        L.orc_cart_to_interp.argtypes = [P, P, D, P]
        L.orc_interp_to_cart.argtypes = [P, P, D, P]
        L.orc_kernel_eval.argtypes = [P, P, P, I64, P]
        L.orc_direct.argtypes = [P, P, I64, P, P, P, P, I64, P, I]
        L.orc_slice_norm.restype = D
        L.orc_slice_norm.argtypes = [P, I64, P, I, I]
        L.orc_estimate_sample.restype = I64
        L.orc_estimate_sample.argtypes = [P, I, I64, P, P, P, P, P]
        _libs[variant] = L
        if variant == "":
            _lib = L
    return _libs[variant]




def _p(a):
    return a.ctypes.data_as(ctypes.c_void_p)


def _pts(a):
    """(3, N) or (N, 3) array -> contiguous (N, 3) float64 (xyz interleaved)."""
    a = np.asarray(a, dtype=np.float64)
    if a.ndim == 2 and a.shape[0] == 3 and a.shape[1] != 3:
        a = a.T
    return np.ascontiguousarray(a)


def default_threads():
    return max(1, min(16, os.cpu_count() or 1))


class Oracle:
    """CPU FP64 IFGF operator with the reference's construct/init/mult semantics."""

    def __init__(self, kind, k, leaf, order, n_elem=1, tol=-1.0, dx=-1, variant=""):
        k = complex(k)
        self.kind = kind
        self._L = lib(variant)
        self.h = self._L.orc_create(kind, k.real, k.imag, int(leaf), int(order), int(n_elem),
                                  float(tol), int(dx))

    def __del__(self):
        try:
            if getattr(self, "h", None):
                self._L.orc_destroy(self.h)
        except Exception:
            pass
        self.h = None

    def set_cutoff(self, dmax):
        """Drop admissible pairs with box distance > dmax at the next init (<= 0: keep all)."""
        self._L.orc_set_cutoff(self.h, float(dmax))
        return self

    def init(self, srcs, tgts, normals=None, nthreads=None):
        s = _pts(srcs)
        t = _pts(tgts)
        self._srcs, self._tgts = s, t
        n = None if normals is None else _pts(normals)
        self._normals = n
        rc = self._L.orc_init(self.h, _p(s), s.shape[0], _p(t), t.shape[0],
                            None if n is None else _p(n), nthreads or default_threads())
        if rc != 0:
            raise RuntimeError(self._L.orc_last_error(self.h).decode())
        return self

    def mult(self, w, nthreads=None):
        w = np.ascontiguousarray(np.asarray(w, dtype=np.complex128))
        out = np.zeros(self._tgts.shape[0], dtype=np.complex128)
        rc = self._L.orc_mult(self.h, _p(w), _p(out), nthreads or default_threads())
        if rc != 0:
            raise RuntimeError(self._L.orc_last_error(self.h).decode())
        return out

    def mult_sample(self, w, stride, nthreads=None):
        """Recursive phase on every stride-th first-recursive-level subtree; returns the
        fraction of source points covered and the summed duration of the subtrees, each of
        which runs on one thread (thread-seconds; for a bounded CPU timing)."""
        w = np.ascontiguousarray(np.asarray(w, dtype=np.complex128))
        ts = np.zeros(1)
        frac = self._L.orc_mult_sample(self.h, _p(w), int(stride), nthreads or default_threads(),
                                       _p(ts))
        return frac, float(ts[0])

    def base_order(self):
        o = np.zeros(3, dtype=np.int32)
        self._L.orc_base_order(self.h, _p(o))
        return o

    def levels(self, tree=0):
        return self._L.orc_levels(self.h, tree)

    def num_boxes(self, tree, level):
        return self._L.orc_num_boxes(self.h, tree, level)

    def perm(self, tree=0):
        n = (self._srcs if tree == 0 else self._tgts).shape[0]
        p = np.zeros(n, dtype=np.int64)
        self._L.orc_perm(self.h, tree, _p(p))
        return p

    def root_box(self, tree=0):
        b = np.zeros(6)
        self._L.orc_root_box(self.h, tree, _p(b))
        return b

    def tree_level(self, tree, level):
        n = self.num_boxes(tree, level)
        begin = np.zeros(n, np.int64)
        end = np.zeros(n, np.int64)
        parent = np.zeros(n, np.int32)
        leaf = np.zeros(n, np.int8)
        bbox = np.zeros((n, 6))
        self._L.orc_tree_level(self.h, tree, level, _p(begin), _p(end), _p(parent), _p(leaf),
                             _p(bbox))
        return dict(begin=begin, end=end, parent=parent, leaf=leaf.astype(bool), bbox=bbox)

    def interactions(self, level, far):
        n = self._L.orc_num_interactions(self.h, level, int(far))
        rows = np.zeros((n, 3), np.int64)
        if n:
            self._L.orc_interactions(self.h, level, int(far), _p(rows))
        return rows

    def cones(self, level, step):
        nb = self.num_boxes(0, level)
        nel = np.zeros((nb, 3), np.int64)
        nact = np.zeros(nb, np.int64)
        self._L.orc_cone_counts(self.h, level, step, _p(nel), _p(nact))
        ids = np.zeros(int(nact.sum()), np.int64)
        if ids.size:
            self._L.orc_cone_active(self.h, level, step, _p(ids))
        return nel, nact, ids

    def estimate_sample(self, level, box, order):
        """One sampling step of the order estimate on source box `box` of `level` at the
        Chebyshev orders `order`: (nodes (n, 3) Cartesian, samples (n,), coefficients (n,),
        info dict with H, centre, smin and the grid's elements)."""
        o = np.asarray(order, dtype=np.int32)
        n = int(np.prod(o))
        pts = np.zeros((n, 3))
        data = np.zeros(n, np.complex128)
        trafo = np.zeros(n, np.complex128)
        info = np.zeros(8)
        self._L.orc_estimate_sample(self.h, int(level), int(box), _p(o), _p(pts), _p(data),
                                  _p(trafo), _p(info))
        return pts, data, trafo, dict(H=info[0], centre=info[1:4].copy(), smin=info[4],
                                      nels=info[5:8].astype(int))

    def min_rec(self):
        return self._L.orc_min_rec(self.h)

    def kernel(self, xs, y, src_idx=0):
        out = np.zeros(2)
        self._L.orc_kernel_eval(self.h, _p(np.ascontiguousarray(xs, np.float64)),
                              _p(np.ascontiguousarray(y, np.float64)), src_idx, _p(out))
        return complex(out[0], out[1])

    def direct_sum(self, srcs, tgts, w, idx, normals=None, nthreads=None):
        """FP64 direct sum of this kernel at tgts[idx] over srcs (no init needed; self term 0)."""
        s = _pts(srcs)
        t = _pts(tgts)
        n = None if normals is None else _pts(normals)
        w = np.ascontiguousarray(np.asarray(w, dtype=np.complex128))
        idx = np.ascontiguousarray(np.asarray(idx, dtype=np.int64))
        out = np.zeros(idx.size, np.complex128)
        self._L.orc_direct(self.h, _p(s), s.shape[0], _p(w), None if n is None else _p(n), _p(t),
                         _p(idx), idx.size, _p(out), nthreads or default_threads())
        return out

    def direct(self, w, idx, nthreads=None):
        """FP64 direct sum at targets[idx] over all sources (self term 0)."""
        w = np.ascontiguousarray(np.asarray(w, dtype=np.complex128))
        idx = np.ascontiguousarray(np.asarray(idx, dtype=np.int64))
        out = np.zeros(idx.size, np.complex128)
        n = self._normals
        self._L.orc_direct(self.h, _p(self._srcs), self._srcs.shape[0], _p(w),
                         None if n is None else _p(n), _p(self._tgts), _p(idx), idx.size,
                         _p(out), nthreads or default_threads())
        return out


def slice_norm(values, ns, axis, layers):
    """util.hpp:214-234: relative norm of the coefficient slice index ns[axis] - layers."""
    ns_a = np.asarray(ns, dtype=np.int32)
    v = np.ascontiguousarray(values, dtype=np.complex128)
    return lib().orc_slice_norm(_p(v), v.size, _p(ns_a), int(axis), int(layers))


def chebnodes1d(n):
    out = np.zeros(n)
    lib().orc_chebnodes1d(n, _p(out))
    return out


def chebtransform(ns, values):
    ns_a = np.asarray(ns, dtype=np.int32)
    v = np.ascontiguousarray(values, dtype=np.complex128)
    out = np.zeros_like(v)
    lib().orc_chebtransform(_p(ns_a), _p(v), _p(out))
    return out


def clenshaw(ns, coef, pts):
    ns_a = np.asarray(ns, dtype=np.int32)
    c = np.ascontiguousarray(coef, dtype=np.complex128)
    p = _pts(pts)
    out = np.zeros(p.shape[0], np.complex128)
    lib().orc_clenshaw(_p(ns_a), _p(c), _p(p), p.shape[0], _p(out))
    return out


def morton_sort(pts):
    p = _pts(pts)
    perm = np.zeros(p.shape[0], np.int64)
    box = np.zeros(6)
    lib().orc_morton_sort(_p(p), p.shape[0], _p(perm), _p(box))
    return perm, box


def morton_less(box, p, q):
    return bool(lib().orc_morton_less(_p(np.ascontiguousarray(box, np.float64)),
                                      _p(np.ascontiguousarray(p, np.float64)),
                                      _p(np.ascontiguousarray(q, np.float64))))


def cart_to_interp(x, xc, H):
    out = np.zeros(3)
    lib().orc_cart_to_interp(_p(np.ascontiguousarray(x, np.float64)),
                             _p(np.ascontiguousarray(xc, np.float64)), float(H), _p(out))
    return out


def interp_to_cart(p, xc, H):
    out = np.zeros(3)
    lib().orc_interp_to_cart(_p(np.ascontiguousarray(p, np.float64)),
                             _p(np.ascontiguousarray(xc, np.float64)), float(H), _p(out))
    return out
