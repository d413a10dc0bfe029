"""Python mirror of TimHRO/ifgf's `pyifgf` module (py_ifgf.cpp:16-58) over the C ABI.

Same class names, constructor arguments, and init / mult semantics as the reference's
pybind11 bindings, so code written against `pyifgf` runs unchanged:

    op = MofifiedHelmholtzIfgfOperator(k, leafSize, order, n_elem, tol)
    op.init(srcs, targets)          # float64 arrays of shape (3, N) and (3, M)
    u = op.mult(weights)            # complex128 (N,) -> complex128 (M,)

The reference exports MofifiedHelmholtzIfgfOperator (sic), GradHelmholtzIfgfOperator
(+ setDx) and LaplaceIfgfOperator; the Helmholtz, double-layer and combined-field classes of
its C++ API are exported here as well.  Extra keyword-only arguments: `precision`
("mixed": FP32 arithmetic with FP64 cone coordinates and phases, the default; "fp64": every
step in FP64).  mult also accepts a 2-D array (N, nrhs) and torch tensors on the GPU
(zero-copy, enqueued on torch's current stream).
"""
import ctypes
import os
import sys

import numpy as np

from . import _lib
from ._lib import check, lib

__version__ = "dev"


def _points(a, name):
    """(3, N) like the reference's Eigen::Array<double,3,Dynamic>; (N, 3) is accepted too."""
    a = np.asarray(a, dtype=np.float64)
    if a.ndim != 2 or 3 not in a.shape:
        raise ValueError(f"{name} must have shape (3, N)")
    pts = a.T if a.shape[0] == 3 else a
    return np.ascontiguousarray(pts)


def _prec(p):
    if p in ("mixed", "fp32", _lib.PREC_MIXED):
        return _lib.PREC_MIXED
    if p in ("fp64", "double", _lib.PREC_FP64):
        return _lib.PREC_FP64
    raise ValueError("precision must be 'mixed' or 'fp64'")


class _IfgfOperator:
    """IfgfOperator<T,3,1,Derived> (ifgfoperator.hpp:25-1936) behind the C ABI."""

    _kernel = _lib.HELMHOLTZ
    _real = False  # Laplace: float64 densities and results

    def __init__(self, k, leafSize, order, n_elem, tol, *, precision="mixed"):
        k = complex(k)
        h = ctypes.c_void_p()
        rc = lib().ifgf_create(self._kernel, k.real, k.imag, int(leafSize), int(order),
                               int(n_elem), float(tol), _prec(precision), ctypes.byref(h))
        if rc != _lib.IFGF_OK:
            raise _lib.IfgfError(f"invalid operator arguments ({rc})")
        self._h = h
        self._n = None
        self._m = None
        self._keep = None
        # mults of device tensors replay the mult's kernels as a HIP graph (one captured graph
        # for single columns, one for batches of several)
        self.use_graph = False

    def __del__(self, _finalizing=sys.is_finalizing):
        h = getattr(self, "_h", None)
        if not h:
            return
        self._h = None
        if _finalizing():
            return  # interpreter exit: the process's device memory goes with it
        lib().ifgf_destroy(h)

    # IfgfOperator::init (ifgfoperator.hpp:65-106)
    def init(self, srcs, targets, cutOff=None, normals=None, *, plan_only=False,
             host_plan=False, level_ring=False):
        s = _points(srcs, "srcs")
        t = _points(targets, "targets")
        n = None if normals is None else _points(normals, "normals")
        cb = _lib.CUTOFF_FN(0) if cutOff is None else _lib.CUTOFF_FN(
            lambda d, ctx: 1 if cutOff(d) else 0)
        self._keep = cb
        rc = lib().ifgf_init(self._h, s.ctypes.data, s.shape[0], t.ctypes.data, t.shape[0],
                             None if n is None else n.ctypes.data, cb, None,
                             (_lib.INIT_PLAN_ONLY if plan_only else 0)
                             | (_lib.INIT_HOST_PLAN if host_plan else 0)
                             | (_lib.INIT_LEVEL_RING if level_ring else 0))
        check(rc, self._h, "init")
        self._n, self._m = s.shape[0], t.shape[0]

    def save_plan(self, path):
        """Write this operator's plan (octrees, interaction lists, active cones, upward
        tables) and launch shape to a file (not in the reference API, whose init rebuilds
        them every time)."""
        check(lib().ifgf_plan_save(self._h, os.fsencode(path)), self._h, "save_plan")

    def load_plan(self, path, *, plan_only=False, level_ring=False):
        """init from a file written by save_plan, for an operator of the same kernel,
        wavenumber, leaf size, order, n_elem and tolerance: the same results, bit for bit,
        without rebuilding the plan."""
        rc = lib().ifgf_init_from_plan(self._h, os.fsencode(path),
                                       (_lib.INIT_PLAN_ONLY if plan_only else 0)
                                       | (_lib.INIT_LEVEL_RING if level_ring else 0))
        check(rc, self._h, "load_plan")
        n_src, n_tgt = ctypes.c_int64(), ctypes.c_int64()
        check(lib().ifgf_sizes(self._h, ctypes.byref(n_src), ctypes.byref(n_tgt)), self._h,
              "sizes")
        self._n, self._m = n_src.value, n_tgt.value
        self._keep = None

    # IfgfOperator::mult (ifgfoperator.hpp:246-867)
    def mult(self, weights):
        return self._apply("ifgf_mult", weights, self._n, self._m)

    def mult_adjoint(self, values):
        """A^H g: the adjoint product, M values -> N (not in the reference API).  For the
        potentials and the gradient kernel; targets other than the sources cost a transposed
        plan on the first call."""
        return self._apply("ifgf_mult_adjoint", values, self._m, self._n)

    def _apply(self, fname, weights, n_in, n_out):
        if self._n is None:
            raise _lib.IfgfError("mult before init")
        try:
            import torch
            if isinstance(weights, torch.Tensor) and weights.is_cuda:
                return self._apply_torch(fname, weights, n_in, n_out)
        except ImportError:
            pass
        w = np.asarray(weights)
        squeeze = w.ndim == 1
        w2 = w.reshape(n_in, -1) if not squeeze else w.reshape(n_in, 1)
        nrhs = w2.shape[1]
        wc = np.ascontiguousarray(w2.T.astype(np.complex128))  # nrhs vectors back to back
        out = np.empty((nrhs, n_out), dtype=np.complex128)
        check(getattr(lib(), fname)(self._h, wc.ctypes.data, nrhs, out.ctypes.data, 0, None),
              self._h, fname[5:])
        res = out.T
        if self._real:
            res = res.real.copy()
        return res[:, 0].copy() if squeeze else np.ascontiguousarray(res)

    def _apply_torch(self, fname, w, n_in, n_out):
        import torch
        squeeze = w.dim() == 1
        w2 = w.reshape(n_in, 1) if squeeze else w.reshape(n_in, -1)
        nrhs = w2.shape[1]
        # conj() and neg() of complex tensors are lazy views: materialise them before the
        # raw pointer goes to the library
        wc = w2.t().to(torch.complex128).resolve_conj().resolve_neg().contiguous()
        out = torch.empty((nrhs, n_out), dtype=torch.complex128, device=w.device)
        stream = torch.cuda.current_stream(w.device).cuda_stream
        flags = _lib.DEVICE_POINTERS
        if getattr(self, "use_graph", False) and fname == "ifgf_mult":
            flags |= _lib.MULT_GRAPH
        check(getattr(lib(), fname)(self._h, wc.data_ptr(), nrhs, out.data_ptr(), flags, stream),
              self._h, fname[5:])
        res = out.t()
        if self._real:
            res = res.real
        return res[:, 0].contiguous() if squeeze else res.contiguous()

    # --- introspection (not in the reference API) ----------------------------------------
    def plan_counts(self):
        c = np.zeros(16, np.int64)
        check(lib().ifgf_plan_counts(self._h, c.ctypes.data), self._h, "plan_counts")
        keys = ["src_levels", "src_boxes", "tgt_levels", "tgt_boxes", "far_pairs", "near_pairs",
                "coarse_cones", "fine_cones", "fine_uncovered", "dropped", "device_bytes",
                "fine_ids", "up_table_levels", "up_table_entries", "level_ring",
                "transform_cones"]
        return {k: int(v) for k, v in zip(keys, c)}

    def plan_orders(self):
        o = np.zeros(12, np.int32)
        check(lib().ifgf_plan_orders(self._h, o.ctypes.data), self._h, "plan_orders")
        return dict(base=o[0:3].copy(), high=o[3:6].copy(), low=o[6:9].copy(),
                    factor=o[9:12].copy())

    def plan_tree(self, tree=0):
        c = self.plan_counts()
        lv = c["src_levels" if tree == 0 else "tgt_levels"]
        nb = c["src_boxes" if tree == 0 else "tgt_boxes"]
        n = self._n if tree == 0 else self._m
        level_off = np.zeros(lv + 1, np.int64)
        begin = np.zeros(nb, np.int64)
        end = np.zeros(nb, np.int64)
        parent = np.zeros(nb, np.int32)
        leaf = np.zeros(nb, np.int8)
        bbox = np.zeros((nb, 6))
        perm = np.zeros(n, np.int64)
        check(lib().ifgf_plan_tree(self._h, tree, level_off.ctypes.data, begin.ctypes.data,
                                   end.ctypes.data, parent.ctypes.data, leaf.ctypes.data,
                                   bbox.ctypes.data, perm.ctypes.data), self._h, "plan_tree")
        return dict(level_off=level_off, begin=begin, end=end, parent=parent,
                    leaf=leaf.astype(bool), bbox=bbox, perm=perm)

    def plan_pairs(self, far):
        c = self.plan_counts()
        n = c["far_pairs" if far else "near_pairs"]
        s = np.zeros(n, np.int32)
        t = np.zeros(n, np.int32)
        if n:
            check(lib().ifgf_plan_pairs(self._h, int(bool(far)), s.ctypes.data, t.ctypes.data),
                  self._h, "plan_pairs")
        return s, t

    def plan_cones(self):
        c = self.plan_counts()
        nb = c["src_boxes"]
        nc = np.zeros((nb, 3), np.int64)
        nf = np.zeros((nb, 3), np.int64)
        cone_off = np.zeros(nb + 1, np.int64)
        cell = np.zeros(c["coarse_cones"], np.int64)
        mask = np.zeros(c["coarse_cones"], np.uint32)
        fine_off = np.zeros(nb + 1, np.int64)
        fine_ids = np.zeros(c["fine_ids"], np.int64)
        check(lib().ifgf_plan_cones(self._h, nc.ctypes.data, nf.ctypes.data, cone_off.ctypes.data,
                                    cell.ctypes.data, mask.ctypes.data, fine_off.ctypes.data,
                                    fine_ids.ctypes.data), self._h, "plan_cones")
        return dict(nc=nc, nf=nf, cone_off=cone_off, cell=cell, mask=mask, fine_off=fine_off,
                    fine_ids=fine_ids)

    def plan_work(self):
        w = np.zeros(10)
        check(lib().ifgf_plan_work(self._h, w.ctypes.data), self._h, "plan_work")
        keys = ["near_pairs", "leaf_pairs", "up_evals", "far_pairs", "coarse_cones", "fine_cones",
                "leaf_cones", "n_hi", "n_lo", "transform_flops_per_cone"]
        return dict(zip(keys, w.tolist()))

    def last_timings(self):
        t = np.zeros(6)
        check(lib().ifgf_last_timings(self._h, t.ctypes.data), self._h, "last_timings")
        return dict(zip(["near", "leaf", "upward", "transform", "far", "total"], t))

    _SHAPE_KEYS = ("xcd_run_log2", "transform_cones", "up_cones", "batch_columns", "far_waves",
                   "overlap")

    def autotune(self, nrhs=1, reps=3):
        """Time the launch-shape candidates on this operator (synthetic densities, nrhs
        columns) and keep the fastest; returns the chosen shape."""
        s = np.zeros(len(self._SHAPE_KEYS), np.int32)
        check(lib().ifgf_autotune(self._h, int(nrhs), int(reps), s.ctypes.data), self._h,
              "autotune")
        return dict(zip(self._SHAPE_KEYS, (int(v) for v in s)))

    @property
    def launch_shape(self):
        s = np.zeros(len(self._SHAPE_KEYS), np.int32)
        check(lib().ifgf_get_launch_shape(self._h, s.ctypes.data), self._h, "launch_shape")
        return dict(zip(self._SHAPE_KEYS, (int(v) for v in s)))

    @launch_shape.setter
    def launch_shape(self, shape):
        cur = self.launch_shape
        cur.update(shape)
        s = np.array([cur[k] for k in self._SHAPE_KEYS], np.int32)
        check(lib().ifgf_set_launch_shape(self._h, s.ctypes.data), self._h, "launch_shape")


class MergedIfgfOperator(_IfgfOperator):
    """Several initialised operators of one class and one set of orders as one (ifgf_merge):
    every phase of mult runs one launch per level for all geometries.  mult takes the parts'
    densities concatenated (or a list of them, returning a list of results)."""

    def __init__(self, ops, *, plan_only=False):
        ops = list(ops)
        if not ops or any(getattr(o, "_n", None) is None for o in ops):
            raise _lib.IfgfError("merge needs initialised operators")
        arr = (ctypes.c_void_p * len(ops))(*[o._h.value for o in ops])
        h = ctypes.c_void_p()
        rc = lib().ifgf_merge(arr, len(ops), _lib.INIT_PLAN_ONLY if plan_only else 0,
                              ctypes.byref(h))
        check(rc, ops[0]._h, "merge")
        self._h = h
        self._real = ops[0]._real
        self._kernel = ops[0]._kernel
        n, m = ctypes.c_int64(), ctypes.c_int64()
        check(lib().ifgf_sizes(h, ctypes.byref(n), ctypes.byref(m)), h, "sizes")
        self._n, self._m = n.value, m.value
        self._keep = None
        self.parts = [(o._n, o._m) for o in ops]

    def mult(self, weights):
        if isinstance(weights, (list, tuple)):
            u = super().mult(np.concatenate([np.asarray(w) for w in weights]))
            out, at = [], 0
            for _, m in self.parts:
                out.append(u[at:at + m])
                at += m
            return out
        return super().mult(weights)


def merge(ops, *, plan_only=False):
    """One operator for several initialised operators of one class and one set of orders."""
    return MergedIfgfOperator(ops, plan_only=plan_only)


class HelmholtzIfgfOperator(_IfgfOperator):
    """helmholtz_ifgf.hpp: G = e^{i kappa d} / (4 pi d), real kappa."""
    _kernel = _lib.HELMHOLTZ

    def __init__(self, waveNumber, leafSize, order, n_elem=1, tol=-1, **kw):
        super().__init__(float(np.real(waveNumber)), leafSize, order, n_elem, tol, **kw)


class MofifiedHelmholtzIfgfOperator(_IfgfOperator):
    """modified_helmholtz_ifgf.hpp: G = e^{-k d} / (4 pi d), complex k (the reference's
    Python class name, py_ifgf.cpp:40-41, spelling included)."""
    _kernel = _lib.MODHELMHOLTZ


ModifiedHelmholtzIfgfOperator = MofifiedHelmholtzIfgfOperator


class GradHelmholtzIfgfOperator(_IfgfOperator):
    """grad_helmholtz_ifgf.hpp: target derivative dG/dy_dx selected by setDx."""
    _kernel = _lib.GRAD

    def setDx(self, dx):
        check(lib().ifgf_set_dx(self._h, int(dx)), self._h, "setDx")


class LaplaceIfgfOperator(_IfgfOperator):
    """laplace_ifgf.hpp: G = 1 / (4 pi d), float64 densities (py_ifgf.cpp:45-48)."""
    _kernel = _lib.LAPLACE
    _real = True

    def __init__(self, leafSize, order, n_elem, tol, **kw):
        super().__init__(0.0, leafSize, order, n_elem, tol, **kw)


class _NormalsOperator(_IfgfOperator):
    def init(self, srcs, targets, normals=None, cutOff=None, **kw):
        if normals is None:
            raise ValueError("this operator needs source normals: init(srcs, targets, normals)")
        if callable(normals) and not hasattr(normals, "shape"):
            raise ValueError("init(srcs, targets, normals): normals is an array")
        return super().init(srcs, targets, cutOff=cutOff, normals=normals, **kw)


class DoubleLayerHelmholtzIfgfOperator(_NormalsOperator):
    """double_layer_helmholtz_ifgf.hpp: -(1/4pi) e^{-kd}(-k - 1/d)((x-y).n_x)/d^2."""
    _kernel = _lib.DOUBLELAYER


class CombinedFieldHelmholtzIfgfOperator(_NormalsOperator):
    """combined_field_helmholtz_ifgf.hpp: DL - i kappa SL with real kappa."""
    _kernel = _lib.COMBINED

    def __init__(self, waveNumber, leafSize, order, n_elem=1, tol=-1, **kw):
        super().__init__(float(np.real(waveNumber)), leafSize, order, n_elem, tol, **kw)
