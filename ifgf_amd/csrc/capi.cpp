// capi.cpp -- the extern "C" boundary (include/ifgf_amd.h) over the plan and the device
// pipeline.  Host C++ only; the HIP code lives in device.hip.
#include <cstring>
#include <exception>
#include <functional>
#include <memory>
#include <string>
#include <vector>

#include "../../include/ifgf_amd.h"
#include "device.hpp"
#include "plan.hpp"
#include "plan_io.hpp"

using namespace ifgf;

struct ifgf_operator
{
    KernelSpec K;
    int64_t leaf = 64;
    int order = 6;
    int n_elem = 1;
    int precision = PREC_MIXED;
    bool planned = false;
    bool on_device = false;
    Plan plan;
    DevicePlan dev;
    std::string err;
    // adjoint products: same_points (targets are the sources, in order) reuses this plan;
    // otherwise a transposed operator (sources and targets swapped) is built on first use
    bool same_points = false;
    bool merged = false; // several geometries (ifgf_merge)
    bool host_plan = false; // octrees and cone marking on the host (IFGF_INIT_HOST_PLAN)
    ifgf_cutoff_fn cutoff = nullptr;
    void* cutoff_ctx = nullptr;
    // the plan was read from a file and built with a cutOff callback that did not travel:
    // a transposed plan (adjoint for other targets) cannot be rebuilt the same way
    bool cutoff_lost = false;
    std::unique_ptr<ifgf_operator> adj;
};

static int fail(ifgf_operator* op, int code, const std::string& msg)
{
    if (op)
        op->err = msg;
    return code;
}

extern "C"
{
    const char* ifgf_version(void) { return "ifgf_amd 0.1 (gfx950)"; }

    int ifgf_device_count(void)
    {
        return device_available() ? 1 : 0;
    }

    int ifgf_create(int kernel, double k_re, double k_im, int64_t leaf_size, int order, int n_elem,
                    double tol, int precision, ifgf_operator** out)
    {
        if (!out)
            return IFGF_EINVAL;
        *out = nullptr;
        if (kernel < IFGF_HELMHOLTZ || kernel > IFGF_GRAD || order < 1 || n_elem < 1 ||
            (precision != IFGF_PREC_FP64 && precision != IFGF_PREC_MIXED))
            return IFGF_EINVAL;
        auto* op = new (std::nothrow) ifgf_operator();
        if (!op)
            return IFGF_EINTERNAL;
        op->K.kind = kernel;
        op->K.tol = tol;
        op->K.kappa = k_re;
        if (kernel == IFGF_HELMHOLTZ || kernel == IFGF_COMBINED)
            op->K.k = cplx(0.0, -k_re); // e^{-k d} = e^{i kappa d}
        else if (kernel == IFGF_LAPLACE)
            op->K.k = cplx(0.0, 0.0);
        else
            op->K.k = cplx(k_re, k_im);
        op->leaf = leaf_size;
        op->order = order;
        op->n_elem = n_elem;
        op->precision = precision;
        *out = op;
        return IFGF_OK;
    }

    int ifgf_set_dx(ifgf_operator* op, int dx)
    {
        if (!op)
            return IFGF_EINVAL;
        if (op->K.kind != IFGF_GRAD || dx < -1 || dx > 2)
            return fail(op, IFGF_EINVAL, "setDx applies to the gradient kernel, dx in -1..2");
        op->K.dx = dx;
        op->plan.K.dx = dx;
        if (op->dev.dx != dx)
            device_drop_graph(op->dev); // the captured kernels carry the old direction
        op->dev.dx = dx;
        return IFGF_OK;
    }

    int ifgf_init(ifgf_operator* op, const double* srcs, int64_t n_src, const double* tgts,
                  int64_t n_tgt, const double* normals, ifgf_cutoff_fn cutoff, void* cutoff_ctx,
                  int flags)
    {
        if (!op)
            return IFGF_EINVAL;
        if (!srcs || !tgts || n_src <= 0 || n_tgt <= 0)
            return fail(op, IFGF_EINVAL, "init needs non-empty source and target arrays");
        if ((op->K.kind == IFGF_DOUBLELAYER || op->K.kind == IFGF_COMBINED) && !normals)
            return fail(op, IFGF_EINVAL, "this kernel needs source normals");
        const bool plan_only = (flags & IFGF_INIT_PLAN_ONLY) != 0;
        if (!plan_only && !device_available())
            return fail(op, IFGF_ENODEV, "no HIP device visible: ifgf_amd has no CPU path");
        try
        {
            device_free(op->dev);
            op->adj.reset();
            op->on_device = false;
            op->planned = false;
            op->same_points =
                n_src == n_tgt && std::memcmp(srcs, tgts, sizeof(double) * 3 * n_src) == 0;
            op->cutoff = cutoff;
            op->cutoff_ctx = cutoff_ctx;
            op->cutoff_lost = false;
            op->plan = Plan();
            op->plan.K = op->K;
            std::function<bool(double)> cb;
            if (cutoff)
                cb = [cutoff, cutoff_ctx](double d) { return cutoff(d, cutoff_ctx) != 0; };
            op->host_plan = plan_only || (flags & IFGF_INIT_HOST_PLAN) != 0;
            std::string e =
                build_plan(op->plan, srcs, n_src, tgts, n_tgt, normals, op->leaf, op->order,
                           op->n_elem, cb, default_threads(),
                           op->host_plan ? PlanBuilders() : device_plan_builders());
            if (!e.empty())
                return fail(op, IFGF_EINVAL, e);
            op->planned = true;
            if (plan_only)
                return IFGF_OK;
            e = device_upload(op->plan, op->precision, op->dev,
                              (flags & IFGF_INIT_LEVEL_RING) ? 1 : 0);
            if (!e.empty())
                return fail(op, IFGF_EINTERNAL, e);
            op->on_device = true;
        }
        catch (const std::exception& ex)
        {
            return fail(op, IFGF_EINTERNAL, ex.what());
        }
        return IFGF_OK;
    }

    int ifgf_plan_save(ifgf_operator* op, const char* path)
    {
        if (!op || !path)
            return IFGF_EINVAL;
        if (!op->planned)
            return fail(op, IFGF_ESTATE, "plan_save before init");
        try
        {
            PlanFileInfo info;
            info.leaf = op->leaf;
            info.order = op->order;
            info.n_elem = op->n_elem;
            const LaunchShape& s = op->dev.shape;
            const int sh[6] = {s.xcd_log2, s.cpb, s.up_k, s.rb, s.far_waves, s.overlap};
            std::memcpy(info.shape, sh, sizeof(sh));
            info.same_points = op->same_points;
            info.merged = op->merged;
            info.had_cutoff = op->cutoff != nullptr || op->cutoff_lost;
            info.host_plan = op->host_plan;
            std::string e = plan_save(op->plan, info, path);
            if (!e.empty())
                return fail(op, IFGF_EINVAL, e);
        }
        catch (const std::exception& ex)
        {
            return fail(op, IFGF_EINTERNAL, ex.what());
        }
        return IFGF_OK;
    }

    int ifgf_init_from_plan(ifgf_operator* op, const char* path, int flags)
    {
        if (!op || !path)
            return IFGF_EINVAL;
        const bool plan_only = (flags & IFGF_INIT_PLAN_ONLY) != 0;
        if (!plan_only && !device_available())
            return fail(op, IFGF_ENODEV, "no HIP device visible: ifgf_amd has no CPU path");
        try
        {
            Plan P;
            PlanFileInfo info;
            std::string e = plan_load(P, info, path);
            if (e.empty())
            {
                e = plan_check(P);
                if (!e.empty())
                    e = std::string(path) + " does not hold a consistent plan (" + e + ")";
            }
            if (!e.empty())
                return fail(op, IFGF_EINVAL, e);
            const KernelSpec& K = P.K;
            if (K.kind != op->K.kind || K.k != op->K.k || K.kappa != op->K.kappa ||
                K.tol != op->K.tol || info.leaf != op->leaf || info.order != op->order ||
                info.n_elem != op->n_elem)
                return fail(op, IFGF_EINVAL,
                            std::string(path) +
                                " holds the plan of an operator with another kernel, "
                                "wavenumber, leaf size, order, n_elem or tolerance");
            device_free(op->dev);
            op->adj.reset();
            op->on_device = false;
            op->planned = false;
            op->plan = std::move(P);
            op->plan.K.dx = op->K.dx; // the direction is the operator's, not the plan's
            op->same_points = info.same_points != 0;
            op->merged = info.merged != 0;
            op->host_plan = info.host_plan != 0;
            op->cutoff = nullptr;
            op->cutoff_ctx = nullptr;
            op->cutoff_lost = info.had_cutoff != 0;
            op->planned = true;
            if (plan_only)
                return IFGF_OK;
            e = device_upload(op->plan, op->precision, op->dev,
                              (flags & IFGF_INIT_LEVEL_RING) ? 1 : 0);
            if (!e.empty())
                return fail(op, IFGF_EINTERNAL, e);
            op->on_device = true;
            LaunchShape s;
            s.xcd_log2 = info.shape[0];
            s.cpb = info.shape[1];
            s.up_k = info.shape[2];
            s.rb = info.shape[3];
            s.far_waves = info.shape[4];
            s.overlap = info.shape[5];
            e = device_set_shape(op->dev, s);
            if (!e.empty())
                return fail(op, IFGF_EINVAL, std::string(path) + ": launch shape: " + e);
        }
        catch (const std::exception& ex)
        {
            return fail(op, IFGF_EINTERNAL, ex.what());
        }
        return IFGF_OK;
    }

    int ifgf_mult(ifgf_operator* op, const double* w, int64_t nrhs, double* out, int flags,
                  void* stream)
    {
        if (!op)
            return IFGF_EINVAL;
        if (!op->on_device)
            return fail(op, IFGF_ESTATE, "mult before a device init");
        if (!w || !out || nrhs < 1)
            return fail(op, IFGF_EINVAL, "mult needs weights, output and nrhs >= 1");
        std::string e = device_mult(op->dev, w, out, nrhs, (flags & IFGF_DEVICE_POINTERS) != 0,
                                    stream, 0, (flags & IFGF_MULT_GRAPH) != 0);
        if (!e.empty())
            return fail(op, IFGF_EINTERNAL, e);
        return IFGF_OK;
    }

    int ifgf_mult_noalloc(ifgf_operator* op, const double* w_dev, double* out_dev, void* stream)
    {
        if (!op)
            return IFGF_EINVAL;
        if (!op->on_device)
            return fail(op, IFGF_ESTATE, "mult before a device init");
        if (!w_dev || !out_dev)
            return fail(op, IFGF_EINVAL, "mult needs device densities and a device output");
        std::string e = device_mult_noalloc(op->dev, w_dev, out_dev, stream);
        if (!e.empty())
            return fail(op, e.rfind("allocation-free mult: ", 0) == 0 ? IFGF_EINVAL
                                                                      : IFGF_EINTERNAL,
                        e);
        return IFGF_OK;
    }

    int ifgf_mult_adjoint(ifgf_operator* op, const double* g, int64_t nrhs, double* out,
                          int flags, void* stream)
    {
        if (!op)
            return IFGF_EINVAL;
        if (!op->on_device)
            return fail(op, IFGF_ESTATE, "mult_adjoint before a device init");
        if (!g || !out || nrhs < 1)
            return fail(op, IFGF_EINVAL, "mult_adjoint needs values, output and nrhs >= 1");
        const int kind = op->K.kind;
        if (kind == IFGF_DOUBLELAYER || kind == IFGF_COMBINED)
            return fail(op, IFGF_EINVAL,
                        "no adjoint for kernels with source normals (it needs target normals)");
        // G(x - y) = G(y - x) for the potentials, and the gradient kernel is odd in x - y:
        // A^H g = s conj(B conj(g)) with B the operator from the targets to the sources
        const int io = IO_CONJ_IN | IO_CONJ_OUT | (kind == IFGF_GRAD ? IO_NEGATE_OUT : 0);
        try
        {
            ifgf_operator* B = op;
            if (!op->same_points && op->merged)
                return fail(op, IFGF_EINVAL,
                            "the adjoint of merged geometries needs targets = sources in each");
            if (!op->same_points && op->cutoff_lost)
                return fail(op, IFGF_EINVAL,
                            "the adjoint for other targets rebuilds the plan with the cutOff "
                            "callback, which a plan file does not hold: init this operator");
            if (!op->same_points)
            {
                if (!op->adj)
                {
                    auto a = std::make_unique<ifgf_operator>();
                    a->K = op->K;
                    a->leaf = op->leaf;
                    a->order = op->order;
                    a->n_elem = op->n_elem;
                    a->precision = op->precision;
                    a->plan.K = op->K;
                    // the caller's point order, from the sorted points and their permutation
                    const Tree &S = op->plan.S, &T = op->plan.T;
                    std::vector<double> xs(3 * S.n), ys(3 * T.n);
                    for (int64_t i = 0; i < S.n; i++)
                        for (int d = 0; d < 3; d++)
                            xs[3 * S.perm[i] + d] = S.pts[3 * i + d];
                    for (int64_t i = 0; i < T.n; i++)
                        for (int d = 0; d < 3; d++)
                            ys[3 * T.perm[i] + d] = T.pts[3 * i + d];
                    std::function<bool(double)> cb;
                    if (op->cutoff)
                    {
                        ifgf_cutoff_fn f = op->cutoff;
                        void* ctx = op->cutoff_ctx;
                        cb = [f, ctx](double d) { return f(d, ctx) != 0; };
                    }
                    a->host_plan = op->host_plan;
                    std::string e =
                        build_plan(a->plan, ys.data(), T.n, xs.data(), S.n, nullptr, a->leaf,
                                   a->order, a->n_elem, cb, default_threads(),
                                   a->host_plan ? PlanBuilders() : device_plan_builders());
                    if (!e.empty())
                        return fail(op, IFGF_EINTERNAL, "adjoint plan: " + e);
                    a->planned = true;
                    e = device_upload(a->plan, a->precision, a->dev);
                    if (!e.empty())
                        return fail(op, IFGF_EINTERNAL, "adjoint upload: " + e);
                    a->on_device = true;
                    op->adj = std::move(a);
                }
                B = op->adj.get();
                B->dev.dx = op->dev.dx;
            }
            std::string e = device_mult(B->dev, g, out, nrhs, (flags & IFGF_DEVICE_POINTERS) != 0,
                                        stream, io);
            if (!e.empty())
                return fail(op, IFGF_EINTERNAL, e);
        }
        catch (const std::exception& ex)
        {
            return fail(op, IFGF_EINTERNAL, ex.what());
        }
        return IFGF_OK;
    }

    int ifgf_merge(ifgf_operator* const* ops, int n_ops, int flags, ifgf_operator** out)
    {
        if (!out)
            return IFGF_EINVAL;
        *out = nullptr;
        if (!ops || n_ops < 1)
            return IFGF_EINVAL;
        std::vector<const Plan*> parts;
        bool same = true;
        for (int i = 0; i < n_ops; i++)
        {
            if (!ops[i] || !ops[i]->planned || ops[i]->merged)
                return IFGF_ESTATE;
            parts.push_back(&ops[i]->plan);
            same = same && ops[i]->same_points;
        }
        const bool plan_only = (flags & IFGF_INIT_PLAN_ONLY) != 0;
        if (!plan_only && !device_available())
            return IFGF_ENODEV;
        auto m = std::make_unique<ifgf_operator>();
        const ifgf_operator& o0 = *ops[0];
        m->K = o0.K;
        m->leaf = o0.leaf;
        m->order = o0.order;
        m->n_elem = o0.n_elem;
        m->precision = o0.precision;
        m->merged = true;
        m->same_points = same;
        try
        {
            std::string e = merge_plans(parts, m->plan);
            if (!e.empty())
            {
                ops[0]->err = e;
                return IFGF_EINVAL;
            }
            m->planned = true;
            if (!plan_only)
            {
                e = device_upload(m->plan, m->precision, m->dev);
                if (!e.empty())
                {
                    ops[0]->err = e;
                    return IFGF_EINTERNAL;
                }
                m->on_device = true;
            }
        }
        catch (const std::exception& ex)
        {
            ops[0]->err = ex.what();
            return IFGF_EINTERNAL;
        }
        *out = m.release();
        return IFGF_OK;
    }

    int ifgf_sizes(const ifgf_operator* op, int64_t* n_src, int64_t* n_tgt)
    {
        if (!op || !n_src || !n_tgt)
            return IFGF_EINVAL;
        if (!op->planned)
            return IFGF_ESTATE;
        *n_src = op->plan.S.n;
        *n_tgt = op->plan.T.n;
        return IFGF_OK;
    }

    void ifgf_destroy(ifgf_operator* op)
    {
        if (!op)
            return;
        if (op->adj)
            device_free(op->adj->dev);
        device_free(op->dev);
        delete op;
    }

    const char* ifgf_last_error(const ifgf_operator* op) { return op ? op->err.c_str() : "null operator"; }

    int ifgf_plan_counts(const ifgf_operator* op, int64_t* c)
    {
        if (!op || !c)
            return IFGF_EINVAL;
        if (!op->planned)
            return IFGF_ESTATE;
        const Plan& P = op->plan;
        c[0] = P.S.levels;
        c[1] = P.S.num_boxes();
        c[2] = P.T.levels;
        c[3] = P.T.num_boxes();
        c[4] = int64_t(P.far.size());
        c[5] = int64_t(P.near.size());
        c[6] = P.num_cones();
        c[7] = P.n_fine;
        c[8] = P.n_fine_uncovered;
        c[9] = P.n_tofar;
        c[10] = op->dev.bytes_total;
        c[11] = int64_t(P.fine_ids.size());
        c[12] = 0;
        c[13] = 0;
        for (const UpLevel& U : P.up)
            if (U.valid)
            {
                c[12]++;
                c[13] += int64_t(U.entries.size());
            }
        for (const auto& pp : P.up_parts)
        {
            c[12]++;
            c[13] += int64_t(pp.U.entries.size());
        }
        c[14] = op->dev.ring ? 1 : 0;
        c[15] = op->on_device ? device_transform_cones(op->dev) : P.num_cones();
        return IFGF_OK;
    }

    int ifgf_plan_orders(const ifgf_operator* op, int* o)
    {
        if (!op || !o)
            return IFGF_EINVAL;
        if (!op->planned)
            return IFGF_ESTATE;
        for (int d = 0; d < 3; d++)
        {
            o[d] = op->plan.base_order[d];
            o[3 + d] = op->plan.ho_order[d];
            o[6 + d] = op->plan.lo_order[d];
            o[9 + d] = op->plan.factor[d];
        }
        return IFGF_OK;
    }

    int ifgf_plan_tree(const ifgf_operator* op, int tree, int64_t* level_off, int64_t* begin,
                       int64_t* end, int32_t* parent, int8_t* leaf, double* bbox6, int64_t* perm)
    {
        if (!op || (tree != 0 && tree != 1))
            return IFGF_EINVAL;
        if (!op->planned)
            return IFGF_ESTATE;
        const Tree& T = tree == 0 ? op->plan.S : op->plan.T;
        const int64_t nb = T.num_boxes();
        if (level_off)
            std::memcpy(level_off, T.level_off.data(), T.level_off.size() * sizeof(int64_t));
        if (begin)
            std::memcpy(begin, T.begin.data(), nb * sizeof(int64_t));
        if (end)
            std::memcpy(end, T.end.data(), nb * sizeof(int64_t));
        if (parent)
            std::memcpy(parent, T.parent.data(), nb * sizeof(int32_t));
        if (leaf)
            for (int64_t g = 0; g < nb; g++)
                leaf[g] = int8_t(T.leaf[g]);
        if (bbox6)
            for (int64_t g = 0; g < nb; g++)
                for (int d = 0; d < 3; d++)
                {
                    bbox6[6 * g + d] = T.blo[3 * g + d];
                    bbox6[6 * g + 3 + d] = T.bhi[3 * g + d];
                }
        if (perm)
            std::memcpy(perm, T.perm.data(), T.perm.size() * sizeof(int64_t));
        return IFGF_OK;
    }

    int ifgf_plan_pairs(const ifgf_operator* op, int far, int32_t* src, int32_t* tgt)
    {
        if (!op || !src || !tgt)
            return IFGF_EINVAL;
        if (!op->planned)
            return IFGF_ESTATE;
        const auto& v = far ? op->plan.far : op->plan.near;
        for (size_t i = 0; i < v.size(); i++)
        {
            src[i] = v[i].src;
            tgt[i] = v[i].tgt;
        }
        return IFGF_OK;
    }

    int ifgf_plan_cones(const ifgf_operator* op, int64_t* nc, int64_t* nf, int64_t* cone_off,
                        int64_t* cone_cell, uint32_t* mask, int64_t* fine_off, int64_t* fine_ids)
    {
        if (!op)
            return IFGF_EINVAL;
        if (!op->planned)
            return IFGF_ESTATE;
        const Plan& P = op->plan;
        const int64_t nb = P.S.num_boxes();
        if (nc)
            std::memcpy(nc, P.nc.data(), 3 * nb * sizeof(int64_t));
        if (nf)
            std::memcpy(nf, P.nf.data(), 3 * nb * sizeof(int64_t));
        if (cone_off)
            std::memcpy(cone_off, P.cone_off.data(), (nb + 1) * sizeof(int64_t));
        if (cone_cell)
            std::memcpy(cone_cell, P.cone_cell.data(), P.cone_cell.size() * sizeof(int64_t));
        if (mask)
            std::memcpy(mask, P.cone_mask.data(), P.cone_mask.size() * sizeof(uint32_t));
        if (fine_off)
        {
            if (int64_t(P.fine_off.size()) != nb + 1)
                return fail(const_cast<ifgf_operator*>(op), IFGF_EINTERNAL,
                            "plan has no fine cone lists");
            std::memcpy(fine_off, P.fine_off.data(), (nb + 1) * sizeof(int64_t));
        }
        if (fine_ids)
            std::memcpy(fine_ids, P.fine_ids.data(), P.fine_ids.size() * sizeof(int64_t));
        return IFGF_OK;
    }

    int ifgf_plan_work(const ifgf_operator* op, double* o)
    {
        if (!op || !o)
            return IFGF_EINVAL;
        if (!op->planned)
            return IFGF_ESTATE;
        const Plan& P = op->plan;
        const Tree& S = P.S;
        const Tree& T = P.T;
        double near = 0, leaf = 0, up = 0, far = 0, leafc = 0;
        const int n_hi = P.ho_order[0] * P.ho_order[1] * P.ho_order[2];
        const int n_lo = P.lo_order[0] * P.lo_order[1] * P.lo_order[2];
        for (const Pair& p : P.near)
            near += double(S.end[p.src] - S.begin[p.src]) * double(T.end[p.tgt] - T.begin[p.tgt]);
        for (const Pair& p : P.far)
            far += double(T.end[p.tgt] - T.begin[p.tgt]);
        for (int64_t c = 0; c < P.num_cones(); c++)
        {
            const int32_t b = P.cone_box[c];
            if (S.leaf[b])
            {
                leafc += 1;
                leaf += double(n_hi) * double(S.end[b] - S.begin[b]);
            }
            else
            {
                int nch = 0;
                for (int j = 0; j < 8; j++)
                    nch += S.child[8 * b + j] >= 0;
                up += double(n_hi) * nch;
            }
        }
        const int h0 = P.ho_order[0], h1 = P.ho_order[1], h2 = P.ho_order[2];
        const int m0 = P.factor[0] * P.lo_order[0], m1 = P.factor[1] * P.lo_order[1],
                  m2 = P.factor[2] * P.lo_order[2];
        o[0] = near;
        o[1] = leaf;
        o[2] = up;
        o[3] = far;
        o[4] = double(P.num_cones());
        o[5] = double(P.n_fine);
        o[6] = leafc;
        o[7] = n_hi;
        o[8] = n_lo;
        o[9] = 4.0 * (double(m0) * h1 * h2 * h0 + double(m0) * m1 * h2 * h1 + double(m0) * m1 * m2 * h2);
        return IFGF_OK;
    }

    int ifgf_last_timings(const ifgf_operator* op, double* ms6)
    {
        if (!op || !ms6)
            return IFGF_EINVAL;
        device_last_timings(op->dev, ms6);
        return IFGF_OK;
    }

    int ifgf_get_launch_shape(const ifgf_operator* op, int* shape)
    {
        if (!op || !shape)
            return IFGF_EINVAL;
        const LaunchShape& s = op->dev.shape;
        shape[0] = s.xcd_log2;
        shape[1] = s.cpb;
        shape[2] = s.up_k;
        shape[3] = s.rb;
        shape[4] = s.far_waves;
        shape[5] = s.overlap;
        return IFGF_OK;
    }

    int ifgf_set_launch_shape(ifgf_operator* op, const int* shape)
    {
        if (!op || !shape)
            return IFGF_EINVAL;
        if (!op->on_device)
            return fail(op, IFGF_ESTATE, "launch shapes need a device init");
        LaunchShape s;
        s.xcd_log2 = shape[0];
        s.cpb = shape[1];
        s.up_k = shape[2];
        s.rb = shape[3];
        s.far_waves = shape[4];
        s.overlap = shape[5];
        std::string e = device_set_shape(op->dev, s);
        if (!e.empty())
            return fail(op, IFGF_EINVAL, e);
        return IFGF_OK;
    }

    int ifgf_autotune(ifgf_operator* op, int64_t nrhs, int reps, int* shape)
    {
        if (!op)
            return IFGF_EINVAL;
        if (!op->on_device)
            return fail(op, IFGF_ESTATE, "autotune needs a device init");
        if (nrhs < 1 || reps < 1)
            return fail(op, IFGF_EINVAL, "autotune needs nrhs >= 1 and reps >= 1");
        LaunchShape best;
        std::string e = device_autotune(op->dev, nrhs, reps, best);
        if (!e.empty())
            return fail(op, IFGF_EINTERNAL, e);
        return shape ? ifgf_get_launch_shape(op, shape) : IFGF_OK;
    }
}
