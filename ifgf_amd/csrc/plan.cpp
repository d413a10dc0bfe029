// plan.cpp -- host construction of the IFGF plan (tree, interaction lists, cone activity).
//
// Follows the reference's init path (ifgfoperator.hpp:65-106) level by level instead of by
// recursion: every stage produces flat arrays that the device kernels read directly.
// Compiled with -ffp-contract=off: the floating-point steps that decide integers (box
// membership, admissibility, cone cells) are written in the reference's operation order so
// the integer results agree exactly with it.
#include "plan.hpp"

#include <algorithm>
#include <atomic>
#include <chrono>
#include <cmath>
#include <cstdio>
#include <cstdlib>
#include <cstring>
#include <limits>
#include <numeric>
#include <thread>

namespace ifgf
{
static const double PI = 3.14159265358979323846;

// IFGF_PLAN_TIMING set: the wall time of each init step on stderr
struct StepClock
{
    bool on = std::getenv("IFGF_PLAN_TIMING") != nullptr;
    std::chrono::steady_clock::time_point t = std::chrono::steady_clock::now();
    void lap(const char* what)
    {
        if (!on)
            return;
        const auto now = std::chrono::steady_clock::now();
        std::fprintf(stderr, "ifgf init: %-28s %8.1f ms\n", what,
                     std::chrono::duration<double, std::milli>(now - t).count());
        t = now;
    }
};

int default_threads()
{
    unsigned n = std::thread::hardware_concurrency();
    return int(std::max(1u, std::min(n, 32u)));
}

template <class F> static void pfor(int64_t n, int nthreads, const F& body)
{
    if (nthreads <= 1 || n < 64)
    {
        for (int64_t i = 0; i < n; i++)
            body(i, 0);
        return;
    }
    std::atomic<int64_t> next{0};
    const int64_t chunk = std::max<int64_t>(1, n / (int64_t(nthreads) * 32));
    std::vector<std::thread> pool;
    for (int t = 0; t < nthreads; t++)
        pool.emplace_back(
            [&, t]()
            {
                for (;;)
                {
                    int64_t b = next.fetch_add(chunk);
                    if (b >= n)
                        return;
                    int64_t e = std::min(n, b + chunk);
                    for (int64_t i = b; i < e; i++)
                        body(i, t);
                }
            });
    for (auto& th : pool)
        th.join();
}

// ---------------------------------------------------------------------------------------
// kernel policies
// ---------------------------------------------------------------------------------------
void KernelSpec::orders(const int base[3], int step, int out[3]) const
{
    for (int i = 0; i < 3; i++)
    {
        out[i] = base[i];
        if (step == 0)
        {
            out[i] = base[i] - 3;
            if (kind == KIND_HELMHOLTZ || kind == KIND_GRAD) // clamp only for these two
                out[i] = std::max(out[i], 2);
        }
    }
}

void KernelSpec::elements(double H, const int64_t base_in[3], int step, int64_t out[3]) const
{
    int64_t base[3] = {base_in[0], base_in[1], base_in[2]};
    if (step == 0)
        for (int i = 0; i < 3; i++)
            base[i] *= fine_factor();
    if (kind == KIND_LAPLACE)
    {
        for (int i = 0; i < 3; i++)
            out[i] = base[i];
        return;
    }
    int mult = 1;
    if (kind == KIND_HELMHOLTZ || kind == KIND_COMBINED)
        mult = int(std::ceil(std::max(kappa * H / 4., 1.)));
    else if (kind == KIND_MODHELMHOLTZ)
        mult = int(std::ceil(std::max(std::abs(k.imag()) * H / (2 * (2 + k.real())), 1.0)));
    else
        mult = int(std::ceil(std::max(std::abs(k.imag()) * H / 4., 1.0) *
                             std::exp(-0.2 * (3.0 / std::sqrt(3.0)) * H * k.real())));
    for (int i = 0; i < 3; i++)
        out[i] = std::max<int64_t>(base[i] * int64_t(mult), 1);
}

double KernelSpec::smin(double H) const
{
    if (kind == KIND_GRAD && tol > 0)
        return std::max(1e-4, std::abs(k.real()) * H / std::abs(std::log(tol)));
    return 1e-4;
}

// ---------------------------------------------------------------------------------------
// Morton keys
// ---------------------------------------------------------------------------------------
static const int KEY_BITS = 40; // fixed-point magnitude bits below the binary point

void morton_key(const double p[3], const double ctr[3], const double diam[3], uint64_t* khi,
                uint64_t* klo)
{
    // per axis: sign class, "nonzero" bit at position 0, and 40 magnitude bits
    uint64_t sign[3], fx[3];
    for (int d = 0; d < 3; d++)
    {
        double v = (p[d] - ctr[d]) / diam[d];
        bool neg = v < 0.0;
        double m = std::fabs(v);
        uint64_t f = uint64_t(std::ldexp(m, KEY_BITS)); // exact: power-of-two scaling
        uint64_t nz = (m != 0.0) ? 1 : 0;
        uint64_t word = (nz << KEY_BITS) | (f & ((uint64_t(1) << KEY_BITS) - 1));
        if (neg)
            word = ~word & ((uint64_t(1) << (KEY_BITS + 1)) - 1);
        sign[d] = neg ? 0 : 1;
        fx[d] = word;
    }
    // 126 bits: [125..123] sign z,y,x ; then positions 0..-40 as (z,y,x) triples
    uint64_t hi = 0, lo = 0;
    auto push = [&](uint64_t bit)
    {
        hi = (hi << 1) | (lo >> 63);
        lo = (lo << 1) | bit;
    };
    push(sign[2]);
    push(sign[1]);
    push(sign[0]);
    for (int b = KEY_BITS; b >= 0; b--)
        for (int d = 2; d >= 0; d--)
            push((fx[d] >> b) & 1);
    *khi = hi;
    *klo = lo;
}

static int fexp(uint64_t bits)
{
    uint64_t u = bits & 0x7fffffffffffffffull;
    if (u == 0 || u >= 0x7ff0000000000000ull)
        return 0;
    u >>= 52;
    return (u == 0) ? -1022 : int(u) - 1023;
}

static int xor_msb(double p, double q)
{
    if (p == q || p == -q)
        return std::numeric_limits<int>::min();
    uint64_t a, b;
    std::memcpy(&a, &p, 8);
    std::memcpy(&b, &q, 8);
    int ea = fexp(a), eb = fexp(b);
    if (ea == eb)
    {
        uint64_t x = (a ^ b) & 0x000fffffffffffffull;
        if (x == 0)
            return ea;
        return ea + (63 - __builtin_clzll(x)) - 52;
    }
    return std::max(ea, eb);
}

bool morton_less_exact(const double* px, const double* qx, const double ctr[3],
                       const double diam[3])
{
    double p[3], q[3];
    for (int i = 0; i < 3; i++)
    {
        p[i] = (px[i] - ctr[i]) / diam[i];
        q[i] = (qx[i] - ctr[i]) / diam[i];
    }
    int best = std::numeric_limits<int>::min(), k = 0;
    for (int j = 2; j >= 0; j--)
    {
        if ((p[j] < 0.0) != (q[j] < 0.0))
            return p[j] < q[j];
        int y = xor_msb(p[j], q[j]);
        if (best < y)
        {
            best = y;
            k = j;
        }
    }
    return p[k] < q[k];
}

// ---------------------------------------------------------------------------------------
// tree (octree.hpp:202-238, 1106-1199)
// ---------------------------------------------------------------------------------------
static double nrm(const double v[3]) { return std::sqrt(v[0] * v[0] + v[1] * v[1] + v[2] * v[2]); }

int Tree::level_of(int64_t g) const
{
    int l = int(std::upper_bound(level_off.begin(), level_off.end(), g) - level_off.begin()) - 1;
    return l;
}

void tree_root(Tree& T, int64_t n, const double mn_in[3], const double mx_in[3], double ctr[3],
               double diam[3])
{
    T = Tree();
    T.n = n;
    double mn[3], mx[3];
    for (int d = 0; d < 3; d++)
    {
        mn[d] = mn_in[d];
        mx[d] = mx_in[d];
    }
    const double a = 1e-8 * nrm(mn), b = 1e-8 * nrm(mx);
    for (int d = 0; d < 3; d++)
    {
        mn[d] -= a;
        mx[d] += b;
        T.lo[d] = mn[d];
        T.hi[d] = mx[d];
    }
    for (int d = 0; d < 3; d++)
    {
        ctr[d] = (mx[d] + mn[d]) / 2.0;
        diam[d] = mx[d] - mn[d];
    }
}

void child_box(const double* blo, const double* bhi, int j, double lo[3], double hi[3])
{
    for (int d = 0; d < 3; d++)
    {
        double size = 0.5 * (bhi[d] - blo[d]);
        double tup = double((j >> d) & 1);
        lo[d] = blo[d] + size * tup;
        hi[d] = lo[d] + size;
    }
}

void grow_tree(Tree& T, int64_t max_leaf, const LevelSplit& split_level)
{
    const int64_t n = T.n;
    // levels: children of every box with more than max_leaf points, greedy closed-box scan
    auto add_box = [&](int64_t b, int64_t e, int32_t parent, const double lo[3], const double hi[3])
    {
        T.begin.push_back(b);
        T.end.push_back(e);
        T.parent.push_back(parent);
        for (int j = 0; j < 8; j++)
            T.child.push_back(-1);
        T.leaf.push_back(1);
        double m = 0;
        for (int d = 0; d < 3; d++)
        {
            T.blo.push_back(lo[d]);
            T.bhi.push_back(hi[d]);
            T.ctr.push_back((hi[d] + lo[d]) / 2.0);
            m = std::max(m, std::abs(hi[d] - lo[d]));
        }
        T.side.push_back(m);
    };
    T.level_off.assign(1, 0);
    T.begin.clear();
    T.end.clear();
    T.parent.clear();
    T.child.clear();
    T.leaf.clear();
    T.blo.clear();
    T.bhi.clear();
    T.ctr.clear();
    T.side.clear();
    add_box(0, n, -1, T.lo, T.hi);
    int64_t lb = 0, le = 1;
    T.levels = 1;
    std::vector<ChildRange> cr;
    std::vector<uint8_t> split;
    while (true)
    {
        cr.assign(le - lb, ChildRange());
        split.assign(le - lb, 0);
        split_level(T, lb, le, max_leaf, cr, split);
        const int64_t nb0 = T.num_boxes();
        for (int64_t q = 0; q < le - lb; q++)
        {
            if (!split[q])
                continue;
            const int64_t g = lb + q;
            for (int j = 0; j < 8; j++)
            {
                if (cr[q].b[j] == cr[q].e[j])
                    continue;
                double cl[3], ch[3];
                child_box(&T.blo[3 * g], &T.bhi[3 * g], j, cl, ch);
                T.child[8 * g + j] = int32_t(T.num_boxes());
                T.leaf[g] = 0;
                add_box(cr[q].b[j], cr[q].e[j], int32_t(g), cl, ch);
            }
        }
        if (T.num_boxes() == nb0)
            break;
        T.level_off.push_back(nb0);
        T.levels++;
        lb = nb0;
        le = T.num_boxes();
    }
    T.level_off.push_back(T.num_boxes());
}

void build_tree(const double* pts, int64_t n, int64_t max_leaf, Tree& T, int nthreads)
{
    double mn[3] = {pts[0], pts[1], pts[2]}, mx[3] = {pts[0], pts[1], pts[2]};
    for (int64_t i = 0; i < n; i++)
        for (int d = 0; d < 3; d++)
        {
            mn[d] = std::min(mn[d], pts[3 * i + d]);
            mx[d] = std::max(mx[d], pts[3 * i + d]);
        }
    double ctr[3], diam[3];
    tree_root(T, n, mn, mx, ctr, diam);

    // sort by 126-bit keys, then order equal-key runs with the exact comparison
    struct KI
    {
        uint64_t hi, lo;
        int64_t idx;
    };
    std::vector<KI> keys(n);
    pfor(n, nthreads,
         [&](int64_t i, int)
         {
             keys[i].idx = i;
             morton_key(pts + 3 * i, ctr, diam, &keys[i].hi, &keys[i].lo);
         });
    std::sort(keys.begin(), keys.end(),
              [](const KI& x, const KI& y)
              {
                  if (x.hi != y.hi)
                      return x.hi < y.hi;
                  if (x.lo != y.lo)
                      return x.lo < y.lo;
                  return x.idx < y.idx;
              });
    for (int64_t i = 0; i < n;)
    {
        int64_t j = i + 1;
        while (j < n && keys[j].hi == keys[i].hi && keys[j].lo == keys[i].lo)
            j++;
        if (j - i > 1)
            std::stable_sort(keys.begin() + i, keys.begin() + j,
                             [&](const KI& x, const KI& y)
                             { return morton_less_exact(pts + 3 * x.idx, pts + 3 * y.idx, ctr, diam); });
        i = j;
    }
    T.perm.resize(n);
    T.pts.resize(3 * size_t(n));
    for (int64_t i = 0; i < n; i++)
    {
        T.perm[i] = keys[i].idx;
        for (int d = 0; d < 3; d++)
            T.pts[3 * i + d] = pts[3 * keys[i].idx + d];
    }
    keys.clear();
    keys.shrink_to_fit();

    grow_tree(T, max_leaf,
              [nthreads](const Tree& T, int64_t lb, int64_t le, int64_t max_leaf,
                         std::vector<ChildRange>& cr, std::vector<uint8_t>& split)
              {
                  pfor(le - lb, nthreads,
                       [&](int64_t q, int)
                       {
                           const int64_t g = lb + q;
                           const int64_t b = T.begin[g], e = T.end[g];
                           if (e - b <= max_leaf)
                               return;
                           split[q] = 1;
                           int64_t pi = b;
                           for (int j = 0; j < 8; j++)
                           {
                               double cl[3], ch[3];
                               child_box(&T.blo[3 * g], &T.bhi[3 * g], j, cl, ch);
                               int64_t ep = pi;
                               while (ep < e)
                               {
                                   const double* p = &T.pts[3 * ep];
                                   if (!(cl[0] <= p[0] && cl[1] <= p[1] && cl[2] <= p[2] &&
                                         p[0] <= ch[0] && p[1] <= ch[1] && p[2] <= ch[2]))
                                       break;
                                   ++ep;
                               }
                               cr[q].b[j] = pi;
                               cr[q].e[j] = ep;
                               pi = ep;
                           }
                       });
              });
}

// ---------------------------------------------------------------------------------------
// interaction lists (octree.hpp:255-349), breadth first
// ---------------------------------------------------------------------------------------
static double sq_ext(const double* lo, const double* hi, const double* p)
{
    double d2 = 0;
    for (int k = 0; k < 3; k++)
    {
        if (lo[k] > p[k])
        {
            double a = lo[k] - p[k];
            d2 += a * a;
        }
        else if (p[k] > hi[k])
        {
            double a = p[k] - hi[k];
            d2 += a * a;
        }
    }
    return d2;
}

void build_interactions(Plan& P, const std::function<bool(double)>& cutoff)
{
    const Tree& S = P.S;
    const Tree& T = P.T;
    P.far.clear();
    P.near.clear();
    P.n_tofar = 0;
    const double eta = std::sqrt(3.0);
    std::vector<Pair> cur{{0, 0}}, nxt;
    while (!cur.empty())
    {
        nxt.clear();
        for (const Pair& pr : cur)
        {
            const int32_t s = pr.src, t = pr.tgt;
            const double ed = std::sqrt(sq_ext(&T.blo[3 * t], &T.bhi[3 * t], &S.ctr[3 * s]));
            if (ed >= eta * S.side[s])
            {
                if (cutoff && cutoff(ed))
                    P.n_tofar++;
                else
                    P.far.push_back(pr);
                continue;
            }
            const bool sl = S.leaf[s], tl = T.leaf[t];
            if (sl && tl)
            {
                P.near.push_back(pr);
                continue;
            }
            bool src_side;
            if (sl)
                src_side = false;
            else if (tl)
                src_side = true;
            else
                src_side = S.side[s] > T.side[t];
            for (int j = 0; j < 8; j++)
            {
                if (src_side)
                {
                    int32_t c = S.child[8 * s + j];
                    if (c >= 0)
                        nxt.push_back({c, t});
                }
                else
                {
                    int32_t c = T.child[8 * t + j];
                    if (c >= 0)
                        nxt.push_back({s, c});
                }
            }
        }
        std::swap(cur, nxt);
    }
    auto by_tgt = [](const Pair& a, const Pair& b)
    { return a.tgt != b.tgt ? a.tgt < b.tgt : a.src < b.src; };
    std::sort(P.far.begin(), P.far.end(), by_tgt);
    std::sort(P.near.begin(), P.near.end(), by_tgt);
    interaction_flags(P);
}

void interaction_flags(Plan& P)
{
    const Tree& S = P.S;
    const int64_t nb = S.num_boxes();
    std::vector<int64_t> nfar(nb, 0);
    for (const Pair& p : P.far)
        nfar[p.src]++;
    P.has_far_anc.assign(nb, 0);
    P.parent_has_far.assign(nb, 0);
    for (int64_t g = 0; g < nb; g++) // parents precede children in global order
    {
        int32_t p = S.parent[g];
        P.parent_has_far[g] = (p >= 0) && (nfar[p] > 0 || P.parent_has_far[p]);
        P.has_far_anc[g] = nfar[g] > 0 || P.parent_has_far[g];
    }
}

// ---------------------------------------------------------------------------------------
// cone geometry (util.hpp:96-171, cone_domain.hpp:145-256)
// ---------------------------------------------------------------------------------------
static inline void interp_to_cart(const double p[3], const double xc[3], double H, double out[3])
{
    double r = H / p[0];
    out[0] = xc[0] + r * std::cos(p[2]) * std::sin(p[1]);
    out[1] = xc[1] + r * std::sin(p[2]) * std::sin(p[1]);
    out[2] = xc[2] + r * std::cos(p[1]);
}

static inline void cart_to_interp(const double x[3], const double xc[3], double H, double out[3])
{
    double p0 = x[0] - xc[0], p1 = x[1] - xc[1], p2 = x[2] - xc[2];
    double a = p0 * p0 + p1 * p1;
    out[2] = std::atan2(p1, p0);
    out[1] = std::atan2(std::sqrt(a), p2);
    out[0] = H / std::sqrt(a + p2 * p2);
}

struct Grid
{
    int64_t n[3];
    double lo[3], hi[3], h[3];
    int64_t total() const { return n[0] * n[1] * n[2]; }
    int64_t cell(const double p[3]) const
    {
        if (sq_ext(lo, hi, p) > 0)
            return -1;
        int64_t id = 0, stride = 1;
        for (int j = 0; j < 3; j++)
        {
            int q = int(std::floor((p[j] - lo[j]) / h[j]));
            int64_t c = std::clamp<int64_t>(q, 0, n[j] - 1);
            id += c * stride;
            stride *= n[j];
        }
        return id;
    }
    void point(int64_t cellid, const double xi[3], double out[3]) const
    {
        for (int i = 0; i < 3; i++)
        {
            int64_t idx = cellid % n[i];
            cellid /= n[i];
            double a = 0.5 * h[i];
            double b = lo[i] + (double(idx) + 0.5) * h[i];
            out[i] = (xi[i] * a) + b;
        }
    }
};

std::vector<double> chebnodes1d(int n)
{
    std::vector<double> x(n);
    for (int i = 0; i < n; i++)
        x[i] = std::cos((2. * i + 1.0) / (2. * n) * PI);
    return x;
}

static std::vector<double> chebnodes3d(const int ns[3])
{
    auto r = chebnodes1d(ns[0]), t = chebnodes1d(ns[1]), p = chebnodes1d(ns[2]);
    std::vector<double> out;
    out.reserve(size_t(3) * ns[0] * ns[1] * ns[2]);
    for (int i = 0; i < ns[2]; i++)
        for (int j = 0; j < ns[1]; j++)
            for (int k = 0; k < ns[0]; k++)
            {
                out.push_back(r[k]);
                out.push_back(t[j]);
                out.push_back(p[i]);
            }
    return out;
}

std::vector<double> axis_map(int o_hi, int o_lo, int f)
{
    typedef long double LD;
    const LD pi = 3.141592653589793238462643383279502884L;
    std::vector<double> M(size_t(f) * o_lo * o_hi, 0.0);
    auto Tk = [](int k, LD x) { return std::cos(LD(k) * std::acos(std::clamp(x, LD(-1), LD(1)))); };
    for (int j = 0; j < f; j++)
        for (int q = 0; q < o_lo; q++)
            for (int s = 0; s < o_hi; s++)
            {
                LD acc = 0;
                for (int m = 0; m < o_lo; m++)
                {
                    LD xl = std::cos((2.0L * m + 1.0L) / (2.0L * o_lo) * pi);
                    LD u = xl / LD(f) + (-1.0L + (2.0L * j + 1.0L) / LD(f));
                    LD val = 0; // value at u of the interpolant of the unit vector e_s
                    for (int k = 0; k < o_hi; k++)
                    {
                        LD tks = std::cos(LD(k) * pi * (2.0L * s + 1.0L) / (2.0L * o_hi));
                        val += (k == 0 ? 1.0L : 2.0L) / LD(o_hi) * tks * Tk(k, u);
                    }
                    LD tq = std::cos(LD(q) * pi * (2.0L * m + 1.0L) / (2.0L * o_lo));
                    acc += (q == 0 ? 1.0L : 2.0L) / LD(o_lo) * tq * val;
                }
                M[(size_t(j) * o_lo + q) * o_hi + s] = double(acc);
            }
    return M;
}

// ---------------------------------------------------------------------------------------
// order estimation for tol > 0 (ifgfoperator.hpp:108-244, util.hpp:214-234)
// ---------------------------------------------------------------------------------------
static cplx factored_host(const Plan& P, const double* xs, const double* y, double dc, int64_t i)
{
    // unit weight (ifgfoperator.hpp:184); the families' evaluateFactoredKernel
    const KernelSpec& K = P.K;
    double z[3] = {xs[0] - y[0], xs[1] - y[1], xs[2] - y[2]};
    double d2 = z[0] * z[0] + z[1] * z[1] + z[2] * z[2];
    double d = std::sqrt(d2);
    const double* n = P.normals.empty() ? nullptr : &P.normals[3 * i];
    switch (K.kind)
    {
    case KIND_HELMHOLTZ: // helmholtz_ifgf.hpp:114-156
        return std::exp(cplx(0, K.kappa) * (d - dc)) * dc / d;
    case KIND_LAPLACE: // laplace_ifgf.hpp:109-131
        return (d == 0) ? cplx(0) : cplx(dc / d);
    case KIND_DOUBLELAYER: // double_layer_helmholtz_ifgf.hpp:131-159
    {
        double id = (d2 > 1e-12) ? 1.0 / d : 0;
        double dd = d2 * id;
        double w = -(1.0 / d2) * dc * (n ? (z[0] * n[0] + z[1] * n[1] + z[2] * n[2]) : 0.0);
        return std::exp(-K.k * (dd - dc)) * (-K.k - id) * w;
    }
    case KIND_COMBINED: // combined_field_helmholtz_ifgf.hpp:131-182 (z = y - x)
    {
        if (!(d2 > 1e-14))
            return 0;
        double nxy = n ? -(z[0] * n[0] + z[1] * n[1] + z[2] * n[2]) : 0.0;
        double id = 1.0 / d, dd = d2 * id;
        double s = std::sin(K.kappa * (dd - dc)), c = std::cos(K.kappa * (dd - dc));
        double f = dc / (d2 * dd);
        return cplx(f * (c * nxy + s * K.kappa * (dd * nxy + d2)),
                    f * (-c * K.kappa * (dd * nxy + d2) + s * nxy));
    }
    case KIND_GRAD: // grad_helmholtz_ifgf.hpp:170-214
    {
        double id = d2 > 1e-12 ? 1.0 / d : 0.0;
        double dd = d2 * id;
        if (K.dx < 0)
            return dc * id * std::exp(-K.k * (dd - dc));
        double w = -(1.0 / d2) * dc * z[K.dx];
        return std::exp(-K.k * (dd - dc)) * (-K.k - id) * w;
    }
    default: // modified_helmholtz_ifgf.hpp:114-136
        return (d == 0) ? cplx(0) : std::exp(-K.k * (d - dc)) * dc / d;
    }
}

static void chebtransform_host(std::vector<cplx>& v, const int ns[3])
{
    // separable c_k = (k==0?1:2)/n sum_s v_s T_k(x_s) along each axis (chebinterp.hpp:176-250)
    int stride = 1;
    for (int d = 0; d < 3; d++)
    {
        const int n = ns[d];
        const int outer = int(v.size()) / (stride * n);
        std::vector<cplx> tmp(v.size());
        for (int o = 0; o < outer; o++)
            for (int in = 0; in < stride; in++)
                for (int k = 0; k < n; k++)
                {
                    cplx acc = 0;
                    for (int s = 0; s < n; s++)
                        acc += v[(size_t(o) * n + s) * stride + in] *
                               std::cos(k * PI * (2. * s + 1.) / (2. * n));
                    tmp[(size_t(o) * n + k) * stride + in] = acc * ((k == 0 ? 1. : 2.) / n);
                }
        v.swap(tmp);
        stride *= n;
    }
}

int estimate_orders(Plan& P, double tol, int out[3])
{
    const Tree& S = P.S;
    const int level = S.levels - 1;
    const int64_t l0 = S.level_off[level], nb = S.level_off[level + 1] - l0;
    const int64_t stride = nb / 10;
    int best[3] = {0, 0, 0};
    for (int i = 0; i < 10; i++)
    {
        const int64_t g = l0 + i * stride;
        const double H = S.side[g];
        const double* xc = &S.ctr[3 * g];
        Grid gr;
        double smin = P.K.smin(H), smax = std::sqrt(3.0) / 3;
        gr.lo[0] = smin;
        gr.hi[0] = smax;
        gr.lo[1] = 0;
        gr.hi[1] = PI;
        gr.lo[2] = -PI;
        gr.hi[2] = PI;
        int base[3] = {0, 0, 0}, new_p[3] = {P.base_order[0], P.base_order[1], P.base_order[2]};
        double err[3] = {1e300, 1e300, 1e300};
        while (std::max(err[0], std::max(err[1], err[2])) > tol &&
               std::max(base[0], std::max(base[1], base[2])) < 10)
        {
            for (int d = 0; d < 3; d++)
                new_p[d] = P.base_order[d] + base[d];
            int plus[3] = {new_p[0] + 4, new_p[1] + 4, new_p[2] + 4}, ord[3];
            P.K.orders(plus, 1, ord);
            P.K.elements(H, P.base_els, 1, gr.n);
            for (int d = 0; d < 3; d++)
                gr.h[d] = (gr.hi[d] - gr.lo[d]) / double(gr.n[d]);
            auto nodes = chebnodes3d(ord);
            const size_t nn = nodes.size() / 3;
            std::vector<cplx> data(nn);
            for (size_t q = 0; q < nn; q++)
            {
                double sp[3], y[3];
                gr.point(0, &nodes[3 * q], sp);
                interp_to_cart(sp, xc, H, y);
                double dv[3] = {y[0] - xc[0], y[1] - xc[1], y[2] - xc[2]};
                double dc = nrm(dv);
                cplx acc = 0;
                for (int64_t s = S.begin[g]; s < S.end[g]; s++)
                    acc += factored_host(P, &S.pts[3 * s], y, dc, s);
                data[q] = acc;
            }
            chebtransform_host(data, ord);
            for (int d = 0; d < 3; d++)
            {
                double v1 = 0, v2 = 0;
                for (size_t idx = 0; idx < nn; idx++)
                {
                    size_t sp[3], jj = idx;
                    for (int a = 0; a < 3; a++)
                    {
                        sp[a] = jj % size_t(ord[a]);
                        jj /= size_t(ord[a]);
                    }
                    double nn2 = std::norm(data[idx]);
                    v1 += nn2;
                    if (sp[d] == size_t(ord[d] - 3))
                        v2 += nn2;
                }
                err[d] = std::sqrt(v2) / std::max(1., std::sqrt(v1));
            }
            for (int d = 0; d < 3; d++)
                if (err[d] > tol)
                    base[d] += 1;
        }
        for (int d = 0; d < 3; d++)
            best[d] = std::max(best[d], new_p[d]);
    }
    for (int d = 0; d < 3; d++)
        out[d] = best[d];
    return 0;
}

// transfer factor e^{-k(d - dp)} dp / d with d = |x - xc|, dp = |x - pxc| (helmholtz_ifgf.hpp:67-92
// and siblings; Laplace: dp / d)
static cplx transfer_host(const KernelSpec& K, const double* x, const double* xc, const double* pxc)
{
    double a[3] = {x[0] - xc[0], x[1] - xc[1], x[2] - xc[2]};
    double b[3] = {x[0] - pxc[0], x[1] - pxc[1], x[2] - pxc[2]};
    const double d = nrm(a), dp = nrm(b);
    if (K.kind == KIND_LAPLACE)
        return cplx(dp / d, 0.0);
    return std::exp(-K.k * (d - dp)) * (dp / d);
}

// ---------------------------------------------------------------------------------------
// cone activity (octree.hpp:351-816, TwoGrid branch)
// ---------------------------------------------------------------------------------------
// grid of box g's coarse (step 1) or fine (step 0) cones, from the plan's per-box arrays
static Grid box_grid(const Plan& P, int64_t g, int step)
{
    Grid G;
    for (int d = 0; d < 3; d++)
    {
        G.n[d] = step ? P.nc[3 * g + d] : P.nf[3 * g + d];
        G.lo[d] = P.dom_lo[3 * g + d];
        G.hi[d] = P.dom_hi[3 * g + d];
        G.h[d] = step ? P.hc[3 * g + d] : P.hf[3 * g + d];
    }
    return G;
}

static bool same_grid(const Plan& P, int64_t a, int64_t b)
{
    for (int d = 0; d < 3; d++)
        if (P.nc[3 * a + d] != P.nc[3 * b + d] || P.nf[3 * a + d] != P.nf[3 * b + d] ||
            P.hc[3 * a + d] != P.hc[3 * b + d] || P.hf[3 * a + d] != P.hf[3 * b + d] ||
            P.dom_lo[3 * a + d] != P.dom_lo[3 * b + d] || P.dom_hi[3 * a + d] != P.dom_hi[3 * b + d])
            return false;
    return true;
}

// upward geometry of parent level l (children on level l + 1), see UpLevel
static void build_up_level(Plan& P, int l, const std::vector<double>& ho_nodes, int nthreads)
{
    const Tree& S = P.S;
    UpLevel& U = P.up[l];
    U = UpLevel();
    const int64_t p0 = S.level_off[l], p1 = S.level_off[l + 1];
    const int64_t c0 = S.level_off[l + 1], c1 = S.level_off[l + 2];
    // one grid class for all parents and all children of the level
    // boxes of a level are translates up to rounding of their extents (a child's extent is
    // (lo + size) - lo); sides must agree to 1e-12 and the grids exactly
    auto close = [](double a, double b) { return std::abs(a - b) <= 1e-12 * std::abs(b); };
    bool any = false;
    for (int64_t g = p0; g < p1; g++)
    {
        if (!same_grid(P, g, p0) || !close(S.side[g], S.side[p0]))
            return;
        any |= (!S.leaf[g] && P.cone_off[g + 1] > P.cone_off[g]);
    }
    if (!any)
        return;
    for (int64_t g = c0; g < c1; g++)
        if (!same_grid(P, g, c0) || !close(S.side[g], S.side[c0]))
            return;
    const size_t n_ho = ho_nodes.size() / 3;
    const Grid pg = box_grid(P, p0, 1);
    U.cellmap.assign(size_t(pg.total()), -1);
    int32_t nu = 0;
    for (int64_t g = p0; g < p1; g++)
        if (!S.leaf[g])
            for (int64_t c = P.cone_off[g]; c < P.cone_off[g + 1]; c++)
                if (U.cellmap[P.cone_cell[c]] < 0)
                    U.cellmap[P.cone_cell[c]] = nu++;
    if (int64_t(nu) * int64_t(n_ho) * 8 > P.up_budget)
    {
        U.cellmap.clear();
        return;
    }
    std::vector<int64_t> cells(nu);
    for (int64_t c = 0; c < int64_t(U.cellmap.size()); c++)
        if (U.cellmap[c] >= 0)
            cells[U.cellmap[c]] = c;
    // representative parent (the first box of the level) and its eight child boxes, formed as
    // build_tree forms children
    const double* pxc = &S.ctr[3 * p0];
    const double pH = S.side[p0];
    double cctr[8][3], cH[8];
    for (int o = 0; o < 8; o++)
    {
        double lo[3], hi[3], m = 0;
        for (int d = 0; d < 3; d++)
        {
            double size = 0.5 * (S.bhi[3 * p0 + d] - S.blo[3 * p0 + d]);
            lo[d] = S.blo[3 * p0 + d] + size * double((o >> d) & 1);
            hi[d] = lo[d] + size;
            cctr[o][d] = (hi[d] + lo[d]) / 2.0;
            m = std::max(m, std::abs(hi[d] - lo[d]));
        }
        cH[o] = m;
        if (!close(m, S.side[c0]))
            return; // children differ from the representative: keep the per-box path
    }
    const Grid fg = box_grid(P, c0, 0), cg = box_grid(P, c0, 1);
    U.entries.assign(size_t(nu) * n_ho * 8, UpEntry());
    const int* f = P.factor;
    pfor(int64_t(nu) * int64_t(n_ho), nthreads,
         [&](int64_t e, int)
         {
             const int64_t u = e / int64_t(n_ho);
             const size_t q = size_t(e % int64_t(n_ho));
             double sp[3], y[3];
             pg.point(cells[u], &ho_nodes[3 * q], sp);
             interp_to_cart(sp, pxc, pH, y);
             for (int o = 0; o < 8; o++)
             {
                 UpEntry& E = U.entries[size_t(e) * 8 + o];
                 double s[3];
                 cart_to_interp(y, cctr[o], cH[o], s);
                 const int64_t fc = fg.cell(s);
                 if (fc < 0)
                 {
                     E.fine = E.coarse = E.cf = -1;
                     E.sub = 0;
                     continue;
                 }
                 E.fine = int32_t(fc);
                 E.coarse = int32_t(cg.cell(s));
                 int64_t t = fc, idx[3];
                 for (int d = 0; d < 3; d++)
                 {
                     idx[d] = t % fg.n[d];
                     t /= fg.n[d];
                 }
                 E.cf = int32_t((idx[0] / f[0]) + cg.n[0] * ((idx[1] / f[1]) + cg.n[1] * (idx[2] / f[2])));
                 E.sub = int32_t((idx[0] % f[0]) + f[0] * ((idx[1] % f[1]) + f[1] * (idx[2] % f[2])));
                 for (int d = 0; d < 3; d++)
                 {
                     const double a = 0.5 * fg.h[d];
                     const double b = fg.lo[d] + (double(idx[d]) + 0.5) * fg.h[d];
                     E.xi[d] = (s[d] - b) / a;
                 }
                 const cplx tf = transfer_host(P.K, y, cctr[o], pxc);
                 E.tf[0] = tf.real();
                 E.tf[1] = tf.imag();
             }
         });
    U.valid = true;
}

void build_cones(Plan& P, int nthreads, const ConeMarker& marker)
{
    const Tree& S = P.S;
    const Tree& T = P.T;
    const int64_t nb = S.num_boxes();
    P.dom_lo.assign(3 * nb, 0);
    P.dom_hi.assign(3 * nb, 0);
    P.nc.assign(3 * nb, 0);
    P.nf.assign(3 * nb, 0);
    P.hc.assign(3 * nb, 0);
    P.hf.assign(3 * nb, 0);
    P.table_off.assign(nb + 1, 0);
    P.cone_off.assign(nb + 1, 0);
    P.fine_off.assign(nb + 1, 0);
    P.cone_box.clear();
    P.cone_cell.clear();
    P.cone_mask.clear();
    P.cone_fine.clear();
    P.table.clear();
    P.fine_ids.clear();
    P.n_fine = 0;
    P.n_fine_uncovered = 0;
    P.up.assign(S.levels, UpLevel());

    // far pairs grouped by source box
    std::vector<int64_t> far_off(nb + 1, 0);
    for (const Pair& p : P.far)
        far_off[p.src + 1]++;
    for (int64_t g = 0; g < nb; g++)
        far_off[g + 1] += far_off[g];
    std::vector<int32_t> far_tgt(P.far.size());
    {
        std::vector<int64_t> fill(far_off.begin(), far_off.end() - 1);
        for (const Pair& p : P.far)
            far_tgt[fill[p.src]++] = p.tgt;
    }

    const int* f = P.factor;
    const auto ho_nodes = chebnodes3d(P.ho_order);
    const size_t n_ho = ho_nodes.size() / 3;

    for (int level = 0; level < S.levels; level++)
    {
        const int64_t l0 = S.level_off[level], l1 = S.level_off[level + 1];
        // cone grids of the level (octree.hpp:487-548)
        for (int64_t g = l0; g < l1; g++)
        {
            const double H = S.side[g];
            double smax = std::sqrt(3.0) / 3;
            double smin = P.K.smin(H);
            if (smin >= smax)
                smin = smax - 0.05;
            const double dlo[3] = {smin, 0, -PI}, dhi[3] = {smax, PI, PI};
            int64_t nf[3], nc[3];
            P.K.elements(H, P.base_els, 0, nf);
            P.K.elements(H, P.base_els, 1, nc);
            for (int d = 0; d < 3; d++)
            {
                P.dom_lo[3 * g + d] = dlo[d];
                P.dom_hi[3 * g + d] = dhi[d];
                P.nf[3 * g + d] = nf[d];
                P.nc[3 * g + d] = nc[d];
                P.hf[3 * g + d] = (dhi[d] - dlo[d]) / double(nf[d]);
                P.hc[3 * g + d] = (dhi[d] - dlo[d]) / double(nc[d]);
            }
        }
        const auto t_a = std::chrono::steady_clock::now();
        if (level > 0)
            build_up_level(P, level - 1, ho_nodes, nthreads);
        const auto t_b = std::chrono::steady_clock::now();
        // activity bitmaps of the level's boxes (64-bit words; fine and coarse grids)
        LevelBits lb;
        lb.foff.assign(l1 - l0 + 1, 0);
        lb.coff.assign(l1 - l0 + 1, 0);
        for (int64_t q = 0; q < l1 - l0; q++)
        {
            const int64_t g = l0 + q;
            const int64_t nfc = P.nf[3 * g] * P.nf[3 * g + 1] * P.nf[3 * g + 2];
            const int64_t ncc = P.nc[3 * g] * P.nc[3 * g + 1] * P.nc[3 * g + 2];
            lb.foff[q + 1] = lb.foff[q] + (nfc + 63) / 64;
            lb.coff[q + 1] = lb.coff[q] + (ncc + 63) / 64;
        }
        lb.fbits.assign(size_t(lb.foff[l1 - l0]), 0);
        lb.cbits.assign(size_t(lb.coff[l1 - l0]), 0);
        if (marker)
            marker(P, level, far_off, far_tgt, ho_nodes, lb);
        else if (l1 > l0)
            pfor(l1 - l0, nthreads,
                 [&](int64_t q, int)
                 {
                     const int64_t g = l0 + q;
                     const double H = S.side[g];
                     const double* xc = &S.ctr[3 * g];
                     const Grid fg = box_grid(P, g, 0), cg = box_grid(P, g, 1);
                     uint64_t* fbits = &lb.fbits[lb.foff[q]];
                     uint64_t* cbits = &lb.cbits[lb.coff[q]];
                     auto mark_ids = [&](int64_t c0, int64_t c1)
                     {
                         if (c1 >= 0)
                             cbits[c1 >> 6] |= uint64_t(1) << (c1 & 63);
                         if (c0 >= 0)
                             fbits[c0 >> 6] |= uint64_t(1) << (c0 & 63);
                     };
                     auto mark = [&](const double s[3]) { mark_ids(fg.cell(s), cg.cell(s)); };
                     for (int64_t r = far_off[g]; r < far_off[g + 1]; r++)
                     {
                         const int32_t t = far_tgt[r];
                         for (int64_t i = T.begin[t]; i < T.end[t]; i++)
                         {
                             double s[3];
                             cart_to_interp(&T.pts[3 * i], xc, H, s);
                             mark(s);
                         }
                     }
                     if (P.parent_has_far[g])
                     {
                         const int32_t pg = S.parent[g];
                         const UpLevel& U = P.up[level - 1];
                         if (U.valid)
                         {
                             int oct = 0;
                             while (S.child[8 * pg + oct] != int32_t(g))
                                 oct++;
                             for (int64_t c = P.cone_off[pg]; c < P.cone_off[pg + 1]; c++)
                             {
                                 const int64_t u = U.cellmap[P.cone_cell[c]];
                                 for (size_t k = 0; k < n_ho; k++)
                                 {
                                     const UpEntry& E = U.entries[(size_t(u) * n_ho + k) * 8 + oct];
                                     if (E.fine >= 0)
                                         mark_ids(E.fine, E.coarse);
                                 }
                             }
                         }
                         else
                         {
                             const double pH = S.side[pg];
                             const double* pxc = &S.ctr[3 * pg];
                             const Grid pgrid = box_grid(P, pg, 1);
                             for (int64_t c = P.cone_off[pg]; c < P.cone_off[pg + 1]; c++)
                             {
                                 const int64_t cell = P.cone_cell[c];
                                 for (size_t k = 0; k < n_ho; k++)
                                 {
                                     double sp[3], y[3], s[3];
                                     pgrid.point(cell, &ho_nodes[3 * k], sp);
                                     interp_to_cart(sp, pxc, pH, y);
                                     cart_to_interp(y, xc, H, s);
                                     mark(s);
                                 }
                             }
                         }
                     }
                 });
        const auto t_m = std::chrono::steady_clock::now();
        struct Out
        {
            std::vector<int64_t> cells;
            std::vector<uint32_t> masks;
            std::vector<int64_t> fine;
            int64_t uncovered = 0;
        };
        std::vector<Out> out(l1 - l0);
        pfor(l1 - l0, nthreads,
             [&](int64_t q, int)
             {
                 const int64_t g = l0 + q;
                 const Grid fg = box_grid(P, g, 0), cg = box_grid(P, g, 1);
                 const uint64_t* fbits = &lb.fbits[lb.foff[q]];
                 const uint64_t* cbits = &lb.cbits[lb.coff[q]];
                 const int64_t nfw = lb.foff[q + 1] - lb.foff[q];
                 const int64_t ncw = lb.coff[q + 1] - lb.coff[q];
                 Out& o = out[q];
                 for (int64_t w = 0; w < ncw; w++)
                     for (uint64_t bits = cbits[w]; bits; bits &= bits - 1)
                         o.cells.push_back(w * 64 + __builtin_ctzll(bits));
                 o.masks.assign(o.cells.size(), 0);
                 for (size_t m = 0; m < o.cells.size(); m++)
                 {
                     int64_t c = o.cells[m];
                     int64_t ci[3];
                     for (int d = 0; d < 3; d++)
                     {
                         ci[d] = c % cg.n[d];
                         c /= cg.n[d];
                     }
                     uint32_t mask = 0;
                     for (int j2 = 0; j2 < f[2]; j2++)
                         for (int j1 = 0; j1 < f[1]; j1++)
                             for (int j0 = 0; j0 < f[0]; j0++)
                             {
                                 int64_t fc = (ci[0] * f[0] + j0) +
                                              fg.n[0] * ((ci[1] * f[1] + j1) + fg.n[1] * (ci[2] * f[2] + j2));
                                 if (fbits[fc >> 6] >> (fc & 63) & 1)
                                     mask |= 1u << (j0 + f[0] * (j1 + f[1] * j2));
                             }
                     o.masks[m] = mask;
                 }
                 for (int64_t w = 0; w < nfw; w++)
                     for (uint64_t bits = fbits[w]; bits; bits &= bits - 1)
                     {
                         int64_t fc = w * 64 + __builtin_ctzll(bits);
                         o.fine.push_back(fc);
                         int64_t t = fc, cc = 0, stride = 1;
                         for (int d = 0; d < 3; d++)
                         {
                             int64_t idx = t % fg.n[d];
                             t /= fg.n[d];
                             cc += (idx / f[d]) * stride;
                             stride *= cg.n[d];
                         }
                         if (!(cbits[cc >> 6] >> (cc & 63) & 1))
                             o.uncovered++;
                     }
             });
        const auto t_c = std::chrono::steady_clock::now();
        if (std::getenv("IFGF_PLAN_TIMING"))
            std::fprintf(stderr,
                         "ifgf init:   level %d: upward table %.1f ms, marking %.1f ms, "
                         "cone lists %.1f ms\n",
                         level, std::chrono::duration<double, std::milli>(t_b - t_a).count(),
                         std::chrono::duration<double, std::milli>(t_m - t_b).count(),
                         std::chrono::duration<double, std::milli>(t_c - t_m).count());
        for (int64_t q = 0; q < l1 - l0; q++)
        {
            const int64_t g = l0 + q;
            Out& o = out[q];
            const int64_t ncell = P.nc[3 * g] * P.nc[3 * g + 1] * P.nc[3 * g + 2];
            P.table_off[g + 1] = P.table_off[g] + ncell;
            P.cone_off[g + 1] = P.cone_off[g] + int64_t(o.cells.size());
            P.fine_off[g + 1] = P.fine_off[g] + int64_t(o.fine.size());
            P.n_fine_uncovered += o.uncovered;
            for (size_t m = 0; m < o.cells.size(); m++)
            {
                P.cone_box.push_back(int32_t(g));
                P.cone_cell.push_back(o.cells[m]);
                P.cone_mask.push_back(o.masks[m]);
                P.cone_fine.push_back(P.n_fine);
                P.n_fine += __builtin_popcount(o.masks[m]);
            }
            P.fine_ids.insert(P.fine_ids.end(), o.fine.begin(), o.fine.end());
            o = Out();
        }
    }
    // the deepest level has no children: no table
    P.table.assign(P.table_off[nb], -1);
    for (int64_t c = 0; c < P.num_cones(); c++)
        P.table[P.table_off[P.cone_box[c]] + P.cone_cell[c]] = int32_t(c);
}

std::string build_plan(Plan& P, const double* srcs, int64_t N, const double* tgts, int64_t M,
                       const double* normals, int64_t leaf, int order, int n_elem,
                       const std::function<bool(double)>& cutoff, int nthreads,
                       const PlanBuilders& dev)
{
    const TreeBuilder& tree_builder = dev.tree;
    const ConeMarker& cone_marker = dev.cones;
    if (N <= 0 || M <= 0)
        return "empty point set";
    if (n_elem <= 0)
        return "n_elem must be positive";
    for (int d = 0; d < 3; d++)
        P.base_order[d] = order;
    P.base_els[0] = 1 * int64_t(n_elem);
    P.base_els[1] = 2 * int64_t(n_elem);
    P.base_els[2] = 4 * int64_t(n_elem);
    int64_t max_leaf = leaf < 0 ? std::numeric_limits<int64_t>::max() : leaf;
    StepClock clk;
    if (tree_builder)
    {
        std::string e = tree_builder(srcs, N, max_leaf, P.S);
        if (e.empty())
            e = tree_builder(tgts, M, max_leaf, P.T);
        if (!e.empty())
            return e;
    }
    else
    {
        build_tree(srcs, N, max_leaf, P.S, nthreads);
        build_tree(tgts, M, max_leaf, P.T, nthreads);
    }
    clk.lap(tree_builder ? "octrees (device)" : "octrees (host)");
    for (int d = 0; d < 3; d++)
        if (!(P.S.hi[d] > P.S.lo[d]) || !(P.T.hi[d] > P.T.lo[d]))
            return "degenerate bounding box (all points share a coordinate plane at the origin)";
    P.normals.clear();
    if (normals) // onOctreeReady: normals in source Morton order (double_layer_helmholtz_ifgf.hpp:29-35)
    {
        P.normals.resize(3 * size_t(N));
        for (int64_t i = 0; i < N; i++)
            for (int d = 0; d < 3; d++)
                P.normals[3 * i + d] = normals[3 * P.S.perm[i] + d];
    }
    if (dev.pairs && !cutoff)
    {
        std::string e = dev.pairs(P);
        if (!e.empty())
            return e;
        clk.lap("interaction lists (device)");
    }
    else
    {
        build_interactions(P, cutoff);
        clk.lap("interaction lists (host)");
    }
    if (P.K.tol > 0)
        estimate_orders(P, P.K.tol, P.base_order);
    else
        P.base_order[0] = std::max(P.base_order[0] - 3, 2);
    P.K.orders(P.base_order, 1, P.ho_order);
    P.K.orders(P.base_order, 0, P.lo_order);
    for (int d = 0; d < 3; d++)
    {
        if (P.lo_order[d] < 1 || P.ho_order[d] < 1)
            return "interpolation order below 1 (the fine order of this kernel is order-3)";
        P.factor[d] = P.K.fine_factor();
    }
    clk.lap("orders");
    build_cones(P, nthreads, cone_marker);
    clk.lap(cone_marker ? "cones (device marking)" : "cones (host marking)");
    return "";
}
// ---------------------------------------------------------------------------------------
// several geometries in one plan
// ---------------------------------------------------------------------------------------
namespace
{
// new ids of every part's boxes: level-major, parts in order within a level
std::vector<std::vector<int32_t>> merged_ids(const std::vector<const Tree*>& trees, int levels,
                                             std::vector<int64_t>& level_off)
{
    const size_t np = trees.size();
    level_off.assign(levels + 1, 0);
    std::vector<std::vector<int32_t>> ids(np);
    int64_t next = 0;
    for (int l = 0; l < levels; l++)
    {
        level_off[l] = next;
        for (size_t g = 0; g < np; g++)
        {
            const Tree& T = *trees[g];
            if (ids[g].empty())
                ids[g].assign(T.num_boxes(), -1);
            if (l >= T.levels)
                continue;
            for (int64_t b = T.level_off[l]; b < T.level_off[l + 1]; b++)
                ids[g][b] = int32_t(next++);
        }
    }
    level_off[levels] = next;
    return ids;
}

void merge_tree(const std::vector<const Tree*>& trees, const std::vector<std::vector<int32_t>>& ids,
                const std::vector<int64_t>& level_off, int levels, Tree& M)
{
    const size_t np = trees.size();
    M = Tree();
    M.levels = levels;
    M.level_off = level_off;
    const int64_t nb = level_off[levels];
    M.begin.assign(nb, 0);
    M.end.assign(nb, 0);
    M.parent.assign(nb, -1);
    M.child.assign(8 * nb, -1);
    M.leaf.assign(nb, 0);
    M.blo.assign(3 * nb, 0);
    M.bhi.assign(3 * nb, 0);
    M.ctr.assign(3 * nb, 0);
    M.side.assign(nb, 0);
    int64_t poff = 0;
    for (size_t g = 0; g < np; g++)
    {
        const Tree& T = *trees[g];
        for (int64_t b = 0; b < T.num_boxes(); b++)
        {
            const int64_t m = ids[g][b];
            M.begin[m] = T.begin[b] + poff;
            M.end[m] = T.end[b] + poff;
            M.parent[m] = T.parent[b] < 0 ? -1 : ids[g][T.parent[b]];
            for (int c = 0; c < 8; c++)
                M.child[8 * m + c] = T.child[8 * b + c] < 0 ? -1 : ids[g][T.child[8 * b + c]];
            M.leaf[m] = T.leaf[b];
            for (int d = 0; d < 3; d++)
            {
                M.blo[3 * m + d] = T.blo[3 * b + d];
                M.bhi[3 * m + d] = T.bhi[3 * b + d];
                M.ctr[3 * m + d] = T.ctr[3 * b + d];
            }
            M.side[m] = T.side[b];
        }
        M.pts.insert(M.pts.end(), T.pts.begin(), T.pts.end());
        for (int64_t i = 0; i < T.n; i++)
            M.perm.push_back(T.perm[i] + poff);
        poff += T.n;
    }
    M.n = poff;
    for (int d = 0; d < 3; d++)
    {
        M.lo[d] = trees[0]->lo[d];
        M.hi[d] = trees[0]->hi[d];
    }
}
} // namespace

std::string merge_plans(const std::vector<const Plan*>& parts, Plan& M)
{
    if (parts.empty())
        return "no plans to merge";
    const Plan& P0 = *parts[0];
    for (const Plan* p : parts)
    {
        if (p->K.kind != P0.K.kind || p->K.dx != P0.K.dx)
            return "merged plans need one kernel family (and one derivative direction)";
        for (int d = 0; d < 3; d++)
            if (p->ho_order[d] != P0.ho_order[d] || p->lo_order[d] != P0.lo_order[d] ||
                p->factor[d] != P0.factor[d])
                return "merged plans need the same interpolation orders";
        if (p->normals.empty() != P0.normals.empty())
            return "merged plans need normals in all parts or in none";
        if (!p->up_parts.empty())
            return "plans to merge are single geometries";
    }
    const size_t np = parts.size();
    M = Plan();
    M.K = P0.K;
    for (int d = 0; d < 3; d++)
    {
        M.base_els[d] = P0.base_els[d];
        M.base_order[d] = P0.base_order[d];
        M.ho_order[d] = P0.ho_order[d];
        M.lo_order[d] = P0.lo_order[d];
        M.factor[d] = P0.factor[d];
    }
    int ls = 0, lt = 0;
    std::vector<const Tree*> st, tt;
    for (const Plan* p : parts)
    {
        ls = std::max(ls, p->S.levels);
        lt = std::max(lt, p->T.levels);
        st.push_back(&p->S);
        tt.push_back(&p->T);
    }
    std::vector<int64_t> loff_s, loff_t;
    const auto sid = merged_ids(st, ls, loff_s);
    const auto tid = merged_ids(tt, lt, loff_t);
    merge_tree(st, sid, loff_s, ls, M.S);
    merge_tree(tt, tid, loff_t, lt, M.T);
    const int64_t nb = M.S.num_boxes();
    M.part_src_off.push_back(0);
    M.part_tgt_off.push_back(0);
    for (const Plan* p : parts)
    {
        M.part_src_off.push_back(M.part_src_off.back() + p->S.n);
        M.part_tgt_off.push_back(M.part_tgt_off.back() + p->T.n);
    }
    // pairs, sorted by target box
    for (size_t g = 0; g < np; g++)
    {
        for (const Pair& q : parts[g]->far)
            M.far.push_back({sid[g][q.src], tid[g][q.tgt]});
        for (const Pair& q : parts[g]->near)
            M.near.push_back({sid[g][q.src], tid[g][q.tgt]});
        M.n_tofar += parts[g]->n_tofar;
        M.normals.insert(M.normals.end(), parts[g]->normals.begin(), parts[g]->normals.end());
    }
    auto by_tgt = [](const Pair& a, const Pair& b) { return a.tgt < b.tgt; };
    std::stable_sort(M.far.begin(), M.far.end(), by_tgt);
    std::stable_sort(M.near.begin(), M.near.end(), by_tgt);
    // per source box data in the new order
    std::vector<std::pair<int32_t, int64_t>> origin(nb); // new box -> (part, old box)
    for (size_t g = 0; g < np; g++)
        for (int64_t b = 0; b < int64_t(sid[g].size()); b++)
            origin[sid[g][b]] = {int32_t(g), b};
    M.has_far_anc.resize(nb);
    M.parent_has_far.resize(nb);
    M.dom_lo.resize(3 * nb);
    M.dom_hi.resize(3 * nb);
    M.nc.resize(3 * nb);
    M.nf.resize(3 * nb);
    M.hc.resize(3 * nb);
    M.hf.resize(3 * nb);
    M.box_k.resize(3 * nb);
    M.box_up_part.assign(nb, -1);
    M.table_off.assign(nb + 1, 0);
    M.cone_off.assign(nb + 1, 0);
    // up tables: one part per (geometry, valid level)
    std::vector<std::vector<int32_t>> part_of(np);
    for (size_t g = 0; g < np; g++)
    {
        part_of[g].assign(parts[g]->up.size(), -1);
        for (size_t l = 0; l < parts[g]->up.size(); l++)
            if (parts[g]->up[l].valid)
            {
                part_of[g][l] = int32_t(M.up_parts.size());
                M.up_parts.push_back({int(l), parts[g]->up[l]});
            }
    }
    std::vector<std::vector<int32_t>> cone_map(np);
    for (size_t g = 0; g < np; g++)
        cone_map[g].assign(parts[g]->num_cones(), -1);
    for (int64_t m = 0; m < nb; m++)
    {
        const Plan& P = *parts[origin[m].first];
        const int64_t b = origin[m].second;
        const int g = origin[m].first;
        M.has_far_anc[m] = P.has_far_anc[b];
        M.parent_has_far[m] = P.parent_has_far[b];
        for (int d = 0; d < 3; d++)
        {
            M.dom_lo[3 * m + d] = P.dom_lo[3 * b + d];
            M.dom_hi[3 * m + d] = P.dom_hi[3 * b + d];
            M.nc[3 * m + d] = P.nc[3 * b + d];
            M.nf[3 * m + d] = P.nf[3 * b + d];
            M.hc[3 * m + d] = P.hc[3 * b + d];
            M.hf[3 * m + d] = P.hf[3 * b + d];
        }
        M.box_k[3 * m] = P.K.k.real();
        M.box_k[3 * m + 1] = P.K.k.imag();
        M.box_k[3 * m + 2] = P.K.kappa;
        const int l = P.S.level_of(b);
        if (l < int(part_of[g].size()))
            M.box_up_part[m] = part_of[g][l];
        M.table_off[m + 1] = M.table_off[m] + (P.table_off[b + 1] - P.table_off[b]);
        for (int64_t c = P.cone_off[b]; c < P.cone_off[b + 1]; c++)
        {
            cone_map[g][c] = int32_t(M.cone_box.size());
            M.cone_box.push_back(int32_t(m));
            M.cone_cell.push_back(P.cone_cell[c]);
            M.cone_mask.push_back(P.cone_mask[c]);
            M.cone_fine.push_back(M.n_fine);
            M.n_fine += __builtin_popcount(P.cone_mask[c]);
        }
        M.cone_off[m + 1] = int64_t(M.cone_box.size());
    }
    M.table.assign(M.table_off[nb], -1);
    for (int64_t m = 0; m < nb; m++)
    {
        const Plan& P = *parts[origin[m].first];
        const int64_t b = origin[m].second;
        const int g = origin[m].first;
        for (int64_t i = P.table_off[b]; i < P.table_off[b + 1]; i++)
            if (P.table[i] >= 0)
                M.table[M.table_off[m] + (i - P.table_off[b])] = cone_map[g][P.table[i]];
    }
    // active fine cone ids per box (introspection), in the new box order
    M.fine_off.assign(nb + 1, 0);
    for (int64_t m = 0; m < nb; m++)
    {
        const Plan& P = *parts[origin[m].first];
        const int64_t b = origin[m].second;
        if (int64_t(P.fine_off.size()) == P.S.num_boxes() + 1)
            M.fine_ids.insert(M.fine_ids.end(), P.fine_ids.begin() + P.fine_off[b],
                              P.fine_ids.begin() + P.fine_off[b + 1]);
        M.fine_off[m + 1] = int64_t(M.fine_ids.size());
    }
    for (const Plan* p : parts)
        M.n_fine_uncovered += p->n_fine_uncovered;
    M.up.assign(ls, UpLevel()); // tables live in up_parts
    return "";
}

} // namespace ifgf
