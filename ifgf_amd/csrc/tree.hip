// tree.hip -- the octrees of init built on the device (octree.hpp:202-238, 1106-1199;
// zorder_less.hpp:212-248; util.hpp:14-71).
//
//   k_minmax           bounding box (first occurrence wins among equal extremes, as the
//                      sequential std::min / std::max scan of the host does)
//   k_keys             126-bit Morton keys, bit for bit the host morton_key
//   radix::sort        hand-written LSD radix sort (radix_sort.hpp) of (key hi, key lo,
//                      index) records: 16 stable passes of 8-bit digits; ties of the whole
//                      key keep index order
//   k_ties             each run of equal keys ordered by the exact comparison (stable
//                      insertion sort = the host's stable_sort), runs up to TIE_RUN_MAX
//   k_gather_pts       the Morton-sorted points
//   k_split            per level, one workgroup per box with more than max_leaf points: the
//                      greedy scan of the closed child boxes as a chunked block-wide search
//                      for the first point outside child j
//
// The O(boxes) bookkeeping of each level (box ids, child boxes, leaf flags) is the host's
// grow_tree, shared with the host builder, so both builders number boxes identically.
#include <hip/hip_runtime.h>

#include <algorithm>
#include <climits>
#include <cstring>
#include <stdexcept>
#include <string>
#include <vector>

#include "device.hpp"
#include "hip_util.hpp"
#include "plan.hpp"
#include "radix_sort.hpp"

#pragma clang fp contract(off)

namespace ifgf
{
namespace
{
constexpr int KEY_BITS = 40;
constexpr int MM_THREADS = 256;
constexpr int RS_PASSES = 16; // 8-bit digits over 128 bits (the top two are zero)
constexpr int64_t TIE_RUN_MAX = 1024;
constexpr int SPLIT_THREADS = 1024;
constexpr int SPLIT_PER = 4; // points per thread per chunk

struct Norm
{
    double ctr[3], diam[3];
};

// a point's Morton key and caller index (the radix sort's record)
struct KeyRec
{
    uint64_t hi, lo;
    int32_t id, pad;
};

// digit of pass p: byte p of lo for p < 8, byte p - 8 of hi after
struct KeyDigit
{
    int use_hi, shift;
    __device__ __forceinline__ uint32_t operator()(const KeyRec& r) const
    {
        return uint32_t((use_hi ? r.hi : r.lo) >> shift) & 255u;
    }
};

// ---------------------------------------------------------------------------------------
// bounding box
// ---------------------------------------------------------------------------------------
struct Ext
{
    double v;
    int64_t i;
};

// v (index i) replaces the current minimum when smaller, or equal with a smaller index
__device__ __forceinline__ void take_min(Ext& a, const Ext& b)
{
    if (b.v < a.v || (b.v == a.v && b.i < a.i))
        a = b;
}
__device__ __forceinline__ void take_max(Ext& a, const Ext& b)
{
    if (a.v < b.v || (b.v == a.v && b.i < a.i))
        a = b;
}

// per block: the first minimum and the first maximum of each coordinate with its index
__global__ __launch_bounds__(MM_THREADS) void k_minmax(const double* __restrict__ pts, int64_t n,
                                                      double* __restrict__ part_v,
                                                      int64_t* __restrict__ part_i)
{
    Ext mn[3], mx[3];
    for (int d = 0; d < 3; d++)
    {
        mn[d] = {0.0, INT64_MAX};
        mx[d] = {0.0, INT64_MAX};
    }
    const int64_t stride = int64_t(gridDim.x) * MM_THREADS;
    for (int64_t i = int64_t(blockIdx.x) * MM_THREADS + threadIdx.x; i < n; i += stride)
        for (int d = 0; d < 3; d++)
        {
            const Ext c = {pts[3 * i + d], i};
            if (mn[d].i == INT64_MAX)
            {
                mn[d] = c;
                mx[d] = c;
            }
            else
            {
                take_min(mn[d], c);
                take_max(mx[d], c);
            }
        }
    __shared__ double sv[6][MM_THREADS];
    __shared__ int64_t si[6][MM_THREADS];
    for (int d = 0; d < 3; d++)
    {
        sv[d][threadIdx.x] = mn[d].v;
        si[d][threadIdx.x] = mn[d].i;
        sv[3 + d][threadIdx.x] = mx[d].v;
        si[3 + d][threadIdx.x] = mx[d].i;
    }
    __syncthreads();
    for (int h = MM_THREADS / 2; h > 0; h >>= 1)
    {
        if (int(threadIdx.x) < h)
            for (int c = 0; c < 6; c++)
            {
                Ext a = {sv[c][threadIdx.x], si[c][threadIdx.x]};
                const Ext b = {sv[c][threadIdx.x + h], si[c][threadIdx.x + h]};
                if (b.i != INT64_MAX)
                {
                    if (a.i == INT64_MAX)
                        a = b;
                    else if (c < 3)
                        take_min(a, b);
                    else
                        take_max(a, b);
                }
                sv[c][threadIdx.x] = a.v;
                si[c][threadIdx.x] = a.i;
            }
        __syncthreads();
    }
    if (threadIdx.x < 6)
    {
        part_v[6 * int64_t(blockIdx.x) + threadIdx.x] = sv[threadIdx.x][0];
        part_i[6 * int64_t(blockIdx.x) + threadIdx.x] = si[threadIdx.x][0];
    }
}

// ---------------------------------------------------------------------------------------
// Morton keys and the exact comparison (the host's morton_key / morton_less_exact)
// ---------------------------------------------------------------------------------------
__device__ void dev_key(const double* p, const Norm& nm, uint64_t& khi, uint64_t& klo)
{
    uint64_t sign[3], fx[3];
    for (int d = 0; d < 3; d++)
    {
        const double v = (p[d] - nm.ctr[d]) / nm.diam[d];
        const bool neg = v < 0.0;
        const double m = fabs(v);
        const uint64_t f = uint64_t(ldexp(m, KEY_BITS));
        const uint64_t nz = (m != 0.0) ? 1 : 0;
        uint64_t word = (nz << KEY_BITS) | (f & ((uint64_t(1) << KEY_BITS) - 1));
        if (neg)
            word = ~word & ((uint64_t(1) << (KEY_BITS + 1)) - 1);
        sign[d] = neg ? 0 : 1;
        fx[d] = word;
    }
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
    khi = hi;
    klo = lo;
}

__device__ __forceinline__ int dev_fexp(uint64_t bits)
{
    uint64_t u = bits & 0x7fffffffffffffffull;
    if (u == 0 || u >= 0x7ff0000000000000ull)
        return 0;
    u >>= 52;
    return (u == 0) ? -1022 : int(u) - 1023;
}

__device__ __forceinline__ int dev_xor_msb(double p, double q)
{
    if (p == q || p == -q)
        return INT_MIN;
    const uint64_t a = uint64_t(__double_as_longlong(p)), b = uint64_t(__double_as_longlong(q));
    const int ea = dev_fexp(a), eb = dev_fexp(b);
    if (ea == eb)
    {
        const uint64_t x = (a ^ b) & 0x000fffffffffffffull;
        if (x == 0)
            return ea;
        return ea + (63 - __builtin_clzll(x)) - 52;
    }
    return ea > eb ? ea : eb;
}

__device__ bool dev_less_exact(const double* px, const double* qx, const Norm& nm)
{
    double p[3], q[3];
    for (int i = 0; i < 3; i++)
    {
        p[i] = (px[i] - nm.ctr[i]) / nm.diam[i];
        q[i] = (qx[i] - nm.ctr[i]) / nm.diam[i];
    }
    int best = INT_MIN, k = 0;
    for (int j = 2; j >= 0; j--)
    {
        if ((p[j] < 0.0) != (q[j] < 0.0))
            return p[j] < q[j];
        const int y = dev_xor_msb(p[j], q[j]);
        if (best < y)
        {
            best = y;
            k = j;
        }
    }
    return p[k] < q[k];
}

__global__ __launch_bounds__(256) void k_keys(const double* __restrict__ pts, int64_t n, Norm nm,
                                             KeyRec* __restrict__ keys)
{
    const int64_t i = int64_t(blockIdx.x) * 256 + threadIdx.x;
    if (i >= n)
        return;
    const double p[3] = {pts[3 * i], pts[3 * i + 1], pts[3 * i + 2]};
    KeyRec k;
    dev_key(p, nm, k.hi, k.lo);
    k.id = int32_t(i);
    k.pad = 0;
    keys[i] = k;
}

// the sorted indices form a permutation of 0..n-1 (seen[] starts at zero); bad is set otherwise
__global__ __launch_bounds__(256) void k_check_perm(const KeyRec* __restrict__ keys, int64_t n,
                                                   uint32_t* __restrict__ seen,
                                                   int* __restrict__ bad)
{
    const int64_t i = int64_t(blockIdx.x) * 256 + threadIdx.x;
    if (i >= n)
        return;
    const int32_t s = keys[i].id;
    if (s < 0 || int64_t(s) >= n || atomicAdd(&seen[s], 1u) != 0)
        atomicOr(bad, 1);
}

__device__ __forceinline__ bool same_key(const KeyRec* k, int64_t a, int64_t b)
{
    return k[a].hi == k[b].hi && k[a].lo == k[b].lo;
}

// one thread per run of equal keys (at its first element): stable insertion sort of the run
// by the exact comparison; runs longer than TIE_RUN_MAX are left to the host (flag)
__global__ __launch_bounds__(256) void k_ties(KeyRec* __restrict__ keys, int64_t n,
                                             const double* __restrict__ pts, Norm nm,
                                             int* __restrict__ long_run)
{
    const int64_t i = int64_t(blockIdx.x) * 256 + threadIdx.x;
    if (i + 1 >= n || !same_key(keys, i, i + 1) || (i > 0 && same_key(keys, i - 1, i)))
        return;
    int64_t j = i + 1;
    while (j < n && same_key(keys, i, j))
        j++;
    if (j - i > TIE_RUN_MAX)
    {
        atomicOr(long_run, 1);
        return;
    }
    for (int64_t a = i + 1; a < j; a++)
    {
        const int32_t x = keys[a].id;
        int64_t b = a;
        while (b > i &&
               dev_less_exact(pts + 3 * int64_t(x), pts + 3 * int64_t(keys[b - 1].id), nm))
        {
            keys[b].id = keys[b - 1].id;
            b--;
        }
        keys[b].id = x;
    }
}

__global__ __launch_bounds__(256) void k_gather_pts(const double* __restrict__ pts,
                                                   const KeyRec* __restrict__ keys, int64_t n,
                                                   double* __restrict__ spts)
{
    const int64_t i = int64_t(blockIdx.x) * 256 + threadIdx.x;
    if (i >= n)
        return;
    const int64_t s = keys[i].id;
    for (int d = 0; d < 3; d++)
        spts[3 * i + d] = pts[3 * s + d];
}

// ---------------------------------------------------------------------------------------
// child ranges of one level
// ---------------------------------------------------------------------------------------
__device__ __forceinline__ void dev_child_box(const double* blo, const double* bhi, int j,
                                              double lo[3], double hi[3])
{
    for (int d = 0; d < 3; d++)
    {
        const double size = 0.5 * (bhi[d] - blo[d]);
        const double tup = double((j >> d) & 1);
        lo[d] = blo[d] + size * tup;
        hi[d] = lo[d] + size;
    }
}

// cr: 16 values per box (ChildRange: begin of the 8 children, then their ends)
__global__ __launch_bounds__(SPLIT_THREADS) void k_split(
    const double* __restrict__ spts, const int64_t* __restrict__ beg,
    const int64_t* __restrict__ end, const double* __restrict__ blo,
    const double* __restrict__ bhi, int64_t max_leaf, int64_t* __restrict__ cr,
    uint8_t* __restrict__ split)
{
    const int64_t q = blockIdx.x;
    const int64_t b = beg[q], e = end[q];
    if (e - b <= max_leaf)
    {
        if (threadIdx.x == 0)
            split[q] = 0;
        return;
    }
    __shared__ int64_t wmin[SPLIT_THREADS / 64];
    const int lane = threadIdx.x & 63, w = threadIdx.x >> 6;
    const double pl[3] = {blo[3 * q], blo[3 * q + 1], blo[3 * q + 2]};
    const double ph[3] = {bhi[3 * q], bhi[3 * q + 1], bhi[3 * q + 2]};
    int64_t pi = b;
    for (int j = 0; j < 8; j++)
    {
        double cl[3], ch[3];
        dev_child_box(pl, ph, j, cl, ch);
        int64_t ep = e;
        for (int64_t c = pi; c < e; c += int64_t(SPLIT_THREADS) * SPLIT_PER)
        {
            int64_t fail = INT64_MAX;
            for (int k = SPLIT_PER - 1; k >= 0; k--)
            {
                const int64_t i = c + int64_t(k) * SPLIT_THREADS + threadIdx.x;
                if (i < e)
                {
                    const double* p = spts + 3 * i;
                    if (!(cl[0] <= p[0] && cl[1] <= p[1] && cl[2] <= p[2] && p[0] <= ch[0] &&
                          p[1] <= ch[1] && p[2] <= ch[2]))
                        fail = i;
                }
            }
            for (int off = 32; off > 0; off >>= 1)
            {
                const int64_t o = __shfl_xor(fail, off, 64);
                fail = o < fail ? o : fail;
            }
            if (lane == 0)
                wmin[w] = fail;
            __syncthreads();
            int64_t f = wmin[0];
            for (int v = 1; v < SPLIT_THREADS / 64; v++)
                f = wmin[v] < f ? wmin[v] : f;
            __syncthreads();
            if (f != INT64_MAX)
            {
                ep = f;
                break;
            }
        }
        if (threadIdx.x == 0)
        {
            cr[16 * q + j] = pi;
            cr[16 * q + 8 + j] = ep;
        }
        pi = ep;
    }
    if (threadIdx.x == 0)
        split[q] = 1;
}

int blocks_for(int64_t n, int per) { return int((n + per - 1) / per); }

} // namespace

std::string device_build_tree(const double* pts, int64_t n, int64_t max_leaf, Tree& T)
{
    try
    {
        if (n <= 0)
            return "device tree: empty point set";
        if (n > int64_t(INT32_MAX))
            return "device tree: more than 2^31 - 1 points";
        DBuf d_pts, d_spts, pv, pi, counts, flag;
        h2d(d_pts, pts, size_t(3 * n));

        // bounding box: per-block first extremes, finished here with the same rule
        const int mm_blocks = std::min(1024, std::max(1, blocks_for(n, MM_THREADS * 8)));
        pv.reserve(sizeof(double) * 6 * mm_blocks);
        pi.reserve(sizeof(int64_t) * 6 * mm_blocks);
        k_minmax<<<mm_blocks, MM_THREADS>>>(d_pts.as<double>(), n, pv.as<double>(),
                                            pi.as<int64_t>());
        HIP_OK(hipGetLastError());
        std::vector<double> hv(6 * mm_blocks);
        std::vector<int64_t> hidx(6 * mm_blocks);
        HIP_OK(hipMemcpy(hv.data(), pv.p, hv.size() * sizeof(double), hipMemcpyDeviceToHost));
        HIP_OK(hipMemcpy(hidx.data(), pi.p, hidx.size() * sizeof(int64_t),
                              hipMemcpyDeviceToHost));
        double mn[3], mx[3];
        for (int c = 0; c < 6; c++)
        {
            double v = 0;
            int64_t at = INT64_MAX;
            for (int bk = 0; bk < mm_blocks; bk++)
            {
                const double x = hv[6 * bk + c];
                const int64_t i = hidx[6 * bk + c];
                if (i == INT64_MAX)
                    continue;
                const bool better = at == INT64_MAX ||
                                    (c < 3 ? (x < v || (x == v && i < at))
                                           : (v < x || (x == v && i < at)));
                if (better)
                {
                    v = x;
                    at = i;
                }
            }
            (c < 3 ? mn[c] : mx[c - 3]) = v;
        }
        Norm nm;
        tree_root(T, n, mn, mx, nm.ctr, nm.diam);

        // keys and the radix sort
        DBuf keys[2];
        for (int b2 = 0; b2 < 2; b2++)
            keys[b2].reserve(sizeof(KeyRec) * n);
        k_keys<<<blocks_for(n, 256), 256>>>(d_pts.as<double>(), n, nm, keys[0].as<KeyRec>());
        HIP_OK(hipGetLastError());
        counts.reserve(sizeof(uint32_t) * radix::counts_needed(n));
        KeyRec* sorted = radix::sort(
            keys[0].as<KeyRec>(), keys[1].as<KeyRec>(), n, RS_PASSES,
            [](int pass) { return KeyDigit{pass >= 8 ? 1 : 0, (pass & 7) * 8}; },
            counts.as<uint32_t>(), nullptr);
        flag.reserve(sizeof(int));
        HIP_OK(hipMemset(flag.p, 0, sizeof(int)));
        {
            // the indices index the points below: they must be a permutation
            DBuf seen;
            seen.reserve(sizeof(uint32_t) * n);
            HIP_OK(hipMemset(seen.p, 0, sizeof(uint32_t) * n));
            k_check_perm<<<blocks_for(n, 256), 256>>>(sorted, n, seen.as<uint32_t>(),
                                                      flag.as<int>());
            HIP_OK(hipGetLastError());
            int bad = 0;
            HIP_OK(hipMemcpy(&bad, flag.p, sizeof(int), hipMemcpyDeviceToHost));
            if (bad)
                return "device tree: the radix sort did not produce a permutation";
        }
        k_ties<<<blocks_for(n, 256), 256>>>(sorted, n, d_pts.as<double>(), nm, flag.as<int>());
        HIP_OK(hipGetLastError());
        int long_run = 0;
        HIP_OK(hipMemcpy(&long_run, flag.p, sizeof(int), hipMemcpyDeviceToHost));
        std::vector<KeyRec> hk(n);
        HIP_OK(hipMemcpy(hk.data(), sorted, sizeof(KeyRec) * n, hipMemcpyDeviceToHost));
        if (long_run)
        {
            // runs of more than TIE_RUN_MAX equal keys (coincident points): the host's stable
            // sort of each such run
            for (int64_t i = 0; i < n;)
            {
                int64_t j = i + 1;
                while (j < n && hk[j].hi == hk[i].hi && hk[j].lo == hk[i].lo)
                    j++;
                if (j - i > TIE_RUN_MAX)
                    std::stable_sort(hk.begin() + i, hk.begin() + j,
                                     [&](const KeyRec& x, const KeyRec& y)
                                     {
                                         return morton_less_exact(pts + 3 * int64_t(x.id),
                                                                  pts + 3 * int64_t(y.id), nm.ctr,
                                                                  nm.diam);
                                     });
                i = j;
            }
            HIP_OK(hipMemcpy(sorted, hk.data(), sizeof(KeyRec) * n, hipMemcpyHostToDevice));
        }
        d_spts.reserve(sizeof(double) * 3 * n);
        k_gather_pts<<<blocks_for(n, 256), 256>>>(d_pts.as<double>(), sorted, n,
                                                  d_spts.as<double>());
        HIP_OK(hipGetLastError());
        T.perm.resize(n);
        for (int64_t i = 0; i < n; i++)
            T.perm[i] = hk[i].id;
        hk.clear();
        hk.shrink_to_fit();
        for (int b2 = 0; b2 < 2; b2++)
            keys[b2].release();
        T.pts.resize(3 * size_t(n));
        HIP_OK(hipMemcpy(T.pts.data(), d_spts.p, sizeof(double) * 3 * n,
                              hipMemcpyDeviceToHost));

        // levels
        DBuf l_beg, l_end, l_lo, l_hi, l_cr, l_split;
        grow_tree(T, max_leaf,
                  [&](const Tree& Tc, int64_t lb, int64_t le, int64_t leaf,
                      std::vector<ChildRange>& cr, std::vector<uint8_t>& split)
                  {
                      const int64_t nl = le - lb;
                      h2d(l_beg, &Tc.begin[lb], size_t(nl));
                      h2d(l_end, &Tc.end[lb], size_t(nl));
                      h2d(l_lo, &Tc.blo[3 * lb], size_t(3 * nl));
                      h2d(l_hi, &Tc.bhi[3 * lb], size_t(3 * nl));
                      l_cr.reserve(sizeof(ChildRange) * nl);
                      l_split.reserve(size_t(nl));
                      static_assert(sizeof(ChildRange) == 16 * sizeof(int64_t), "layout");
                      k_split<<<int(nl), SPLIT_THREADS>>>(
                          d_spts.as<double>(), l_beg.as<int64_t>(), l_end.as<int64_t>(),
                          l_lo.as<double>(), l_hi.as<double>(), leaf, l_cr.as<int64_t>(),
                          l_split.as<uint8_t>());
                      HIP_OK(hipGetLastError());
                      HIP_OK(hipMemcpy(cr.data(), l_cr.p, sizeof(ChildRange) * nl,
                                            hipMemcpyDeviceToHost));
                      HIP_OK(hipMemcpy(split.data(), l_split.p, size_t(nl),
                                            hipMemcpyDeviceToHost));
                  });
    }
    catch (const std::exception& ex)
    {
        return std::string("device tree: ") + ex.what();
    }
    return "";
}

} // namespace ifgf
