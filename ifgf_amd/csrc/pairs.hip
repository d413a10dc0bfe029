// pairs.hip -- the interaction lists of init on the device (octree.hpp:192-197, 255-349).
//
// The dual traversal breadth first, one launch per generation of (source box, target box)
// pairs: k_expand classifies each pair of the generation with the host's admissibility test
// (FP64, no contraction: extDist(target box, source centre) >= sqrt(3) H_src), appends far
// and near pairs to their lists and the refined pairs (children of the larger non-leaf side)
// to the next generation, slots taken by atomic counters.  The lists are then put in
// (target, source) order by the radix sort of radix_sort.hpp on 64-bit (target, source) keys,
// which is the order build_interactions sorts its lists into, so both builders agree exactly.
// cutOff callbacks are host functions: with one set, init uses build_interactions.
#include <hip/hip_runtime.h>

#include <algorithm>
#include <cmath>
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
struct TreeDev
{
    const double* ctr;   // 3 per box (source tree)
    const double* blo;   // 3 per box (target tree)
    const double* bhi;
    const double* side;
    const uint8_t* leaf;
    const int32_t* child; // 8 per box
};

// (target << 32) | source: one 64-bit word per pair, ordered as (target, source)
__device__ __forceinline__ uint64_t pair_key(int32_t src, int32_t tgt)
{
    return (uint64_t(uint32_t(tgt)) << 32) | uint64_t(uint32_t(src));
}

// counters: [0] far pairs, [1] near pairs, [2] next generation
__global__ __launch_bounds__(256) void k_expand(const uint64_t* __restrict__ cur, int64_t n,
                                               TreeDev S, TreeDev T, double eta,
                                               uint64_t* __restrict__ far_out,
                                               uint64_t* __restrict__ near_out,
                                               uint64_t* __restrict__ nxt,
                                               unsigned long long* __restrict__ counters)
{
    const int64_t i = int64_t(blockIdx.x) * 256 + threadIdx.x;
    if (i >= n)
        return;
    const uint64_t key = cur[i];
    const int32_t t = int32_t(key >> 32), s = int32_t(key & 0xffffffffu);
    double d2 = 0;
    for (int k = 0; k < 3; k++)
    {
        const double lo = T.blo[3 * t + k], hi = T.bhi[3 * t + k], p = S.ctr[3 * s + k];
        if (lo > p)
        {
            const double a = lo - p;
            d2 += a * a;
        }
        else if (p > hi)
        {
            const double a = p - hi;
            d2 += a * a;
        }
    }
    const double ed = sqrt(d2);
    if (ed >= eta * S.side[s])
    {
        far_out[atomicAdd(counters + 0, 1ull)] = key;
        return;
    }
    const bool sl = S.leaf[s], tl = T.leaf[t];
    if (sl && tl)
    {
        near_out[atomicAdd(counters + 1, 1ull)] = key;
        return;
    }
    bool src_side;
    if (sl)
        src_side = false;
    else if (tl)
        src_side = true;
    else
        src_side = S.side[s] > T.side[t];
    const int32_t* ch = src_side ? S.child + 8 * int64_t(s) : T.child + 8 * int64_t(t);
    int nc = 0;
    for (int j = 0; j < 8; j++)
        nc += ch[j] >= 0;
    uint64_t at = atomicAdd(counters + 2, (unsigned long long)nc);
    for (int j = 0; j < 8; j++)
    {
        const int32_t c = ch[j];
        if (c >= 0)
            nxt[at++] = src_side ? pair_key(c, t) : pair_key(s, c);
    }
}

struct U64Digit
{
    int shift;
    __device__ __forceinline__ uint32_t operator()(uint64_t k) const
    {
        return uint32_t(k >> shift) & 255u;
    }
};

int bits_for(int64_t v)
{
    int b = 0;
    while (b < 63 && (int64_t(1) << b) <= v)
        b++;
    return b;
}

// the sorted keys of one list back as pairs
void sort_list(DBuf& list, DBuf& tmp, DBuf& counts, int64_t n, int key_bits,
               std::vector<Pair>& out)
{
    out.resize(n);
    if (n == 0)
        return;
    tmp.reserve(sizeof(uint64_t) * n);
    counts.reserve(sizeof(uint32_t) * radix::counts_needed(n));
    const int passes = (key_bits + 7) / 8;
    uint64_t* sorted = radix::sort(list.as<uint64_t>(), tmp.as<uint64_t>(), n, passes,
                                   [](int pass) { return U64Digit{8 * pass}; },
                                   counts.as<uint32_t>(), nullptr);
    std::vector<uint64_t> keys(n);
    HIP_OK(hipMemcpy(keys.data(), sorted, sizeof(uint64_t) * n, hipMemcpyDeviceToHost));
    for (int64_t i = 0; i < n; i++)
        out[i] = {int32_t(keys[i] & 0xffffffffu), int32_t(keys[i] >> 32)};
}

} // namespace

std::string device_build_interactions(Plan& P)
{
    try
    {
        const Tree& S = P.S;
        const Tree& T = P.T;
        P.far.clear();
        P.near.clear();
        P.n_tofar = 0;
        DBuf s_ctr, s_side, s_leaf, s_child, t_lo, t_hi, t_side, t_leaf, t_child;
        h2d(s_ctr, S.ctr);
        h2d(s_side, S.side);
        h2d(s_leaf, S.leaf);
        h2d(s_child, S.child);
        h2d(t_lo, T.blo);
        h2d(t_hi, T.bhi);
        h2d(t_side, T.side);
        h2d(t_leaf, T.leaf);
        h2d(t_child, T.child);
        TreeDev SD{s_ctr.as<double>(), nullptr, nullptr, s_side.as<double>(),
                   s_leaf.as<uint8_t>(), s_child.as<int32_t>()};
        TreeDev TD{nullptr, t_lo.as<double>(), t_hi.as<double>(), t_side.as<double>(),
                   t_leaf.as<uint8_t>(), t_child.as<int32_t>()};
        const double eta = std::sqrt(3.0);
        DBuf gen[2], far_l, near_l, ctr;
        ctr.reserve(3 * sizeof(unsigned long long));
        HIP_OK(hipMemset(ctr.p, 0, 3 * sizeof(unsigned long long)));
        const uint64_t root = 0; // (target 0, source 0)
        gen[0].reserve(sizeof(uint64_t));
        HIP_OK(hipMemcpy(gen[0].p, &root, sizeof(uint64_t), hipMemcpyHostToDevice));
        int64_t n = 1, nfar = 0, nnear = 0;
        int cur = 0;
        while (n > 0)
        {
            // room for every pair of the generation in either list, 8 children each
            far_l.reserve(sizeof(uint64_t) * (nfar + n), sizeof(uint64_t) * nfar);
            near_l.reserve(sizeof(uint64_t) * (nnear + n), sizeof(uint64_t) * nnear);
            gen[1 - cur].reserve(sizeof(uint64_t) * 8 * n);
            HIP_OK(hipMemset(static_cast<unsigned long long*>(ctr.p) + 2, 0,
                             sizeof(unsigned long long)));
            k_expand<<<dim3(unsigned((n + 255) / 256)), dim3(256)>>>(
                gen[cur].as<uint64_t>(), n, SD, TD, eta, far_l.as<uint64_t>(),
                near_l.as<uint64_t>(), gen[1 - cur].as<uint64_t>(),
                ctr.as<unsigned long long>());
            HIP_OK(hipGetLastError());
            unsigned long long c[3];
            HIP_OK(hipMemcpy(c, ctr.p, sizeof(c), hipMemcpyDeviceToHost));
            nfar = int64_t(c[0]);
            nnear = int64_t(c[1]);
            n = int64_t(c[2]);
            cur = 1 - cur;
        }
        const int key_bits = 32 + bits_for(T.num_boxes());
        DBuf tmp, counts;
        sort_list(far_l, tmp, counts, nfar, key_bits, P.far);
        sort_list(near_l, tmp, counts, nnear, key_bits, P.near);
        interaction_flags(P);
    }
    catch (const std::exception& ex)
    {
        return std::string("device interaction lists: ") + ex.what();
    }
    return "";
}

} // namespace ifgf
