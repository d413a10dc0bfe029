// radix_sort.hpp -- hand-written LSD radix sort of records on the device, 8-bit digits.
//
// One pass = three launches: k_rs_hist (digit counts of every 2048-record tile, stored
// digit-major), k_rs_scan (exclusive prefix sum of those counts: where each (digit, tile)
// starts in the output), k_rs_scatter (stable scatter of each tile).  In the scatter, wave w
// of a tile owns records [w * 512, (w + 1) * 512) taken 64 at a time in order; a record's rank
// among the equal digits of its wave is the wave's running count of that digit (LDS) plus the
// lanes below it with the same digit (eight ballots); per-wave prefix sums then place the
// waves' runs one after another.  Every pass is stable, so the passes together sort by the
// digits of the last pass first, ties keeping the input order.
#pragma once
#include <hip/hip_runtime.h>

#include <algorithm>
#include <cstdint>
#include <stdexcept>
#include <string>

namespace ifgf
{
namespace radix
{
constexpr int THREADS = 256; // four waves
constexpr int ITEMS = 8;     // records per lane
constexpr int WAVE_SEG = 64 * ITEMS;
constexpr int TILE = THREADS * ITEMS;

// Dig: a functor giving the 8-bit digit (0..255) of a record for the current pass
template <class Rec, class Dig>
__global__ __launch_bounds__(THREADS) void k_rs_hist(const Rec* __restrict__ in, int64_t n,
                                                    Dig dig, uint32_t* __restrict__ counts,
                                                    int nblk)
{
    __shared__ uint32_t h[256];
    h[threadIdx.x] = 0;
    __syncthreads();
    const int64_t base = int64_t(blockIdx.x) * TILE;
    for (int r = 0; r < ITEMS; r++)
    {
        const int64_t i = base + int64_t(r) * THREADS + threadIdx.x;
        if (i < n)
            atomicAdd(&h[dig(in[i])], 1u);
    }
    __syncthreads();
    counts[int64_t(threadIdx.x) * nblk + blockIdx.x] = h[threadIdx.x];
}

// exclusive prefix sum of m counts in place (one workgroup; a contiguous chunk per thread)
template <int UNUSED = 0>
__global__ __launch_bounds__(1024) void k_rs_scan(uint32_t* __restrict__ c, int64_t m)
{
    __shared__ uint32_t s[1024];
    const int t = threadIdx.x;
    const int64_t per = (m + 1023) / 1024;
    const int64_t a = std::min<int64_t>(m, int64_t(t) * per), b = std::min<int64_t>(m, a + per);
    uint32_t sum = 0;
    for (int64_t i = a; i < b; i++)
        sum += c[i];
    s[t] = sum;
    __syncthreads();
    for (int off = 1; off < 1024; off <<= 1)
    {
        const uint32_t v = t >= off ? s[t - off] : 0u;
        __syncthreads();
        s[t] += v;
        __syncthreads();
    }
    uint32_t run = s[t] - sum;
    for (int64_t i = a; i < b; i++)
    {
        const uint32_t v = c[i];
        c[i] = run;
        run += v;
    }
}

template <class Rec, class Dig>
__global__ __launch_bounds__(THREADS) void k_rs_scatter(const Rec* __restrict__ in,
                                                       Rec* __restrict__ out, int64_t n, Dig dig,
                                                       const uint32_t* __restrict__ offs,
                                                       int nblk)
{
    __shared__ uint32_t cnt[THREADS / 64][256];
    __shared__ uint32_t gofs[256];
    const int lane = threadIdx.x & 63, w = threadIdx.x >> 6;
    for (int i = threadIdx.x; i < (THREADS / 64) * 256; i += THREADS)
        (&cnt[0][0])[i] = 0;
    gofs[threadIdx.x] = offs[int64_t(threadIdx.x) * nblk + blockIdx.x];
    __syncthreads();
    const int64_t seg = int64_t(blockIdx.x) * TILE + int64_t(w) * WAVE_SEG;
    const uint64_t below_me = (uint64_t(1) << lane) - 1;
    Rec rec[ITEMS];
    uint32_t dg[ITEMS], rk[ITEMS];
#pragma unroll
    for (int r = 0; r < ITEMS; r++)
    {
        const int64_t i = seg + int64_t(r) * 64 + lane;
        const bool valid = i < n;
        uint32_t d = 0;
        if (valid)
        {
            rec[r] = in[i];
            d = dig(rec[r]);
        }
        uint64_t peers = __ballot(valid);
#pragma unroll
        for (int b = 0; b < 8; b++)
        {
            const bool set = (d >> b) & 1u;
            const uint64_t m = __ballot(set);
            peers &= set ? m : ~m;
        }
        dg[r] = d;
        rk[r] = 0;
        if (valid)
        {
            const uint32_t below = uint32_t(__popcll(peers & below_me));
            const uint32_t base = cnt[w][d];
            rk[r] = base + below;
            if (below == 0)
                cnt[w][d] = base + uint32_t(__popcll(peers));
        }
    }
    __syncthreads();
    {
        const int d = threadIdx.x;
        uint32_t s = 0;
        for (int v = 0; v < THREADS / 64; v++)
        {
            const uint32_t t = cnt[v][d];
            cnt[v][d] = s;
            s += t;
        }
    }
    __syncthreads();
#pragma unroll
    for (int r = 0; r < ITEMS; r++)
    {
        const int64_t i = seg + int64_t(r) * 64 + lane;
        const int64_t pos = int64_t(gofs[dg[r]]) + cnt[w][dg[r]] + rk[r];
        if (i < n && pos < n)
            out[pos] = rec[r];
    }
}

// Sort n records of a by passes digits (digit_of_pass(p) gives the functor of pass p, least
// significant first), b as the ping-pong buffer and counts as scratch of counts_needed(n)
// words.  Returns the buffer holding the result (a or b).
inline int64_t counts_needed(int64_t n) { return int64_t(256) * ((n + TILE - 1) / TILE); }

template <class Rec, class DigitOfPass>
Rec* sort(Rec* a, Rec* b, int64_t n, int passes, const DigitOfPass& digit_of_pass,
          uint32_t* counts, hipStream_t st)
{
    if (n <= 1)
        return a;
    const int nblk = int((n + TILE - 1) / TILE);
    for (int p = 0; p < passes; p++)
    {
        const auto dig = digit_of_pass(p);
        k_rs_hist<<<nblk, THREADS, 0, st>>>(a, n, dig, counts, nblk);
        k_rs_scan<<<1, 1024, 0, st>>>(counts, int64_t(256) * nblk);
        k_rs_scatter<<<nblk, THREADS, 0, st>>>(a, b, n, dig, counts, nblk);
        const hipError_t e = hipGetLastError();
        if (e != hipSuccess)
            throw std::runtime_error(std::string("radix sort: ") + hipGetErrorString(e));
        std::swap(a, b);
    }
    return a;
}

} // namespace radix
} // namespace ifgf
