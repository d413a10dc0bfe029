// cones.hip -- cone activity of init marked on the device (octree.hpp:487-548, 620-700;
// util.hpp:156-171 cartToInterp2, cone_domain.hpp:232-256 elementForPoint).
//
// Per source level, after the host has laid out the level's grids and bitmaps:
//   k_mark_far     one workgroup per (source box, far target box) pair: every target point's
//                  fine and coarse cells of the source box's grids, or'ed into the bitmaps
//   k_mark_parent  one workgroup per box whose parent has far work: every high-order node of
//                  every active coarse cone of the parent, located in the box's grids (from
//                  the parent level's upward table where the host built one, otherwise by
//                  the FP64 cone map), or'ed into the bitmaps
// The arithmetic is the host marking's, operation for operation in FP64 without contraction
// (IEEE division and square root; atan2, sin and cos from the device library, which agree
// with the host's except within an ulp or two of a cell edge).  The host turns the bitmaps
// into cone lists exactly as it does its own.
#include <hip/hip_runtime.h>

#include <algorithm>
#include <memory>
#include <stdexcept>
#include <string>
#include <vector>

#include "device.hpp"
#include "hip_util.hpp"
#include "plan.hpp"

#pragma clang fp contract(off)

namespace ifgf
{
namespace
{
// one source box of the level: centre, side, the shared cone domain and both grids
struct MarkBox
{
    double ctr[3];
    double H;
    double lo[3], hi[3];
    double hf[3], hc[3];
    int32_t nf[3], nc[3];
    int64_t foff, coff; // first bitmap word of the box (fine, coarse)
};

// a box of the level whose parent's active coarse cones are marked into it
struct ParentWork
{
    int32_t q;       // box of the level (index from the level's first box)
    int32_t oct;     // its octant in the parent
    int64_t c0, c1;  // the parent's cones, as indices into the parent level's cone arrays
    // the parent's geometry (the per-box path without a table)
    double pctr[3];
    double pH;
    double plo[3], phc[3];
    int32_t pnc[3];
};

__device__ __forceinline__ void cart_to_interp_exact(const double x[3], const double xc[3],
                                                     double H, double out[3])
{
    const double p0 = x[0] - xc[0], p1 = x[1] - xc[1], p2 = x[2] - xc[2];
    const double a = p0 * p0 + p1 * p1;
    out[2] = atan2(p1, p0);
    out[1] = atan2(sqrt(a), p2);
    out[0] = H / sqrt(a + p2 * p2);
}

__device__ __forceinline__ void interp_to_cart_exact(const double p[3], const double xc[3],
                                                     double H, double out[3])
{
    const double r = H / p[0];
    out[0] = xc[0] + r * cos(p[2]) * sin(p[1]);
    out[1] = xc[1] + r * sin(p[2]) * sin(p[1]);
    out[2] = xc[2] + r * cos(p[1]);
}

// Grid::cell of the host plan: -1 outside the closed domain box, else the clamped cell id
// (dimension 0 fastest)
__device__ __forceinline__ int64_t grid_cell(const double p[3], const double lo[3],
                                             const double hi[3], const double h[3],
                                             const int32_t n[3])
{
    double d2 = 0;
    for (int k = 0; k < 3; k++)
    {
        if (lo[k] > p[k])
        {
            const double a = lo[k] - p[k];
            d2 += a * a;
        }
        else if (p[k] > hi[k])
        {
            const double a = p[k] - hi[k];
            d2 += a * a;
        }
    }
    if (d2 > 0)
        return -1;
    int64_t id = 0, stride = 1;
    for (int j = 0; j < 3; j++)
    {
        const int q = int(floor((p[j] - lo[j]) / h[j]));
        const int64_t c = std::min<int64_t>(std::max<int64_t>(q, 0), n[j] - 1);
        id += c * stride;
        stride *= n[j];
    }
    return id;
}

__device__ __forceinline__ void set_bit(unsigned long long* words, int64_t c)
{
    atomicOr(words + (c >> 6), 1ull << (c & 63));
}

__device__ __forceinline__ void mark_point(const MarkBox& B, const double s[3],
                                           unsigned long long* fbits,
                                           unsigned long long* cbits)
{
    const int64_t cf = grid_cell(s, B.lo, B.hi, B.hf, B.nf);
    const int64_t cc = grid_cell(s, B.lo, B.hi, B.hc, B.nc);
    if (cc >= 0)
        set_bit(cbits + B.coff, cc);
    if (cf >= 0)
        set_bit(fbits + B.foff, cf);
}

__global__ __launch_bounds__(256) void k_mark_far(const MarkBox* __restrict__ boxes,
                                                 const int32_t* __restrict__ pair_box,
                                                 const int32_t* __restrict__ pair_tgt,
                                                 const int64_t* __restrict__ tbeg,
                                                 const int64_t* __restrict__ tend,
                                                 const double* __restrict__ tpts,
                                                 unsigned long long* __restrict__ fbits,
                                                 unsigned long long* __restrict__ cbits)
{
    const int64_t r = blockIdx.x;
    const MarkBox B = boxes[pair_box[r]];
    const int32_t t = pair_tgt[r];
    const int64_t i1 = tend[t];
    for (int64_t i = tbeg[t] + threadIdx.x; i < i1; i += 256)
    {
        const double x[3] = {tpts[3 * i], tpts[3 * i + 1], tpts[3 * i + 2]};
        double s[3];
        cart_to_interp_exact(x, B.ctr, B.H, s);
        mark_point(B, s, fbits, cbits);
    }
}

// tab_fc: (fine, coarse) of the parent level's upward-table entries, ((u * n_ho) + node) * 8
// + octant; cone_u: compact table index of each parent cone (table path); cone_cell: coarse
// cell of each parent cone (per-box path)
__global__ __launch_bounds__(256) void k_mark_parent(
    const MarkBox* __restrict__ boxes, const ParentWork* __restrict__ work,
    const int32_t* __restrict__ tab_fc, const int32_t* __restrict__ cone_u,
    const int64_t* __restrict__ cone_cell, const double* __restrict__ ho_nodes, int n_ho,
    int use_table, unsigned long long* __restrict__ fbits, unsigned long long* __restrict__ cbits)
{
    const ParentWork W = work[blockIdx.x];
    const MarkBox B = boxes[W.q];
    const int64_t n = (W.c1 - W.c0) * n_ho;
    for (int64_t e = threadIdx.x; e < n; e += 256)
    {
        const int64_t c = W.c0 + e / n_ho;
        const int k = int(e % n_ho);
        if (use_table)
        {
            const int64_t at = ((int64_t(cone_u[c]) * n_ho + k) * 8 + W.oct) * 2;
            const int32_t fine = tab_fc[at], coarse = tab_fc[at + 1];
            if (fine >= 0)
            {
                if (coarse >= 0)
                    set_bit(cbits + B.coff, coarse);
                set_bit(fbits + B.foff, fine);
            }
            continue;
        }
        // the parent's coarse grid point of this node (Grid::point), its physical position
        // and its cone coordinates in this box
        int64_t cell = cone_cell[c];
        double sp[3];
        for (int i = 0; i < 3; i++)
        {
            const int64_t idx = cell % W.pnc[i];
            cell /= W.pnc[i];
            const double a = 0.5 * W.phc[i];
            const double b = W.plo[i] + (double(idx) + 0.5) * W.phc[i];
            sp[i] = (ho_nodes[3 * k + i] * a) + b;
        }
        double y[3], s[3];
        interp_to_cart_exact(sp, W.pctr, W.pH, y);
        cart_to_interp_exact(y, B.ctr, B.H, s);
        mark_point(B, s, fbits, cbits);
    }
}

// device buffers kept across the levels of one build_cones
struct MarkerState
{
    bool ready = false;
    DBuf tpts, tbeg, tend, far_tgt, ho;
    DBuf boxes, pair_box, fbits, cbits, work, tab_fc, cone_u, cone_cell;
};

void mark_level(MarkerState& st, const Plan& P, int level, const std::vector<int64_t>& far_off,
                const std::vector<int32_t>& far_tgt, const std::vector<double>& ho_nodes,
                LevelBits& lb)
{
    const Tree& S = P.S;
    const Tree& T = P.T;
    if (!st.ready)
    {
        h2d(st.tpts, T.pts.data(), T.pts.size());
        h2d(st.tbeg, T.begin.data(), T.begin.size());
        h2d(st.tend, T.end.data(), T.end.size());
        h2d(st.far_tgt, far_tgt.data(), far_tgt.size());
        h2d(st.ho, ho_nodes.data(), ho_nodes.size());
        st.ready = true;
    }
    const int64_t l0 = S.level_off[level], l1 = S.level_off[level + 1];
    const int64_t nl = l1 - l0;
    std::vector<MarkBox> boxes(nl);
    for (int64_t q = 0; q < nl; q++)
    {
        const int64_t g = l0 + q;
        MarkBox& B = boxes[q];
        for (int d = 0; d < 3; d++)
        {
            B.ctr[d] = S.ctr[3 * g + d];
            B.lo[d] = P.dom_lo[3 * g + d];
            B.hi[d] = P.dom_hi[3 * g + d];
            B.hf[d] = P.hf[3 * g + d];
            B.hc[d] = P.hc[3 * g + d];
            B.nf[d] = int32_t(P.nf[3 * g + d]);
            B.nc[d] = int32_t(P.nc[3 * g + d]);
        }
        B.H = S.side[g];
        B.foff = lb.foff[q];
        B.coff = lb.coff[q];
    }
    // far pairs of the level's boxes: far_tgt[far_off[l0] .. far_off[l1]) in box order
    const int64_t r0 = far_off[l0], r1 = far_off[l1];
    std::vector<int32_t> pair_box(r1 - r0);
    for (int64_t q = 0; q < nl; q++)
        for (int64_t r = far_off[l0 + q]; r < far_off[l0 + q + 1]; r++)
            pair_box[r - r0] = int32_t(q);
    // parent work
    std::vector<ParentWork> work;
    const UpLevel* U = level > 0 ? &P.up[level - 1] : nullptr;
    int64_t pc_base = 0;
    if (level > 0)
    {
        const int64_t p0 = S.level_off[level - 1];
        pc_base = P.cone_off[p0];
        for (int64_t q = 0; q < nl; q++)
        {
            const int64_t g = l0 + q;
            if (!P.parent_has_far[g])
                continue;
            const int32_t pg = S.parent[g];
            if (P.cone_off[pg + 1] == P.cone_off[pg])
                continue;
            ParentWork W{};
            W.q = int32_t(q);
            int oct = 0;
            while (S.child[8 * pg + oct] != int32_t(g))
                oct++;
            W.oct = oct;
            W.c0 = P.cone_off[pg] - pc_base;
            W.c1 = P.cone_off[pg + 1] - pc_base;
            for (int d = 0; d < 3; d++)
            {
                W.pctr[d] = S.ctr[3 * pg + d];
                W.plo[d] = P.dom_lo[3 * pg + d];
                W.phc[d] = P.hc[3 * pg + d];
                W.pnc[d] = int32_t(P.nc[3 * pg + d]);
            }
            W.pH = S.side[pg];
            work.push_back(W);
        }
    }
    if (r1 == r0 && work.empty())
        return; // nothing marks this level: the bitmaps stay zero
    h2d(st.boxes, boxes.data(), boxes.size());
    st.fbits.reserve(sizeof(uint64_t) * lb.fbits.size());
    st.cbits.reserve(sizeof(uint64_t) * lb.cbits.size());
    if (!lb.fbits.empty())
        HIP_OK(hipMemset(st.fbits.p, 0, sizeof(uint64_t) * lb.fbits.size()));
    if (!lb.cbits.empty())
        HIP_OK(hipMemset(st.cbits.p, 0, sizeof(uint64_t) * lb.cbits.size()));
    if (r1 > r0)
    {
        h2d(st.pair_box, pair_box.data(), pair_box.size());
        k_mark_far<<<dim3(unsigned(r1 - r0)), dim3(256)>>>(
            st.boxes.as<MarkBox>(), st.pair_box.as<int32_t>(), st.far_tgt.as<int32_t>() + r0,
            st.tbeg.as<int64_t>(), st.tend.as<int64_t>(), st.tpts.as<double>(),
            st.fbits.as<unsigned long long>(), st.cbits.as<unsigned long long>());
        HIP_OK(hipGetLastError());
    }
    if (!work.empty())
    {
        const int64_t p0 = S.level_off[level - 1], p1 = S.level_off[level];
        const int64_t npc = P.cone_off[p1] - pc_base;
        const int n_ho = int(ho_nodes.size() / 3);
        const bool table = U->valid;
        h2d(st.work, work.data(), work.size());
        if (table)
        {
            std::vector<int32_t> cu(npc), fc(U->entries.size() * 2);
            for (int64_t c = 0; c < npc; c++)
                cu[c] = U->cellmap[P.cone_cell[pc_base + c]];
            for (size_t e = 0; e < U->entries.size(); e++)
            {
                fc[2 * e] = U->entries[e].fine;
                fc[2 * e + 1] = U->entries[e].coarse;
            }
            h2d(st.cone_u, cu.data(), cu.size());
            h2d(st.tab_fc, fc.data(), fc.size());
        }
        else
            h2d(st.cone_cell, P.cone_cell.data() + pc_base, size_t(npc));
        (void)p0;
        k_mark_parent<<<dim3(unsigned(work.size())), dim3(256)>>>(
            st.boxes.as<MarkBox>(), st.work.as<ParentWork>(), st.tab_fc.as<int32_t>(),
            st.cone_u.as<int32_t>(), st.cone_cell.as<int64_t>(), st.ho.as<double>(), n_ho,
            table ? 1 : 0, st.fbits.as<unsigned long long>(), st.cbits.as<unsigned long long>());
        HIP_OK(hipGetLastError());
    }
    if (!lb.fbits.empty())
        HIP_OK(hipMemcpy(lb.fbits.data(), st.fbits.p, sizeof(uint64_t) * lb.fbits.size(),
                               hipMemcpyDeviceToHost));
    if (!lb.cbits.empty())
        HIP_OK(hipMemcpy(lb.cbits.data(), st.cbits.p, sizeof(uint64_t) * lb.cbits.size(),
                               hipMemcpyDeviceToHost));
}

} // namespace

ConeMarker make_device_cone_marker()
{
    auto st = std::make_shared<MarkerState>();
    return [st](const Plan& P, int level, const std::vector<int64_t>& far_off,
                const std::vector<int32_t>& far_tgt, const std::vector<double>& ho_nodes,
                LevelBits& lb) { mark_level(*st, P, level, far_off, far_tgt, ho_nodes, lb); };
}

} // namespace ifgf
