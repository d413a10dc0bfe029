// plan.hpp -- the integer plan of an IFGF operator as flat, level-major arrays.
//
// Reference layers replaced: octree.hpp (Octree::build, buildOctreeNode,
// buildInteractionList, calculateInterpolationRange), zorder_less.hpp (Morton order),
// cone_domain.hpp (ConeDomain) and the per-kernel discretisation policies
// (orderForBox / elementsForBox / cutoff_limit of helmholtz_ifgf.hpp and siblings).
// Instead of shared_ptr trees and per-box hash sets the plan is a set of arrays indexed
// by a global box id (level-major) and a global coarse-cone id (box-major), so it can be
// uploaded to the device as-is.
#pragma once

#include <complex>
#include <cstdint>
#include <functional>
#include <string>
#include <vector>

namespace ifgf
{
using cplx = std::complex<double>;

enum KernelKind
{
    KIND_HELMHOLTZ = 0,     // helmholtz_ifgf.hpp (real kappa)
    KIND_MODHELMHOLTZ = 1,  // modified_helmholtz_ifgf.hpp (complex k)
    KIND_LAPLACE = 2,       // laplace_ifgf.hpp
    KIND_DOUBLELAYER = 3,   // double_layer_helmholtz_ifgf.hpp (complex k, source normals)
    KIND_COMBINED = 4,      // combined_field_helmholtz_ifgf.hpp (real kappa, source normals)
    KIND_GRAD = 5           // grad_helmholtz_ifgf.hpp (complex k, setDx)
};

// Discretisation policy of one kernel family.  All families are written with a complex
// wavenumber k and the factor e^{-k d}; real-kappa Helmholtz is k = -i kappa.
struct KernelSpec
{
    int kind = KIND_HELMHOLTZ;
    cplx k{0, 0};     // e^{-k d}
    double kappa = 0; // real wavenumber for the real-kappa families
    int dx = -1;      // gradient component (-1: potential)
    double tol = -1;

    // orderForBox (helmholtz_ifgf.hpp:158-170, modified_helmholtz_ifgf.hpp:138-153,
    // laplace_ifgf.hpp:133-145, double_layer_helmholtz_ifgf.hpp:161-186,
    // combined_field_helmholtz_ifgf.hpp:184-198, grad_helmholtz_ifgf.hpp:238-264)
    void orders(const int base[3], int step, int out[3]) const;
    // elementsForBox (helmholtz_ifgf.hpp:172-195 and siblings)
    void elements(double H, const int64_t base_els[3], int step, int64_t out[3]) const;
    // cutoff_limit (ifgfoperator.hpp:1914, grad_helmholtz_ifgf.hpp:294-304)
    double smin(double H) const;
    // fine/coarse element factor per axis (3, or 1 for the modified Helmholtz kernel)
    int fine_factor() const { return kind == KIND_MODHELMHOLTZ ? 1 : 3; }
};

struct Tree
{
    int64_t n = 0;
    std::vector<double> pts;    // Morton-sorted points, xyz interleaved
    std::vector<int64_t> perm;  // sorted position -> caller index
    double lo[3], hi[3];        // padded root box
    int levels = 0;
    std::vector<int64_t> level_off;   // global box id of the first box of each level
    // per global box
    std::vector<int64_t> begin, end;  // point range
    std::vector<int32_t> parent;      // global id or -1
    std::vector<int32_t> child;       // 8 per box, global id or -1
    std::vector<uint8_t> leaf;
    std::vector<double> blo, bhi;     // 3 per box
    std::vector<double> ctr;          // 3 per box: (hi+lo)/2
    std::vector<double> side;         // max extent

    int64_t num_boxes() const { return int64_t(begin.size()); }
    int level_of(int64_t g) const;
};

struct Pair
{
    int32_t src, tgt; // global box ids in the source / target trees
};

// Upward geometry of one parent level.  All boxes of a level are translates of one another,
// so where the high-order node q of parent coarse cell c lands in the fine grid of the child
// in octant o (cell, local coordinates) and the transfer factor e^{-k(d - dp)} dp/d do not
// depend on the box.  Entries cover the union of active coarse cells of the level's parents.
struct UpEntry
{
    int32_t fine;   // fine cell of the child; -1: outside the cone domain
    int32_t coarse; // coarse cell of the child (the reference marks both grids independently)
    int32_t cf;     // coarse cell containing the fine cell (the cone the device reads)
    int32_t sub;    // sub-cone index of the fine cell within cf
    double xi[3];   // local coordinates in the fine cell
    double tf[2];   // transfer factor (complex)
};

struct UpLevel
{
    bool valid = false;
    std::vector<int32_t> cellmap;  // parent coarse cell -> compact index (or -1)
    std::vector<UpEntry> entries;  // ((compact * n_hi) + node) * 8 + octant
};

struct Plan
{
    KernelSpec K;
    int64_t base_els[3];
    int base_order[3];     // after init (tolerance estimate or radial reduction)
    int ho_order[3];       // coarse (step 1) order
    int lo_order[3];       // fine (step 0) order
    int factor[3];         // fine elements / coarse elements per axis
    Tree S, T;
    std::vector<double> normals; // source normals in source Morton order (DL, CF)

    // interaction lists (octree.hpp:255-349)
    std::vector<Pair> far, near;
    int64_t n_tofar = 0;

    // per source box (global id): ancestor flags (octree.hpp:947-978)
    std::vector<uint8_t> has_far_anc, parent_has_far;

    // per source box: cone grids.  Both grids share the (s, theta, phi) domain box.
    std::vector<double> dom_lo, dom_hi;        // 3 per box
    std::vector<int64_t> nc, nf;               // coarse / fine elements, 3 per box
    std::vector<double> hc, hf;                // cell widths, 3 per box
    std::vector<int64_t> table_off;            // per box: first coarse-cell slot (size nbox+1)
    std::vector<int64_t> cone_off;             // per box: first global coarse cone (size nbox+1)

    // per global coarse cone (box-major, cell id ascending within a box)
    std::vector<int32_t> cone_box;
    std::vector<int64_t> cone_cell;            // cell id in the box's coarse grid
    std::vector<uint32_t> cone_mask;           // active fine sub-cones (bit = sub index)
    std::vector<int64_t> cone_fine;            // first fine slot of this cone
    int64_t n_fine = 0;                        // number of active, covered fine cones
    int64_t n_fine_uncovered = 0;              // active fine cones whose coarse cone is inactive

    // dense coarse-cell -> global cone id (or -1), concatenated over boxes
    std::vector<int32_t> table;

    // fine active cone ids per box (sorted) for parity dumps: CSR
    std::vector<int64_t> fine_off;
    std::vector<int64_t> fine_ids;

    // per source level: precomputed upward geometry (valid when the level's boxes share one
    // grid class and the table fits the budget)
    std::vector<UpLevel> up;
    int64_t up_budget = int64_t(64) << 20; // entries

    // several geometries in one plan (merge_plans): the upward tables of each (geometry, level)
    // with the parent boxes they serve, and the wavenumber of every source box (k_re, k_im,
    // kappa).  Empty for a plan of one geometry (tables per level in `up`, P.K for all boxes).
    struct UpPart
    {
        int level;
        UpLevel U;
    };
    std::vector<UpPart> up_parts;
    std::vector<int32_t> box_up_part; // per source box: its up_parts index or -1
    std::vector<double> box_k;        // 3 per source box
    std::vector<int64_t> part_src_off, part_tgt_off; // caller-order offsets of the geometries

    int64_t num_cones() const { return int64_t(cone_box.size()); }
};

// Morton key of one point as two 64-bit words (hi, lo): 3 sign bits (z, y, x), then for
// bit positions 0, -1, ..., -40 of the box-normalised magnitudes the (z, y, x) bits.  The
// order of keys is the order of zorder_knn::Less (zorder_less.hpp:212-248) except among
// points whose normalised coordinates agree to 2^-40; those are ordered by the exact
// comparison below.
void morton_key(const double p[3], const double ctr[3], const double diam[3], uint64_t* hi,
                uint64_t* lo);
bool morton_less_exact(const double* px, const double* qx, const double ctr[3],
                       const double diam[3]);

// The tree in three steps (octree.hpp:202-238, 1106-1199): tree_root pads the bounding box
// [mn, mx] (boundingbox.hpp:13-26) and gives the Morton normalisation (ctr, diam); the points
// are then sorted (perm, pts); grow_tree adds levels until no box has more than max_leaf
// points, asking split_level for the child point ranges of each level (the greedy scan of
// the closed child boxes, lower octant first) and forming the child boxes with child_box.
struct ChildRange
{
    int64_t b[8], e[8];
};
using LevelSplit = std::function<void(const Tree& T, int64_t lb, int64_t le, int64_t max_leaf,
                                      std::vector<ChildRange>& cr, std::vector<uint8_t>& split)>;
void tree_root(Tree& T, int64_t n, const double mn[3], const double mx[3], double ctr[3],
               double diam[3]);
void child_box(const double* blo, const double* bhi, int j, double lo[3], double hi[3]);
void grow_tree(Tree& T, int64_t max_leaf, const LevelSplit& split_level);
void build_tree(const double* pts, int64_t n, int64_t max_leaf, Tree& T, int nthreads);
// builds the trees when set (the device builder); build_tree otherwise
using TreeBuilder = std::function<std::string(const double* pts, int64_t n, int64_t max_leaf,
                                              Tree& T)>;
void build_interactions(Plan& P, const std::function<bool(double)>& cutoff);
// after the far and near lists are sorted by (target, source): has_far_anc, parent_has_far
void interaction_flags(Plan& P);
// builds P.far and P.near sorted by (target, source) and calls interaction_flags (the device
// builder; used when no cutOff callback is set).  Returns "" or an error.
using PairBuilder = std::function<std::string(Plan& P)>;
int estimate_orders(Plan& P, double tol, int out[3]); // ifgfoperator.hpp:108-244
// Cone activity of one source level: for box q of the level (global id level_off + q) the
// fine-grid and coarse-grid bitmaps (bit = cell id) in 64-bit words fbits[foff[q] ..
// foff[q + 1]) and cbits[coff[q] .. coff[q + 1]), zeroed by build_cones.  A cell is active
// when a far target of the box, or a high-order node of an active coarse cone of its parent,
// lies in it (octree.hpp:487-548, 620-700).  The marker (when set) fills them; build_cones
// otherwise marks on the host.  Then each box's active coarse cones (ascending cell), fine
// sub-cone masks and fine cell lists follow from the bitmaps.
struct LevelBits
{
    std::vector<int64_t> foff, coff;
    std::vector<uint64_t> fbits, cbits;
};
using ConeMarker = std::function<void(const Plan& P, int level,
                                      const std::vector<int64_t>& far_off,
                                      const std::vector<int32_t>& far_tgt,
                                      const std::vector<double>& ho_nodes, LevelBits& bits)>;
void build_cones(Plan& P, int nthreads, const ConeMarker& marker = nullptr);

// the device builders of init (unset members: the host code)
struct PlanBuilders
{
    TreeBuilder tree;
    PairBuilder pairs;
    ConeMarker cones;
};

// Build everything init() needs (ifgfoperator.hpp:65-106).  Returns "" or an error.
std::string build_plan(Plan& P, const double* srcs, int64_t N, const double* tgts, int64_t M,
                       const double* normals, int64_t leaf, int order, int n_elem,
                       const std::function<bool(double)>& cutoff, int nthreads,
                       const PlanBuilders& dev = PlanBuilders());

// One plan for several planned geometries of one kernel family and one set of orders: the
// boxes of every level are those of part 0, then part 1, ...; points, cones, coarse cells and
// pairs follow; sources and targets are the parts' concatenated in order.  Returns "" or an
// error (families, orders or normals that differ).
std::string merge_plans(const std::vector<const Plan*>& parts, Plan& M);

// Chebyshev helpers shared by the host plan and the device tables
std::vector<double> chebnodes1d(int n); // x_i = cos((2i+1) pi / 2n), chebinterp.hpp:27-39

// the (f*o_lo) x o_hi matrix mapping high-order nodal values along one axis to the
// low-order Chebyshev coefficients of the f sub-intervals (chebtransform o coarseToFine o
// chebtransform along that axis; chebinterp.hpp:176-250, ifgfoperator.hpp:1539-1637)
std::vector<double> axis_map(int o_hi, int o_lo, int f);

int default_threads();
} // namespace ifgf
