// device.hpp -- device-resident plan and the mult pipeline (host-side declarations).
#pragma once

#include <cstdint>
#include <string>
#include <vector>

#include "plan.hpp"

namespace ifgf
{
enum Precision
{
    PREC_FP64 = 0,  // every step in FP64 (the plain form)
    PREC_MIXED = 1  // FP32 arithmetic; FP64 for cone coordinates, phases and box geometry
};

// geometry and cone grid of one source box, as read by the kernels
struct BoxGeo
{
    double ctr[3];
    double H;
    double lo[3];     // cone domain (s, theta, phi)
    double hi[3];
    double hf[3];     // fine cell widths
    double inv_hf[3]; // 1 / hf (cell search; exact division near cell boundaries)
    double hc[3];     // coarse cell widths
    int32_t nf[3];    // fine cells per axis
    int32_t nc[3];    // coarse cells per axis
    int64_t table_off;
    int32_t th_off;   // node-angle tables of the coarse grid (cos/sin per cell and node)
    int32_t ph_off;
    int32_t fth_off;  // fine-grid theta table: (cos, sin) of the nf1+1 cell edges, then centres
    int32_t fph_off;  // fine-grid phi table: the same for the nf2 cells
    double k_re, k_im, kappa; // wavenumber of the box's geometry (e^{-k d}; kappa: real form)
};

struct ConeDev
{
    int32_t box;
    int32_t cell;      // coarse cell id within the box grid
    uint32_t mask;     // active fine sub-cones
    int32_t fine_base; // first fine slot
};

// coarse cell of a box grid, as the cone lookups read it: the fine slots of its cone in one
// load (mask 0: no cone in the cell)
struct CellRef
{
    int32_t fine_base;
    uint32_t mask;
};

struct WorkItem // a chunk of at most 64 consecutive target points of one target box
{
    int32_t box;   // target box (global id in the target tree)
    int32_t first; // first target point (sorted order)
    int32_t count; // 1..64
    int32_t pad;
};

// Launch shapes of the mult's kernels: choices that change speed but not results, picked by
// device_autotune (measured on the operator itself) or set by the caller.
struct LaunchShape
{
    int xcd_log2 = 4; // work items dealt to the XCDs in runs of 2^xcd_log2 (0: plain order)
    int cpb = 0;      // transform: cones per workgroup (0: the order class's default)
    int up_k = 16;    // upward: parent cones of one cell per workgroup (2, 4, 8 or 16)
    int rb = 0;       // batched: columns per pass (8 or 32; 0: 8 up to 8 columns, else 32)
    int far_waves = 1; // mixed far field: waves (items) per workgroup sharing the LDS tables
    int overlap = 1;   // far field on a second stream beside the upward chain, with the near
                       // field there too (1) or first on the main stream (2); 0: one stream
};
constexpr int UP_K_CHOICES[4] = {2, 4, 8, 16};
// tiles of the grouped upward kernel's work list: parent boxes (in Morton order within a level)
// x cells (table index) per tile
constexpr int UP_TILE_BOXES = 16;
constexpr int UP_TILE_CELLS = 128;

// What k_up_grp needs of a parent cone before its tasks, per entry of the tiled upward list
// (FP32, compiled orders), computed at upload from the plan's fixed tables: read as one
// 16-byte-aligned record instead of a chain of dependent loads (cone -> box -> children ->
// their tables -> the (cell, octant) rows -> the listed child cones' slots).
constexpr int UP_DMAX = 32; // listed child cones per (parent cone, child) with a slot here
struct alignas(16) UpDesc
{
    int64_t toff[8];          // child c's coarse-cell table offset (-1: no child c)
    int32_t tg0[8];           // the group record of child c's task t is tg0[c] + t
    int32_t tend[8];          // child c's tasks end at tend[c] (prefix over the children)
    int32_t g;                // the parent cone
    uint32_t mask;            // its active fine sub-cones
    int32_t fine_base;        // its first fine slot
    int32_t pad;
    int32_t slot[8][UP_DMAX]; // fine slot of listed child cone d of child c (-1: inactive)
};
static_assert(sizeof(UpDesc) == 1168, "UpDesc is 73 16-byte words");

// bytes of the fixed device scratch of the allocation-free mult (device_mult_noalloc): the
// Morton-order densities and results of one column (64 MiB: n_src + n_tgt up to 8,388,576
// points in the mixed form, 4,194,272 in FP64)
constexpr size_t NOALLOC_SCRATCH_BYTES = size_t(64) << 20;

struct DevicePlan
{
    int device = -1; // HIP device of the upload
    int prec = PREC_MIXED;
    int kind = 0;
    double k_re = 0, k_im = 0; // e^{-k d}
    double kappa = 0;
    int dx = -1;
    int ho[3], lo[3], f[3];
    int n_hi = 0, n_lo = 0, nsub = 0;
    int64_t nsrc = 0, ntgt = 0;
    int levels = 0;
    int64_t n_cones = 0, n_fine = 0;

    // device arrays (raw pointers owned by this struct)
    double* src_pts = nullptr;   // sorted, xyz
    double* tgt_pts = nullptr;
    double* normals = nullptr;   // sorted source normals (DL, CF) or null
    int64_t* src_perm = nullptr;
    int64_t* tgt_perm = nullptr;
    int64_t* src_begin = nullptr; // per source box
    int64_t* src_end = nullptr;
    int32_t* child = nullptr;     // 8 per source box
    BoxGeo* geo = nullptr;
    ConeDev* cones = nullptr;
    CellRef* ctab = nullptr;      // per box: dense coarse-cell table (BoxGeo::table_off)
    double* ho_nodes = nullptr;   // n_hi x 3 reference nodes in [-1,1]^3
    double* trig_th = nullptr;    // (cos, sin) of theta at coarse cell / node, per grid class
    double* trig_ph = nullptr;    // (cos, sin) of phi
    double* fine_ang = nullptr;   // fine-grid edge / centre (cos, sin) tables
    void* wmat = nullptr;         // axis maps (R): W0 | W1 | W2
    int woff[4] = {0, 0, 0, 0};

    // upward geometry tables: entries {cf, sub, xi[3], tf[2]} in R of every part (a level of
    // one geometry), up_part_ent[part] = the part's first entry
    void* up_entries = nullptr;
    int64_t* up_part_ent = nullptr;
    int64_t n_up_parts = 0;
    // FP32 tables: per (cell, octant) the distinct child cones (cf, sub) its entries fall into,
    // in order of first appearance (UpEntF::d indexes them): up_dkey pairs, (cell, octant) row
    // r = (up_part_cell[part] + cell) * 8 + octant spans up_dkey[2 up_doff[r] .. 2 up_doff[r+1])
    int32_t* up_dkey = nullptr;
    int32_t* up_doff = nullptr;
    // the same tables regrouped by child cone (UpGrp records; FP32): the groups of row r span
    // up_grp[up_goff[r] .. up_goff[r + 1])
    int32_t* up_goff = nullptr;
    void* up_grp = nullptr;
    int64_t* up_part_cell = nullptr;

    // (cell, cones of that cell, part) items of at most UP_K_CHOICES[i] cones, one list per
    // choice; up_items / up_item_off name the list of the current shape
    WorkItem* up_items_k[4] = {nullptr, nullptr, nullptr, nullptr};
    std::vector<int64_t> up_item_off_k[4];
    WorkItem* up_items = nullptr;
    std::vector<int64_t> up_item_off;        // per level, into up_items
    int32_t* up_glist = nullptr;             // the items' cones, level-major
    std::vector<int64_t> up_gl_off;          // per level, into up_glist
    // the same cones for k_up_grp (which takes an item's cones one at a time): ordered by tiles
    // of UP_TILE_BOXES boxes x UP_TILE_CELLS cells, items of at most UP_K_CHOICES[i] cones
    // within a tile, each cone with its cell's table index (tl_gcell); same per-level offsets
    WorkItem* up_tl_items_k[4] = {nullptr, nullptr, nullptr, nullptr};
    std::vector<int64_t> up_tl_item_off_k[4];
    WorkItem* up_tl_items = nullptr;
    std::vector<int64_t> up_tl_item_off;
    int32_t* up_tl_glist = nullptr;
    int32_t* up_tl_gcell = nullptr;
    UpDesc* up_tl_desc = nullptr; // per entry of up_tl_glist (FP32, compiled orders)

    int32_t* leaf_cones = nullptr;
    int64_t n_leaf_cones = 0;
    // per level, the cones whose values -> coefficients transform runs on its own (non-leaf
    // cones without an upward table) when leaf sampling and k_up_grp transform theirs
    int32_t* tr_list = nullptr;
    std::vector<int64_t> tr_list_off;
    WorkItem* leaf_boxes = nullptr; // leaf boxes with active cones: (box, first cone, count)
    int64_t n_leaf_boxes = 0;
    int32_t* up_cones = nullptr;             // non-leaf cones without a table, level-major
    std::vector<int64_t> up_off;             // per level, into up_cones
    std::vector<int64_t> cone_lvl_off;       // per level, global cone range

    int32_t* far_src = nullptr;   // far source boxes, by target box then source box
    int32_t* far_srange = nullptr; // per far item: its sources far_src[2i] .. far_src[2i + 1]
    int32_t* near_off = nullptr;
    int32_t* near_src = nullptr;
    WorkItem* far_items = nullptr;
    // far work in launch groups of one source level and one target level, deepest source level
    // first: a group can run as soon as its source level's coefficients exist; groups run one
    // after another (a target point is in boxes of several levels)
    struct FarGroup
    {
        int64_t item0, item1; // far_items[item0 .. item1)
        int ang_base, ang_n;  // the sources' angle tables staged in LDS (mixed form; 0: none)
        int src_level, tgt_level;
    };
    std::vector<FarGroup> far_groups;
    WorkItem* near_items = nullptr;
    int64_t n_near_items = 0;

    // work buffers of the single-column mult
    void* wsorted = nullptr;  // N complex (R)
    void* result = nullptr;   // M complex (R)
    void* V = nullptr;        // n_cones * n_hi complex (R)
    void* F = nullptr;        // n_fine * n_lo complex (R)
    // level storage (ring): instead of every level's values and coefficients, V holds the values
    // of one level and F_bank[l & 1] the fine coefficients of level l (a level's coefficients
    // are read only by the upward pass of the level above and the far field of that source
    // level); for operators whose coefficients of all levels exceed the device memory
    bool ring = false;
    void* F_bank[2] = {nullptr, nullptr};
    std::vector<int64_t> fine_lvl_off;    // per level, global fine-slot range
    std::vector<int64_t> leaf_lvl_off;    // per level, range of leaf_boxes
    std::vector<void*> ev_far_done;       // ring on two streams: far groups of a source level done
    int64_t bytes_total = 0;

    // work buffers of the batched mult (b_rb columns, column-minor), allocated on first use;
    // b_in / b_out stage host densities and results
    // the single-column mult's core captured as a graph (on first use with graphs enabled),
    // and the batched mult's core for the current batch buffers
    void* graph = nullptr;      // hipGraph_t
    void* graph_exec = nullptr; // hipGraphExec_t
    void* b_graph = nullptr;
    void* b_graph_exec = nullptr;
    void* cap_stream = nullptr; // hipStream_t of the capture
    // the second stream of the overlapped near field and its fork / join events
    void* side_stream = nullptr;
    void* ev_fork = nullptr;
    void* ev_near = nullptr;
    std::vector<void*> ev_level; // per source level: its coefficients are written

    LaunchShape shape;
    // per-phase device times of this operator's last timed mult (ms: near, leaf, upward,
    // transform, far, total); force_timing: time the phases of the next mults (autotuner)
    float last_ms[6] = {0, 0, 0, 0, 0, 0};
    bool force_timing = false;

    // staging of host densities and results of single-column mults (first use)
    double *h_in = nullptr, *h_out = nullptr;

    int b_rb = 0;
    void *b_ws = nullptr, *b_res = nullptr, *b_V = nullptr, *b_F = nullptr;
    double *b_in = nullptr, *b_out = nullptr;
    int64_t b_bytes = 0;
};

// ring: 0 = every level resident when it fits the device memory, else level storage;
// 1 = level storage (DevicePlan::ring)
std::string device_upload(const Plan& P, int precision, DevicePlan& D, int ring = 0);
void device_free(DevicePlan& D);
// coarse cones of one mult's standalone values -> coefficients transform (all of them unless
// leaf sampling and the grouped upward kernel transform their own)
int64_t device_transform_cones(const DevicePlan& D);
// batched mult (batched.hip): ncol <= rb columns (rb = 8 or 32) of device densities with the
// batch buffers allocated for rb; the compiled order classes only (batched_supported)
bool batched_supported(const DevicePlan& D);
void mult_batched(DevicePlan& D, const double* w_dev, double* out_dev, int rb, int ncol,
                  void* stream, int io, bool graph = false);
// w: complex128 (host or device), out: complex128 (host or device); nrhs columns of length
// N / M stored one after another.  io: IO_* bits applied while the densities are gathered and
// the results scattered (conjugations and a sign, for adjoint products).
enum
{
    IO_CONJ_IN = 1,
    IO_CONJ_OUT = 2,
    IO_NEGATE_OUT = 4
};
std::string device_mult(DevicePlan& D, const void* w, void* out, int64_t nrhs, bool device_ptrs,
                        void* stream, int io = 0, bool graph = false);
// One column of device densities into device results with no device allocation: the staging
// vectors live in the library's fixed per-device scratch (NOALLOC_SCRATCH_BYTES; larger
// operators are refused), everything on `stream`, no phase timing, no graph.  Calls that
// overlap on one GPU race on the scratch.  The current device must be the upload's.
std::string device_mult_noalloc(DevicePlan& D, const double* w_dev, double* out_dev,
                                void* stream);
// Apply a launch shape (validated; an unknown choice is an error).  Drops a captured graph.
std::string device_set_shape(DevicePlan& D, const LaunchShape& s);
// Drop the captured graphs of the single-column and batched mults (after a change of their
// kernel arguments).
void device_drop_graph(DevicePlan& D);
// Time the candidate shapes on this operator with synthetic densities (nrhs columns, each
// candidate `reps` times, the fastest kept), one choice after another: the XCD run length on
// the whole mult, then the transform's cones per workgroup, the upward cones per item and the
// far field's waves per workgroup on their phases, then (nrhs >= 2) the batched columns per
// pass.  Leaves the best shape applied.
std::string device_autotune(DevicePlan& D, int64_t nrhs, int reps, LaunchShape& best);
// per-phase timings of the last mult (ms): near, leaf, upward, transform, far, total
void device_last_timings(const DevicePlan& D, double* out6);
bool device_available();

// The source or target octree of init on the device (tree.hip): the same Tree as build_tree
// (identical permutation, boxes and point ranges).  Returns "" or an error.
std::string device_build_tree(const double* pts, int64_t n, int64_t max_leaf, Tree& T);
// The cone marking of build_cones on the device (cones.hip): the same bitmaps as the host's.
ConeMarker make_device_cone_marker();
// The interaction lists on the device (pairs.hip): the same sorted lists as build_interactions.
std::string device_build_interactions(Plan& P);
// all three device builders
inline PlanBuilders device_plan_builders()
{
    PlanBuilders b;
    b.tree = device_build_tree;
    b.pairs = device_build_interactions;
    b.cones = make_device_cone_marker();
    return b;
}
} // namespace ifgf
