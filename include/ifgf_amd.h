/* ifgf_amd.h -- C ABI of the MI355X IFGF operator library (libifgf_amd.so).
 *
 * Drop-in boundary for TimHRO/ifgf's Python module `pyifgf` (py_ifgf.cpp:16-48): one opaque
 * operator per IfgfOperator instance, with the reference's construct / init / mult
 * semantics.  Plain pointers and sizes only; no exception crosses the ABI: every call
 * returns IFGF_OK or a negative code, and ifgf_last_error() describes the failure.
 *
 * Points are xyz-interleaved float64 arrays (N x 3, i.e. the column-major 3 x N
 * Eigen::Array<double,3,Dynamic> the reference takes).  Densities and results are
 * complex128 (interleaved re, im); nrhs density vectors are stored one after another.
 */
#ifndef IFGF_AMD_H
#define IFGF_AMD_H

#include <stdint.h>

#ifdef __cplusplus
extern "C" {
#endif

#define IFGF_OK 0
#define IFGF_EINVAL (-1)    /* bad argument */
#define IFGF_ENODEV (-2)    /* no HIP device: the library has no CPU path */
#define IFGF_ESTATE (-3)    /* call out of order (mult before init, ...) */
#define IFGF_EINTERNAL (-4) /* HIP or allocation failure */

/* kernel families (the L5 classes of the reference) */
#define IFGF_HELMHOLTZ 0     /* helmholtz_ifgf.hpp: e^{i kappa d}/(4 pi d), k_re = kappa      */
#define IFGF_MODHELMHOLTZ 1  /* modified_helmholtz_ifgf.hpp: e^{-k d}/(4 pi d), complex k     */
#define IFGF_LAPLACE 2       /* laplace_ifgf.hpp: 1/(4 pi d)                                  */
#define IFGF_DOUBLELAYER 3   /* double_layer_helmholtz_ifgf.hpp: complex k, source normals    */
#define IFGF_COMBINED 4      /* combined_field_helmholtz_ifgf.hpp: kappa = k_re, normals      */
#define IFGF_GRAD 5          /* grad_helmholtz_ifgf.hpp: complex k, ifgf_set_dx               */

/* arithmetic of mult */
#define IFGF_PREC_FP64 0  /* every step in FP64 */
#define IFGF_PREC_MIXED 1 /* FP32 arithmetic; FP64 cone coordinates, phases, geometry */

/* ifgf_init flags */
#define IFGF_INIT_PLAN_ONLY 1 /* build the host plan only (introspection); mult is refused */
#define IFGF_INIT_HOST_PLAN 2 /* sort the points, build the octrees and mark the active cones
                                 with the host code (by default a device init does these on
                                 the GPU; both give the same plan) */
#define IFGF_INIT_LEVEL_RING 4 /* keep the interpolation values of one level and the fine
                                  coefficients of two levels on the device instead of every
                                  level's (same results; the far field of a level runs as soon
                                  as its coefficients are written, and mults of several columns
                                  run one column at a time).  Chosen by init when every level's
                                  do not fit the free device memory. */

/* ifgf_mult flags */
#define IFGF_DEVICE_POINTERS 1 /* w and out are device pointers; work is enqueued on stream */
#define IFGF_MULT_GRAPH 2      /* with IFGF_DEVICE_POINTERS: replay the mult's kernels between
                                  the gather and the scatter as a captured HIP graph, one for
                                  the single-column mult and one for the batched mult of several
                                  columns (each captured on its first such call and reused;
                                  captured again after ifgf_set_launch_shape / ifgf_autotune
                                  change a shape, ifgf_set_dx changes the direction or the batch
                                  buffers change width) */

typedef struct ifgf_operator ifgf_operator;

/* cutOff callback of init (ifgfoperator.hpp:65-67): nonzero drops an admissible pair */
typedef int (*ifgf_cutoff_fn)(double dist, void* ctx);

/* Replaces the IfgfOperator constructors (ifgfoperator.hpp:36-59; py_ifgf.cpp:19,46):
 * (k, leafSize, order, n_elem, tolerance).  k_re/k_im: complex wavenumber (k_re = kappa
 * for the real-kappa families; ignored for Laplace).  leaf_size < 0 means unbounded. */
int ifgf_create(int kernel, double k_re, double k_im, int64_t leaf_size, int order, int n_elem,
                double tol, int precision, ifgf_operator** out);

/* GradHelmholtzIfgfOperator::setDx (grad_helmholtz_ifgf.hpp:60): -1 potential, 0..2 d/dy_i */
int ifgf_set_dx(ifgf_operator* op, int dx);

/* IfgfOperator::init (ifgfoperator.hpp:65-106) and the normals overload of the
 * double-layer / combined-field operators (double_layer_helmholtz_ifgf.hpp:22-26).
 * normals may be NULL; cutoff may be NULL (never drop). */
int ifgf_init(ifgf_operator* op, const double* srcs, int64_t n_src, const double* tgts,
              int64_t n_tgt, const double* normals, ifgf_cutoff_fn cutoff, void* cutoff_ctx,
              int flags);

/* Plan reuse across processes.  No counterpart in the reference, whose init rebuilds the
 * octrees, interaction lists and cone domains every time (ifgfoperator.hpp:65-106,
 * octree.hpp:202-816); these replace that work when the same points come back.
 * ifgf_plan_save writes the operator's plan (both octrees, interaction lists, active cones,
 * upward tables, the geometries of a merged operator) and its launch shape to a file.
 * ifgf_init_from_plan reads one into an operator created with the same kernel, wavenumber,
 * leaf size, order, n_elem and tolerance (IFGF_EINVAL otherwise), then uploads it as
 * ifgf_init does (flags: IFGF_INIT_PLAN_ONLY, IFGF_INIT_LEVEL_RING) and applies the saved
 * launch shape; the gradient direction stays the operator's (ifgf_set_dx).  Results equal
 * those of the operator that wrote the file, bit for bit.  The file is checked (format,
 * version, byte order, length, a 64-bit checksum, array sizes and indices); a damaged or
 * foreign file gives IFGF_EINVAL and leaves the operator as it was.  Load only files this
 * library wrote: the checks catch damage, not a file forged to pass them.  A plan built
 * with a cutoff callback keeps its interaction lists, but the callback does not travel: the
 * adjoint for targets other than the sources then needs ifgf_init. */
int ifgf_plan_save(ifgf_operator* op, const char* path);
int ifgf_init_from_plan(ifgf_operator* op, const char* path, int flags);

/* IfgfOperator::mult (ifgfoperator.hpp:246-867) for nrhs density vectors.
 * stream: a hipStream_t (NULL = default stream).  With host pointers the call returns
 * after the result is copied back; with IFGF_DEVICE_POINTERS it only enqueues. */
int ifgf_mult(ifgf_operator* op, const double* w, int64_t nrhs, double* out, int flags,
              void* stream);

/* Allocation-free mult: one column of device densities w_dev (complex128, N) into device
 * results out_dev (complex128, M), enqueued on stream (NULL = default stream); allocates no
 * device memory, creates no stream, event or graph, and runs every kernel on that one stream.
 * No counterpart in the reference: for callers that cannot allocate device memory (code under
 * HIP graph capture, tight serving loops, harnesses that forbid hipMalloc).  The Morton-order
 * densities and results live in a fixed 64 MiB device array of the library (one per GPU, held
 * until the process ends, its address looked up at ifgf_init): operators with
 * n_src + n_tgt > 8,388,576 (mixed) or 4,194,272 (FP64) get IFGF_EINVAL.  The results are
 * those of ifgf_mult with IFGF_DEVICE_POINTERS, bit for bit.  Calls that overlap on one GPU
 * (concurrent streams or threads) race on the array.  The current device must be the one of
 * ifgf_init. */
int ifgf_mult_noalloc(ifgf_operator* op, const double* w_dev, double* out_dev, void* stream);

/* The adjoint product A^H g of mult (out: N values per column, g: M).  No counterpart in the
 * reference, which has no adjoint; it serves the autograd wrapper's backward (the gradient
 * with respect to the densities).  For the potentials G(x - y) = G(y - x) and for the
 * gradient kernel, odd in x - y, A^H g = (+-) conj(B conj(g)) with B the IFGF operator from the
 * targets to the sources: this plan when the targets are the sources, else a transposed plan
 * built and uploaded on the first call (the cutoff callback of init must still be valid).
 * The double-layer and combined-field kernels return IFGF_EINVAL (their adjoints need
 * target normals).  Flags and stream as for ifgf_mult. */
int ifgf_mult_adjoint(ifgf_operator* op, const double* g, int64_t nrhs, double* out, int flags,
                      void* stream);

/* Several operators of one kernel family and one set of interpolation orders, each after
 * ifgf_init (IFGF_INIT_PLAN_ONLY is enough), as one operator on the device: every phase runs
 * one launch per level for all of them.  mult takes the parts' densities concatenated in order
 * (sum of N) and returns their results concatenated (sum of M); every part keeps its own
 * wavenumber and points.  Precision and kernel family are those of ops[0].  No counterpart in
 * the reference (one geometry per IfgfOperator); for batches of many geometries.  The parts
 * stay the caller's; errors are reported on ops[0].  flags: IFGF_INIT_PLAN_ONLY merges the
 * host plans only (introspection). */
int ifgf_merge(ifgf_operator* const* ops, int n_ops, int flags, ifgf_operator** out);

/* Numbers of sources and targets of a planned operator (a merged one: the sums). */
int ifgf_sizes(const ifgf_operator* op, int64_t* n_src, int64_t* n_tgt);

void ifgf_destroy(ifgf_operator* op);
const char* ifgf_last_error(const ifgf_operator* op);
const char* ifgf_version(void);
int ifgf_device_count(void);

/* --- plan introspection (integer parity with Octree / ConeDomain) ---------------------
 * counts: [0] source levels [1] source boxes [2] target levels [3] target boxes
 *         [4] far pairs [5] near pairs [6] coarse cones [7] fine cones (covered)
 *         [8] fine cones active but not covered [9] dropped (cutOff) pairs
 *         [10] device bytes [11] fine cone ids listed
 *         [12] levels with precomputed upward geometry [13] entries of those tables
 *         [14] level storage (IFGF_INIT_LEVEL_RING) in use
 *         [15] coarse cones of a mult's standalone values -> coefficients transform (fewer than
 *              [6] when leaf sampling and the upward pass transform the cones they compute);
 *              the array holds 16 values */
int ifgf_plan_counts(const ifgf_operator* op, int64_t* counts);
/* orders: base[3], coarse (high) order[3], fine (low) order[3], factor[3] */
int ifgf_plan_orders(const ifgf_operator* op, int* out12);
/* tree 0 = sources, 1 = targets; arrays sized by the counts above (level_off: levels+1) */
int ifgf_plan_tree(const ifgf_operator* op, int tree, int64_t* level_off, int64_t* begin,
                   int64_t* end, int32_t* parent, int8_t* leaf, double* bbox6, int64_t* perm);
/* interaction pairs (global source box, global target box), sorted by target then source */
int ifgf_plan_pairs(const ifgf_operator* op, int far, int32_t* src, int32_t* tgt);
/* per source box: coarse/fine elements (3 each), cone_off/fine_off (boxes+1); per coarse
 * cone: cell id and fine sub-cone mask; fine_ids: active fine cells per box, ascending */
int ifgf_plan_cones(const ifgf_operator* op, int64_t* nc, int64_t* nf, int64_t* cone_off,
                    int64_t* cone_cell, uint32_t* mask, int64_t* fine_off, int64_t* fine_ids);
/* algorithmic work of one mult (per density vector), from the plan:
 * [0] near point pairs  [1] leaf (node, source) pairs  [2] upward (node, child) evaluations
 * [3] far (target, box) pairs  [4] coarse cones  [5] fine cones  [6] leaf coarse cones
 * [7] n_hi  [8] n_lo  [9] transform flops per coarse cone */
int ifgf_plan_work(const ifgf_operator* op, double* out10);

/* per-phase device times of this operator's last mult in ms (near, leaf, upward, transform,
 * far, total); filled only when the environment variable IFGF_PHASE_TIMING is set (adjoint
 * products of other targets run on the transposed operator and do not update them) */
int ifgf_last_timings(const ifgf_operator* op, double* ms6);

/* --- launch shapes (no counterpart in the reference: its TBB loops have no tile shapes) ---
 * shape: IFGF_LAUNCH_SHAPE_FIELDS ints
 *         [0] XCD run log2 (work items dealt to the 8 XCDs in runs of 2^v; 0 = plain order)
 *         [1] transform cones per workgroup (0 = default; the compiled order classes only)
 *         [2] upward parent cones per workgroup (2, 4, 8, 16; default 16)
 *         [3] batched columns per pass (0 = by nrhs, 8, 32)
 *         [4] mixed far field: waves per workgroup sharing the staged angle tables (1, 2, 4, 8)
 *         [5] second stream for the far field beside the upward chain, with the near field
 *             there (1) or first on the main stream (2); 0: one stream
 * Shapes change speed, never results.  ifgf_autotune times the candidates on the operator
 * itself (synthetic densities, nrhs columns, the fastest of reps mults per candidate), one
 * choice after another, applies the best and returns it; after ifgf_init, not thread-safe
 * with concurrent mults of the same operator. */
#define IFGF_LAUNCH_SHAPE_FIELDS 6
int ifgf_autotune(ifgf_operator* op, int64_t nrhs, int reps, int* shape);
int ifgf_get_launch_shape(const ifgf_operator* op, int* shape);
int ifgf_set_launch_shape(ifgf_operator* op, const int* shape);

#ifdef __cplusplus
}
#endif
#endif /* IFGF_AMD_H */
