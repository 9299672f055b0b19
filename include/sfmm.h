/*
 * sfmm — MI355X-native Laplace FMM evaluator, C-ABI.
 *
 * Drop-in for the evaluation path of icyspirit/SimpleFMM: the calls below are
 * what that project's header-only C++ API (fmm.hpp, driven as in main.cpp)
 * does, split into a host tree (Box + OctreePartitioner) and a device plan
 * (FMM3D).  Plain pointers and sizes only; every function returns 0 on success
 * or a non-zero status whose text sfmm_last_error() returns.
 *
 * Particle arrays are caller-owned, in the caller's ("original") order, AoS:
 * positions double[3n]; charges double[qdim*n]; outputs double[nout*n].
 * Outputs are ACCUMULATED (U += ...), never zeroed, as the reference's rinv
 * (fmm.hpp:767,788,792,801,805).
 */
#ifndef SFMM_H
#define SFMM_H

#include <stdint.h>

#ifdef __cplusplus
extern "C" {
#endif

typedef struct sfmm_tree_s* sfmm_tree;
typedef struct sfmm_plan_s* sfmm_plan;

enum {
    SFMM_OK = 0,
    SFMM_ERR_ARG = 1,      /* invalid argument */
    SFMM_ERR_HIP = 2,      /* HIP runtime error (no device, launch failure, ...) */
    SFMM_ERR_ALLOC = 3,    /* device allocation failed */
    SFMM_ERR_STATE = 4     /* call not valid in the object's state */
};

/* Evaluation kinds.  qdim/nout are per particle. */
enum {
    SFMM_POTENTIAL = 0, /* rinv<false>: ncharge charge components -> ncharge potentials, sum Q_j/|r| */
    SFMM_CURL = 1,      /* rinv<true>: 3-vector charge -> sum Q_j x R/|R|^3, R = x_i - x_j          */
    SFMM_POTGRAD = 2    /* scalar charge -> (phi, dphi/dx, dphi/dy, dphi/dz), nout = 4               */
};

/* Which part of the sum (rinv = both, rinv_nonear = far, N2N = near). */
enum { SFMM_FAR = 1, SFMM_NEAR = 2, SFMM_BOTH = 3 };
/* or-ed into `which`: the near field excludes the pairs of the plan's self lists
 * (rinv(Q, U, is_self) with is_self(i, j) = j in self_generator(i), fmm.hpp:864-873);
 * without it only j == i is excluded (rinv(Q, U), fmm.hpp:875-881).  Listed pairs in
 * the far field are subtracted either way (_slist, fmm.hpp:398-408,798-807). */
enum { SFMM_SELF_LIST = 4 };

enum { SFMM_FP64 = 0, SFMM_MIXED = 1 };

typedef struct {
    int64_t n;                 /* number of particles                                 */
    const double* positions;   /* double[3n], AoS                                     */
    int max_level;             /* refine(max_level, ...) — at most 20                 */
    int max_per_leaf;          /* refine(..., max_particles_per_node)                 */
    int vertical_axis;         /* Box(..., vertical_axis, phideg)                     */
    double phideg;
    int box_given;             /* 0: Box(positions, ...); 1: Box(corner, width, ...)  */
    double corner[3];
    double width;
    int threads;               /* host threads for the build (0 = hardware)           */
    int on_device;             /* 1: build on the current GPU (same tree, bit for bit) */
} sfmm_tree_desc;

/* Box(positions, vertical_axis, phideg) alone (box.hpp:84-107): out4 = width, corner x, y, z. */
int sfmm_box_from_positions(int64_t n, const double* positions, int vertical_axis, double phideg, double out4[4]);

/* Box(...) + OctreePartitioner(positions, box) + refine(max_level, cap):
 * box.hpp:73-107, partition.hpp:397-446.  The tree copies the positions. */
int sfmm_tree_create(const sfmm_tree_desc* desc, sfmm_tree* out);
int sfmm_tree_destroy(sfmm_tree tree);

/* FNV-1a hash over every array of the tree (permutation, per-level codes, ranges,
 * links, leaves, neighbour and interaction lists, near ranges): equal digests mean
 * the same tree.  No counterpart in the reference; it compares host and device builds. */
int sfmm_tree_digest(sfmm_tree tree, uint64_t* digest);

/* Near-field sources of original particle i: the particles of the leaves in its
 * leaf's near list (partition.hpp:324-366), the pairs N2N visits (fmm.hpp:779-785),
 * as original indices in sorted-slot order; i itself is among them.  *count receives
 * the number; out (may be NULL) receives up to `capacity` of them, and the call fails
 * with SFMM_ERR_ARG when capacity is smaller than the count. */
int sfmm_tree_near_sources(sfmm_tree tree, int64_t i, int64_t* out, int64_t capacity, int64_t* count);

typedef struct {
    int depth;             /* partitioner.level()            */
    int64_t n_node;        /* over all levels                */
    int64_t n_leaf;
    int64_t m2l_pairs;     /* sum of interaction list sizes  */
    int64_t near_pairs;    /* near-field (i, j) pairs, i != j */
    double width;          /* box.width()                    */
    double corner[3];      /* box.corner()                   */
} sfmm_tree_info;
int sfmm_tree_get_info(sfmm_tree tree, sfmm_tree_info* info);

/* Per-level counts: n_node, n_leaf, nnz(indices), nnz(ilist), nnz(clist), nnz(nlist). */
int sfmm_tree_level_counts(sfmm_tree tree, int level, int64_t counts[6]);
/* Per-level lists in the reference's layout (partition.hpp:155-166); any pointer may be NULL.
 * ilist values: (src, di, dj, dk); clist: (child, code); plist: (parent, code); nlist: (level, leaf). */
int sfmm_tree_level_lists(sfmm_tree tree, int level, int64_t* c_node, int64_t* leaf2node, int64_t* idx_row,
                          int64_t* idx_val, int64_t* ilist_row, int64_t* ilist_val, int64_t* clist_row,
                          int64_t* clist_val, int64_t* plist, int64_t* nlist_row, int64_t* nlist_val);

typedef struct {
    int p;                    /* expansion order, 1..16 (FMM3D template p)            */
    int precision;            /* SFMM_FP64 or SFMM_MIXED (FP32 MFMA M2L stages, FP64 accumulation) */
    int max_ncharge;          /* largest ncharge later passed to SFMM_POTENTIAL (1..4) */
    /* optional self lists (FMM3D(partitioner, comm, self_generator), fmm.hpp:204):
     * self_generator(i) = self_val[self_row[i] .. self_row[i+1]) (original indices,
     * self_row[0] = 0); NULL means self_generator(i) = {i}.  Copied at creation. */
    const int64_t* self_row;
    const int64_t* self_val;
    int m2l_kernel;           /* 0: offset-grouped MFMA (default); 1: row-parallel VALU kernel */
} sfmm_plan_desc;

/* FMM3D(partitioner, comm[, self_generator]): fmm.hpp:204-427.  Uploads the tree
 * and the operator tables to the current HIP device.  The tree must outlive the plan. */
int sfmm_plan_create(sfmm_tree tree, const sfmm_plan_desc* desc, sfmm_plan* out);
int sfmm_plan_destroy(sfmm_plan plan);

/* rinv / rinv<true> / rinv_nonear / N2N with host arrays (fmm.hpp:811-881):
 * U[nout*i + c] += value, original particle order.  Synchronous. */
int sfmm_evaluate(sfmm_plan plan, int kind, int ncharge, const double* Q, double* U, int which);

/* Same with device arrays, asynchronous on the plan's stream (sfmm_plan_stream).
 * dQ: double[qdim*n] AoS original order; dU: double[nout*n], accumulated. */
int sfmm_evaluate_device(sfmm_plan plan, int kind, int ncharge, const double* dQ, double* dU, int which);

/* FMM3D::slist() (fmm.hpp:398-408,439-442): per original particle i, the entries of
 * self_generator(i) that lie outside i's near field (subtracted from every near-field
 * sum), ascending, duplicates kept.  row: int64[n+1] or NULL; val: int64[*nnz] or NULL. */
int sfmm_plan_slist(sfmm_plan plan, int64_t* row, int64_t* val, int64_t* nnz);

/* The plan's HIP stream (hipStream_t) for sfmm_evaluate_device. */
int sfmm_plan_stream(sfmm_plan plan, void** stream);

/* No counterpart in the reference (its kernels have no tile shapes).  Times the
 * kernels' tile shapes on this plan with one evaluation's device charges (as for
 * sfmm_evaluate_device): the near field's targets per lane (1, 2, 4) and, for the
 * mixed path at p <= 12, the M2L's wavefronts per workgroup (4, 8); keeps the
 * fastest for later evaluations.  A shape changes the rounding, not the method.
 * Synchronous; allocates scratch outputs while it runs. */
int sfmm_plan_tune(sfmm_plan plan, int kind, int ncharge, const double* dQ);
/* The plan's tile shapes (m2l_waves 0: the plan does not use the dense M2L). */
int sfmm_plan_tiles(sfmm_plan plan, int* p2p_targets_per_lane, int* m2l_waves);
/* Sets the tile shapes explicitly (p2p 1, 2 or 4; m2l 4 or 8, or 0 to keep; the
 * M2L shape applies to plans that use the dense M2L).  Waits for the plan's stream. */
int sfmm_plan_set_tiles(sfmm_plan plan, int p2p_targets_per_lane, int m2l_waves);
int sfmm_synchronize(sfmm_plan plan);

/* Device milliseconds of each stage of the last evaluation:
 * [0] gather, [1] P2M, [2] M2M, [3] M2L, [4] L2L, [5] L2P, [6] P2P, [7] scatter, [8] total. */
int sfmm_last_stage_ms(sfmm_plan plan, float ms[9]);

/* Records the stage times of each of the next `count` evaluations in its own set of
 * HIP events on the plan's stream (instead of only the last one); count <= 0 stops. */
int sfmm_record_stages(sfmm_plan plan, int count);
/* Stage times (as sfmm_last_stage_ms, 9 per evaluation) of the evaluations recorded
 * since sfmm_record_stages; ms holds room for `capacity` evaluations. */
int sfmm_recorded_stage_ms(sfmm_plan plan, float* ms, int capacity, int* recorded);

/* FP64 direct sum over all sources j != i for the given original-order targets
 * (main.cpp:70-96): out[nout*t + c] = value (overwritten).  Host arrays, synchronous. */
int sfmm_direct(sfmm_plan plan, int kind, int ncharge, const double* Q, int64_t ntarget, const int64_t* targets,
                double* out);

/* No counterpart in the reference.  For callers that cannot allocate device
 * memory (code under HIP graph capture, serving loops, harnesses that forbid
 * hipMalloc): sfmm_direct with device arrays and no allocation.
 * dQ: charges in original order as for sfmm_evaluate_device; dtargets: ntarget
 * original indices; dout[nout*t + c] overwritten.  Asynchronous on the plan's
 * stream; the same sums as sfmm_direct, bit for bit.  The partial sums live in
 * a static device array of 16 MiB per GPU (held until the process ends), so
 * ntarget <= SFMM_DIRECT_NOALLOC_MAX_TARGETS (else SFMM_ERR_ARG).  Targets are
 * not checked on the host: one outside [0, n) gives NaN outputs.  Every call on
 * one GPU shares that array: calls that overlap on one GPU (other streams,
 * other threads, other plans) race and must be ordered by the caller. */
#define SFMM_DIRECT_NOALLOC_MAX_TARGETS 524288
int sfmm_direct_device_noalloc(sfmm_plan plan, int kind, int ncharge, const double* dQ, int64_t ntarget,
                               const int64_t* dtargets, double* dout);

/* ---- Batches of independent problems.
 * No counterpart in the reference (one problem per FMM3D object): many plans
 * (each its own tree, order and charges) evaluated by one call, spread over
 * concurrent streams. */
typedef struct sfmm_batch_s* sfmm_batch;
typedef struct {
    sfmm_plan plan;     /* distinct plans, one device */
    int kind;           /* SFMM_POTENTIAL, SFMM_CURL or SFMM_POTGRAD */
    int ncharge;
    int which;          /* as for sfmm_evaluate */
    const double* dQ;   /* device charges, AoS original order */
    double* dU;         /* device outputs, accumulated */
} sfmm_batch_item;

/* Prepares every plan for its item, assigns the items to concurrent streams
 * (largest first) and captures the launch sequence into a HIP graph that every
 * launch replays.  streams: 1..8, or 0 to tune first (into scratch outputs): the
 * tile shapes of each expansion order on its largest problem (as
 * sfmm_plan_tune; set on every plan of that order), then 1, 2, 4, 8 streams and
 * every item on its plan's own stream, each timed as graph replays, keeping the
 * fastest.  The items' arrays must stay valid while the batch is used; a batch
 * keeps the tile shapes its plans had when it was created. */
int sfmm_batch_create(const sfmm_batch_item* items, int count, int streams, sfmm_batch* out);
/* Evaluates every item (one graph launch) after the work queued on `stream`
 * (hipStream_t; NULL: the batch's own stream) and on each plan's own stream, both
 * of which then wait for all of them; asynchronous. */
int sfmm_batch_launch(sfmm_batch batch, void* stream);
/* Waits for launches on the batch's own stream. */
int sfmm_batch_synchronize(sfmm_batch batch);
/* The number of concurrent streams of the batch (the item count when every item
 * runs on its plan's own stream). */
int sfmm_batch_streams(sfmm_batch batch, int* streams);
int sfmm_batch_destroy(sfmm_batch batch);

/* Text of the last error on this thread. */
const char* sfmm_last_error(void);

#ifdef __cplusplus
}
#endif

#endif
