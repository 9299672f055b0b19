// HIP kernels of the FMM evaluation (gfx950).  Launch wrappers only; the
// kernels live in kernels.hip.
#pragma once

#include <hip/hip_runtime.h>

#include <cstdint>

namespace sfmm {

// What a charge is and what is written per particle.
enum Mode : int {
    MODE_POT = 0,      // nc charge components -> nc potentials (rinv)
    MODE_CURL = 1,     // 3-vector charge -> curl field Q x R/|R|^3 (rinv<true>)
    MODE_POTGRAD = 2,  // scalar charge -> potential and its gradient (4 outputs)
};

// Device view of everything a launch needs.  All pointers are device memory.
struct DevPlan {
    int p = 0, K = 0, depth = 0;
    int64_t n = 0;
    // particles in sorted ("slot") order
    const double *px = nullptr, *py = nullptr, *pz = nullptr;  // raw positions
    const double *ux = nullptr, *uy = nullptr, *uz = nullptr;  // (rotated - leaf centre) / leaf width
    const int32_t* perm = nullptr;                             // slot -> original index
    const int32_t* iperm = nullptr;                            // original index -> slot
    // leaves with far field (level >= 2)
    int64_t n_far_leaf = 0;
    const int32_t *fl_gnode = nullptr, *fl_level = nullptr;
    const int64_t *fl_begin = nullptr, *fl_end = nullptr;
    const int32_t* slot_fleaf = nullptr;  // sorted slot -> far leaf index, -1 for near-only leaves
    // all leaves, near field
    int64_t n_leaf = 0;
    const int64_t *lf_begin = nullptr, *lf_end = nullptr;
    const int64_t *near_row = nullptr, *near_begin = nullptr, *near_end = nullptr;
    const int64_t* near_pref = nullptr;  // per range: sources of the leaf's earlier ranges
    // nodes
    const int8_t* octant = nullptr;     // code & 7 per gnode
    const int32_t* parent = nullptr;    // parent gnode (-1 at level 0)
    const int32_t* child0 = nullptr;    // first child gnode (-1 for leaves)
    const int8_t* nchild = nullptr;
    // M2L work: targets (gnodes at levels >= 2) and their interaction lists
    int64_t n_target = 0;
    const int32_t* tgt_gnode = nullptr;
    const int64_t* il_row = nullptr;
    const int32_t* il_src = nullptr;   // source gnode
    const int16_t* il_off = nullptr;   // offset id
    const double* inv_width = nullptr;  // per level 1/w_l
    // tables
    const double *m2mT = nullptr, *l2lT = nullptr;  // [oct][col][row]
    const double *rot = nullptr, *coax = nullptr, *off_cs = nullptr;
    const int32_t *off_class = nullptr, *off_d2 = nullptr;
    const int32_t *om = nullptr, *om_rows = nullptr, *rot_off = nullptr, *coax_off = nullptr;
    int rot_size = 0, coax_size = 0;
    const double* rec = nullptr;  // recurrence coefficients, see kernels.hip
};

// Offset-grouped M2L work (see m2l_mfma.hip): blocks of 16/nc targets of one
// level and child octant, each with its list of (offset, 16/nc sources).
struct M2LArgs {
    int64_t n_block = 0;
    const int32_t* blk_tgt = nullptr;  // [n_block][16/nc] target gnodes, -1 = none
    const int64_t* blk_row = nullptr;  // [n_block+1] entry ranges
    const int32_t* ent_meta = nullptr; // per entry: offset id | polar class << 9 | distance class << 16
    const int32_t* ent_src = nullptr;  // [entries][16/nc] source gnodes, -1 = none
    const void* rotA = nullptr;        // rotation fragments [70][2][nfrag][64], double or float
    const void* coaxA = nullptr;       // coaxial fragments [28][nfrag][64]
    const double* off_cs = nullptr;    // [343][p+1] (cos, sin) of m * azimuth
    const double* zero = nullptr;      // 4 (p+1)^2 zeros: the columns of absent sources
    const float* M32 = nullptr;        // FP32 copy of the multipoles (FP32 stages gather from it)
    int fp32 = 0;                      // 1: FP32 MFMA stages (mixed precision)
};

// Mixed-precision M2L on the f16 matrix cores (m2l_dense.hpp): blocks of
// targets of one level and child octant; per block the interaction offsets of
// its targets grouped by azimuth, each with one source per target.
struct DenseArgs {
    int64_t n_block = 0;
    const int32_t* blk_tgt = nullptr;   // [n_block][targets per block] target gnodes, -1 = none
    const int64_t* blk_row = nullptr;   // [n_block+1] entry ranges
    const int32_t* ent_meta = nullptr;  // per entry: polar class | azimuth group << 8
    const int32_t* ent_src = nullptr;   // [entries][targets per block] source gnodes, -1 = none
    const void* ops = nullptr;          // per class: [fragment][hi 64 | lo 64] x 8 halves (f16)
    const float* az = nullptr;          // [azimuth group][p+1][2] cos, sin of m phi
    const float* nodemax = nullptr;     // [gnode][nc] largest |value| of each FP32 multipole component
    const float* zero = nullptr;        // zeros: the columns of absent sources
    int op_exp = 0;                     // the operators are stored scaled by 2^op_exp
    int waves = 8;                      // wavefronts per workgroup (4 or 8): 16 targets per wavefront group
};

// Near-field work: items of (leaf, run of its concatenated near sources).
struct P2PItems {
    int64_t n_item = 0;
    const int64_t* leaf = nullptr;       // global leaf id
    const int64_t* src_begin = nullptr;  // first source (index into the concatenated near list)
    const int64_t* src_count = nullptr;
    const int64_t* partial = nullptr;    // -1: add into the output; else first partial row
    // split leaves: per target, the item partial rows first + i*stride, i < nitem
    int64_t n_split_targets = 0;
    int64_t n_partial_rows = 0;
    const int64_t* split_slot = nullptr;
    const int64_t* split_first = nullptr;
    const int64_t* split_nitem = nullptr;
    const int64_t* split_stride = nullptr;
};

// Self exclusions from a self generator (fmm.hpp:398-408): per target slot,
// the sources (slots) of its list other than itself and whether each lies in
// the target's near field (partition.hpp:484-488).
struct SelfCorr {
    int64_t n_row = 0;
    const int64_t* tgt = nullptr;  // target slot per row
    const int64_t* row = nullptr;  // [n_row+1]
    const int64_t* src = nullptr;  // source slot
    const int8_t* near = nullptr;  // 1: the source is in the target's near field
};

struct LevelSpan {
    int64_t first = 0, count = 0;  // gnode range of one level
};

// M2M / L2L on matrix cores: blocks of 16/nc output nodes, each with entries
// (child octant, 16/nc input nodes); the operator of an octant is a dense
// (p+1)^2 x (p+1)^2 matrix stored as MFMA A fragments.
struct TransArgs {
    // M2M and L2L share these lists: per block 16/nc parents, per entry the
    // children of one octant of those parents
    const int64_t* blk_row = nullptr;   // [n_block+1] entry ranges
    const int32_t* blk_node = nullptr;  // [n_block][16/nc] parent gnodes, -1 = none
    const int32_t* ent_oct = nullptr;   // child octant per entry
    const int32_t* ent_node = nullptr;  // [entries][16/nc] child gnodes, -1 = none
    const double* frag = nullptr;      // [8][nfrag][64]
};

// All launches are asynchronous on `stream`.
void launch_gather_charges(const DevPlan& d, int qdim, const double* Q, double* q, hipStream_t s);
void launch_scatter_outputs(const DevPlan& d, int nout, const double* out, double* U, hipStream_t s);
void launch_p2m(const DevPlan& d, Mode mode, int nc, const double* q, double* M, hipStream_t s);
void launch_m2m(const DevPlan& d, LevelSpan parents, int nc, double* M, hipStream_t s);
void launch_m2l(const DevPlan& d, int nc, const double* M, double* L, hipStream_t s);
void launch_l2l(const DevPlan& d, LevelSpan children, int nc, double* L, hipStream_t s);
void launch_l2p(const DevPlan& d, int nc, const double* L, double* out, hipStream_t s);
void launch_m2l_mfma(const M2LArgs& a, int p, int nc, const double* M, double* L, hipStream_t s);
// p <= 12; M32: the multipoles in FP32, [gnode][nc][dense_order(p).kd] (m2l_dense.hpp)
void launch_m2l_dense(const DenseArgs& a, int p, int nc, const float* M32, double* L, hipStream_t s);
// The FP32 copy (F32LayoutOf<p>: the dense M2L's layout for p <= 12, else the
// real-form order) of `count` nodes of nc x (p+1)^2 FP64 coefficients, with
// their per-component largest |value| into nodemax when it is not null.
void launch_fp32_copy(const double* M, int64_t count, int nc, int p, float* M32, float* nodemax, hipStream_t s);
// upper = false: M2M (parents written); true: L2L (children added to).
void launch_translate(const TransArgs& a, LevelSpan blocks, int p, int nc, bool upper, double* io, hipStream_t s);
int translate_nfrag(int p, bool upper);
// A-fragment index of element (row, col) of a (p+1)^2 operator, -1 if not stored.
int translate_frag_index(int p, bool upper, int row, int col, int* lane);
// mixed: FP32 pair arithmetic with FP64 accumulation per source tile.
// partial: P2PItems::n_partial_rows x nout doubles of scratch.
// tpl: targets per lane (1, 2 or 4), the kernel's tile shape.
void launch_p2p(const DevPlan& d, const P2PItems& w, Mode mode, int nc, const double* q, double* out,
                double* partial, hipStream_t s, bool mixed, int tpl = 2);
// Subtracts the pairs of the self lists: far ones always, near ones when
// use_near (the near field then used the lists as its exclusion predicate).
void launch_self_correction(const DevPlan& d, const SelfCorr& c, Mode mode, int nc, bool use_near, const double* q,
                            double* out, hipStream_t s);
// q: charges in slot order ([c][slot]), or, with qperm = the slot -> original
// index map, the caller's original-order array ([index][c]).  Targets outside
// [0, n) give NaN.
void launch_direct(const DevPlan& d, Mode mode, int nc, const double* q, int64_t ntarget, const int64_t* targets,
                   double* out, double* partial, int nsplit, hipStream_t s, const int32_t* qperm = nullptr);
void launch_fill(double* x, int64_t count, double v, hipStream_t s);

// Number of double slots the direct sum needs as scratch for `ntarget` targets.
int64_t direct_scratch_size(int64_t ntarget, int nout, int nsplit);
// Source splits of the direct sum for `ntarget` targets among n particles:
// ntarget * nsplit <= 2^19 whenever ntarget <= 2^19.
int direct_nsplit(int64_t ntarget, int64_t n);

}  // namespace sfmm
