// C-ABI of the FMM evaluator: host tree, device plan, evaluation orchestration.
//
// The plan mirrors FMM3D (icyspirit/SimpleFMM fmm.hpp:134-882): built once from
// a partitioner, then evaluated any number of times with new charges.  The
// level sweep of rinv_nonear (fmm.hpp:811-862) becomes one stream of kernels:
// P2M, M2M per level bottom-up, M2L over all levels, L2L per level top-down,
// L2P; rinv adds the near field (N2N, fmm.hpp:773-809).
#include "../../include/sfmm.h"
#include "kernels.hpp"
#include "m2l_dense.hpp"
#include "tables.hpp"
#include "tree.hpp"

#include <hip/hip_runtime.h>

#include <sched.h>

#include <algorithm>
#include <cmath>
#include <cstdlib>
#include <cstring>
#include <functional>
#include <memory>
#include <mutex>
#include <string>
#include <thread>
#include <vector>

namespace {

thread_local std::string g_last_error;

// Host threads for plan construction: the CPUs this process may run on.
int host_threads()
{
    cpu_set_t set;
    int n = 0;
    if (sched_getaffinity(0, sizeof set, &set) == 0) n = CPU_COUNT(&set);
    if (n <= 0) n = static_cast<int>(std::thread::hardware_concurrency());
    return std::max(1, n);
}

// f(begin, end, chunk) over [0, count) cut into `chunks` contiguous ranges, one thread each.
void parallel_ranges(int64_t count, int chunks, const std::function<void(int64_t, int64_t, int)>& f)
{
    chunks = static_cast<int>(std::max<int64_t>(1, std::min<int64_t>(chunks, count / 1024)));
    if (chunks <= 1) {
        f(0, count, 0);
        return;
    }
    std::vector<std::thread> pool;
    for (int c = 0; c < chunks; ++c)
        pool.emplace_back([&f, count, chunks, c] { f(count * c / chunks, count * (c + 1) / chunks, c); });
    for (auto& t : pool) t.join();
}

int fail(int code, const std::string& msg)
{
    g_last_error = msg;
    return code;
}

// Every device allocation of the library goes through here.  Fault injection
// for tests: while the environment variable SFMM_FAULT_ALLOC holds a count k,
// the first k allocations after it was set (or changed) succeed and every
// later one fails as if the device were out of memory.
hipError_t dev_malloc(void** p, size_t bytes)
{
    static std::mutex mu;
    static std::string armed;
    static long made = 0;
    if (const char* f = std::getenv("SFMM_FAULT_ALLOC")) {
        std::lock_guard<std::mutex> lock(mu);
        if (armed != f) {
            armed = f;
            made = 0;
        }
        if (made++ >= std::strtol(f, nullptr, 10)) {
            *p = nullptr;
            return hipErrorOutOfMemory;
        }
    } else {
        std::lock_guard<std::mutex> lock(mu);
        armed.clear();
    }
    return hipMalloc(p, bytes);
}
template <class T>
hipError_t dev_malloc(T** p, size_t bytes)
{
    void* v = nullptr;
    const hipError_t e = dev_malloc(&v, bytes);
    *p = static_cast<T*>(v);
    return e;
}

#define HIP_TRY(expr)                                                                            \
    do {                                                                                         \
        hipError_t e_ = (expr);                                                                  \
        if (e_ != hipSuccess) return fail(SFMM_ERR_HIP, std::string(#expr) + ": " + hipGetErrorString(e_)); \
    } while (0)

// An uninitialised host array: its pages are first touched by the threads that
// fill it (a zero-initialising std::vector of the plan's sizes spends most of
// its time in a single-threaded memset).
template <class T>
struct HostArray {
    std::unique_ptr<T[]> data;
    size_t size = 0;
    void resize(size_t n)
    {
        data.reset(new T[n]);
        size = n;
    }
    T& operator[](size_t i) { return data[i]; }
    const T& operator[](size_t i) const { return data[i]; }
};

// Host tables depend only on the expansion order: built once per order per
// process and shared by every plan (plans copy what they keep).
struct DenseHost {
    std::vector<uint16_t> ops;
    std::vector<float> az;
    std::vector<int32_t> grp;
    int op_exp = 0;
};

std::mutex g_tables_mu;
std::shared_ptr<const sfmm::Tables> g_tables[17];
std::shared_ptr<const DenseHost> g_dense[17];

std::shared_ptr<const sfmm::Tables> tables_for(int p)
{
    std::lock_guard<std::mutex> lock(g_tables_mu);
    if (!g_tables[p]) {
        auto t = std::make_shared<sfmm::Tables>();
        sfmm::build_tables(p, *t);
        g_tables[p] = t;
    }
    return g_tables[p];
}

std::shared_ptr<const DenseHost> dense_for(int p, const sfmm::Tables& t)
{
    std::lock_guard<std::mutex> lock(g_tables_mu);
    if (!g_dense[p]) {
        auto d = std::make_shared<DenseHost>();
        sfmm::build_dense_m2l(t, d->ops, d->op_exp, d->az, d->grp);
        g_dense[p] = d;
    }
    return g_dense[p];
}

}  // namespace

struct sfmm_tree_s {
    sfmm::Tree t;
    std::vector<double> pos;  // copy of the caller's positions
    // original index -> sorted slot, sorted slot -> global leaf (built on first use)
    std::vector<int32_t> iperm, slot_leaf;
    std::once_flag maps_once;

    void build_maps()
    {
        std::call_once(maps_once, [this] {
            iperm.assign(t.n, 0);
            for (int64_t s = 0; s < t.n; ++s) iperm[t.perm[s]] = static_cast<int32_t>(s);
            slot_leaf.assign(t.n, -1);
            for (size_t g = 0; g < t.leaf_node.size(); ++g) {
                const auto& lv = t.lv[t.leaf_level[g]];
                const int32_t a = t.leaf_node[g];
                for (int64_t s = lv.pbegin[a]; s < lv.pend[a]; ++s) slot_leaf[s] = static_cast<int32_t>(g);
            }
        });
    }
};

struct sfmm_plan_s {
    sfmm_tree_s* tree = nullptr;
    int device = 0;
    int p = 0, K = 0, precision = SFMM_FP64, max_ncharge = 1;
    hipStream_t stream = nullptr;
    sfmm::Tables tables;
    sfmm::DevPlan d;
    std::vector<void*> owned;
    std::vector<int64_t> level_first;  // gnode of the first node of each level (size depth+2)
    double* M = nullptr;
    double* L = nullptr;
    float* M32 = nullptr;  // FP32 copy of M for the mixed-precision M2L
    int nc_alloc = 0;
    double* qbuf = nullptr;
    double* obuf = nullptr;
    int qdim_alloc = 0, nout_alloc = 0;
    double* hQ = nullptr;  // device staging for host-array calls
    double* hU = nullptr;
    int64_t stage_q = 0, stage_u = 0;
    hipEvent_t ev[9] = {};
    bool timed = false;
    int m2l_kernel = 0;
    sfmm::P2PItems p2p;                // near-field work items
    int p2p_tpl = 2;                   // near-field tile shape: targets per lane (1, 2, 4)
    double* p2p_partial = nullptr;     // p2p.n_partial_rows x 4 doubles
    sfmm::M2LArgs m2l[5];              // offset-grouped M2L work per component count (built on first use)
    const double* rotA64 = nullptr;
    const double* coaxA64 = nullptr;
    const float* rotA32 = nullptr;
    const float* coaxA32 = nullptr;
    const double* zero_cols = nullptr;  // 4 (p+1)^2 zeros: M2L columns of absent sources
    const double* m2m_frag = nullptr;   // matrix-core M2M / L2L operators [8][nfrag][64]
    const double* l2l_frag = nullptr;
    sfmm::TransArgs m2m[5], l2l[5];     // per component count (built on first use)
    // mixed precision, p <= 12: M2L on the f16 matrix cores (m2l_dense.hpp)
    bool dense = false;
    sfmm::DenseArgs dns[5];             // per component count (built on first use)
    const void* dense_ops = nullptr;    // class operators, f16 hi/lo fragments
    const float* dense_az = nullptr;    // cos, sin of m phi per azimuth group
    const float* dense_zero = nullptr;  // zero columns of absent sources
    float* dense_nodemax = nullptr;     // largest |value| per node and component of the FP32 multipoles
    int dense_op_exp = 0;
    int dense_kd = 0;                   // floats per (node, component) of the FP32 copy
    int dense_waves = 8;                // the dense M2L's tile shape: wavefronts per workgroup (4, 8)
    std::vector<int32_t> dense_grp;     // azimuth group per offset id
    sfmm::SelfCorr self;                // self-list corrections (empty without a self generator)
    std::vector<int64_t> slist_row, slist_val;  // far self pairs per particle, original order (_slist)
    std::vector<int64_t> m2m_lvl[5], l2l_lvl[5];  // first block of each level (+ end)
    hipEvent_t* last_ev = nullptr;   // event set of the last evaluation
    std::vector<hipEvent_t> rec_ev;  // sets of 9 events, one set per recorded evaluation
    int rec_count = 0, rec_used = 0;

    ~sfmm_plan_s()
    {
        for (void* p : owned) (void)hipFree(p);
        (void)hipFree(M);
        (void)hipFree(L);
        (void)hipFree(M32);
        (void)hipFree(qbuf);
        (void)hipFree(obuf);
        (void)hipFree(hQ);
        (void)hipFree(hU);
        (void)hipFree(p2p_partial);
        for (auto& e : ev)
            if (e) (void)hipEventDestroy(e);
        for (auto& e : rec_ev) (void)hipEventDestroy(e);
        if (stream) (void)hipStreamDestroy(stream);
    }

    template <class T>
    int upload_ptr(const T* v, size_t count, const T** out)
    {
        *out = nullptr;
        if (count == 0) return SFMM_OK;
        void* p = nullptr;
        if (dev_malloc(&p, count * sizeof(T)) != hipSuccess) return fail(SFMM_ERR_ALLOC, "hipMalloc failed (plan tables)");
        owned.push_back(p);
        HIP_TRY(hipMemcpy(p, v, count * sizeof(T), hipMemcpyHostToDevice));
        *out = static_cast<const T*>(p);
        return SFMM_OK;
    }
    template <class T>
    int upload(const std::vector<T>& v, const T** out)
    {
        return upload_ptr(v.data(), v.size(), out);
    }
    template <class T>
    int upload(const HostArray<T>& v, const T** out)
    {
        return upload_ptr(v.data.get(), v.size, out);
    }
};

// A batch: the evaluations of several plans, spread over concurrent streams.
struct sfmm_batch_s {
    int device = 0;
    std::vector<sfmm_batch_item> items;
    std::vector<int> where;            // stream of each item
    std::vector<hipStream_t> streams;  // [0] is the batch's own stream
    std::vector<hipEvent_t> fork, join;
    int nstream = 1;
    bool own = false;               // every item on its plan's own stream
    std::vector<hipEvent_t> pjoin;      // per item, for `own`
    std::vector<hipEvent_t> pre, post;  // per item: its plan's stream before / after the item
    // the launch sequence captured once: a HIP graph replayed by every launch
    hipStream_t cap = nullptr;
    hipGraph_t graph = nullptr;
    hipGraphExec_t exec = nullptr;
    hipEvent_t done = nullptr;

    ~sfmm_batch_s()
    {
        if (exec) (void)hipGraphExecDestroy(exec);
        if (graph) (void)hipGraphDestroy(graph);
        if (done) (void)hipEventDestroy(done);
        if (cap) (void)hipStreamDestroy(cap);
        for (auto e : fork) (void)hipEventDestroy(e);
        for (auto e : join) (void)hipEventDestroy(e);
        for (auto e : pjoin) (void)hipEventDestroy(e);
        for (auto e : pre) (void)hipEventDestroy(e);
        for (auto e : post) (void)hipEventDestroy(e);
        for (auto st : streams) (void)hipStreamDestroy(st);
    }
};

namespace {

int mode_of(int kind, int ncharge, sfmm::Mode& mode, int& qdim, int& nc)
{
    switch (kind) {
    case SFMM_POTENTIAL:
        if (ncharge < 1 || ncharge > 4) return fail(SFMM_ERR_ARG, "ncharge must be in 1..4");
        mode = sfmm::MODE_POT;
        qdim = nc = ncharge;
        return SFMM_OK;
    case SFMM_CURL:
        if (ncharge != 3) return fail(SFMM_ERR_ARG, "the curl field needs 3-vector charges (fmm.hpp:550)");
        mode = sfmm::MODE_CURL;
        qdim = nc = 3;
        return SFMM_OK;
    case SFMM_POTGRAD:
        if (ncharge != 1) return fail(SFMM_ERR_ARG, "potential+gradient takes scalar charges");
        mode = sfmm::MODE_POTGRAD;
        qdim = 1;
        nc = 4;
        return SFMM_OK;
    default: return fail(SFMM_ERR_ARG, "unknown evaluation kind");
    }
}

int ensure_buffers(sfmm_plan_s* P, int nc, int qdim)
{
    const int64_t nnode = P->level_first.back();
    const int64_t n = P->tree->t.n;
    if (nc > P->nc_alloc) {
        (void)hipFree(P->M);
        (void)hipFree(P->L);
        (void)hipFree(P->M32);
        P->M = P->L = nullptr;
        P->M32 = nullptr;
        P->nc_alloc = 0;
        const size_t bytes = sizeof(double) * static_cast<size_t>(nnode) * nc * P->K;
        if (dev_malloc(&P->M, bytes) != hipSuccess || dev_malloc(&P->L, bytes) != hipSuccess)
            return fail(SFMM_ERR_ALLOC, "hipMalloc failed (expansions)");
        if (P->precision == SFMM_MIXED && P->m2l_kernel == 0) {
            // the dense layout pads each (node, component) to dense_kd floats
            // (the FP32 copy kernel writes the padding as zeros)
            const int row = P->dense ? P->dense_kd : P->K;
            const size_t b32 = sizeof(float) * static_cast<size_t>(nnode) * nc * row;
            if (dev_malloc(&P->M32, b32) != hipSuccess) return fail(SFMM_ERR_ALLOC, "hipMalloc failed (FP32 multipoles)");
            HIP_TRY(hipMemset(P->M32, 0, b32));
        }
        P->nc_alloc = nc;
    }
    if (qdim > P->qdim_alloc) {
        (void)hipFree(P->qbuf);
        P->qbuf = nullptr;
        if (dev_malloc(&P->qbuf, sizeof(double) * n * qdim) != hipSuccess)
            return fail(SFMM_ERR_ALLOC, "hipMalloc failed (charges)");
        P->qdim_alloc = qdim;
    }
    if (P->p2p.n_partial_rows > 0 && !P->p2p_partial) {
        if (dev_malloc(&P->p2p_partial, sizeof(double) * 4 * P->p2p.n_partial_rows) != hipSuccess)
            return fail(SFMM_ERR_ALLOC, "hipMalloc failed (near-field partial sums)");
    }
    if (nc > P->nout_alloc) {
        (void)hipFree(P->obuf);
        P->obuf = nullptr;
        if (dev_malloc(&P->obuf, sizeof(double) * n * nc) != hipSuccess)
            return fail(SFMM_ERR_ALLOC, "hipMalloc failed (outputs)");
        P->nout_alloc = nc;
    }
    return SFMM_OK;
}

// Targets of each level grouped by child octant into blocks of TB (per level,
// the q-th chunks of the 8 octants, children of nearby parents, are adjacent);
// per block, the offsets that occur with any of its targets, in `order`, each
// with meta(offset) and the source of every target (-1: none).  Blocks are cut
// on all host threads; fails with more than 189 offsets in a block.
int cut_m2l_blocks(sfmm_plan_s* P, int TB, const std::vector<int>& order, const std::function<int32_t(int)>& meta,
                   HostArray<int32_t>& blk_tgt, std::vector<int64_t>& blk_row, HostArray<int32_t>& ent_meta,
                   HostArray<int32_t>& ent_src)
{
    const sfmm::Tree& t = P->tree->t;
    struct Blk {
        int level, octant;
        int64_t first;  // index of the first target in the level's octant list
    };
    std::vector<std::vector<int32_t>> oct_nodes(static_cast<size_t>(8) * (t.depth + 1));
    std::vector<Blk> blocks;
    for (int l = 2; l <= t.depth; ++l) {
        const auto& lv = t.lv[l];
        for (int32_t a = 0; a < lv.n_node(); ++a) oct_nodes[8 * l + (lv.code[a] & 7)].push_back(a);
        size_t nchunk = 0;
        for (int o = 0; o < 8; ++o) nchunk = std::max(nchunk, (oct_nodes[8 * l + o].size() + TB - 1) / TB);
        for (size_t q = 0; q < nchunk; ++q)
            for (int o = 0; o < 8; ++o)
                if (q * TB < oct_nodes[8 * l + o].size()) blocks.push_back({l, o, static_cast<int64_t>(q * TB)});
    }
    const int64_t nb = static_cast<int64_t>(blocks.size());
    const int nth = host_threads();
    std::vector<char> in_order(343, 0);
    for (int off : order) in_order[off] = 1;
    // pass 1: the offsets present in each block; pass 2 writes every block's
    // entries at its place in the final arrays (both over host threads)
    std::vector<int64_t> cnt(nb, 0);
    parallel_ranges(nb, nth, [&](int64_t b0, int64_t b1, int) {
        std::vector<char> used(343);
        for (int64_t bi = b0; bi < b1; ++bi) {
            const Blk& B = blocks[bi];
            const auto& lv = t.lv[B.level];
            const auto& nodes = oct_nodes[8 * B.level + B.octant];
            std::fill(used.begin(), used.end(), 0);
            int64_t c = 0;
            for (int tt = 0; tt < TB && B.first + tt < static_cast<int64_t>(nodes.size()); ++tt) {
                const int32_t a = nodes[B.first + tt];
                for (int64_t k = lv.il_row[a]; k < lv.il_row[a + 1]; ++k) {
                    const int off = lv.il_off[k];
                    if (!used[off] && in_order[off]) ++c;
                    used[off] = 1;
                }
            }
            cnt[bi] = c;
        }
    });
    blk_row.assign(static_cast<size_t>(nb) + 1, 0);
    for (int64_t b = 0; b < nb; ++b) {
        if (cnt[b] > 189) return fail(SFMM_ERR_STATE, "M2L block with more than 189 interaction offsets");
        blk_row[b + 1] = blk_row[b] + cnt[b];
    }
    blk_tgt.resize(static_cast<size_t>(nb) * TB);
    ent_meta.resize(static_cast<size_t>(blk_row[nb]));
    ent_src.resize(static_cast<size_t>(blk_row[nb]) * TB);
    parallel_ranges(nb, nth, [&](int64_t b0, int64_t b1, int) {
        std::vector<int32_t> src(static_cast<size_t>(343) * TB);
        std::vector<char> used(343, 0);
        for (int64_t bi = b0; bi < b1; ++bi) {
            const Blk& B = blocks[bi];
            const auto& lv = t.lv[B.level];
            const auto& nodes = oct_nodes[8 * B.level + B.octant];
            for (int tt = 0; tt < TB; ++tt) {
                if (B.first + tt >= static_cast<int64_t>(nodes.size())) {
                    blk_tgt[bi * TB + tt] = -1;
                    continue;
                }
                const int32_t a = nodes[B.first + tt];
                blk_tgt[bi * TB + tt] = static_cast<int32_t>(P->level_first[B.level] + a);
                for (int64_t k = lv.il_row[a]; k < lv.il_row[a + 1]; ++k) {
                    const int off = lv.il_off[k];
                    int32_t* row = src.data() + static_cast<size_t>(off) * TB;
                    if (!used[off]) {
                        std::fill(row, row + TB, -1);
                        used[off] = 1;
                    }
                    row[tt] = static_cast<int32_t>(P->level_first[B.level] + lv.il_src[k]);
                }
            }
            int64_t e = blk_row[bi];
            for (int off : order) {
                if (!used[off]) continue;
                ent_meta[e] = meta(off);
                std::copy(src.begin() + static_cast<size_t>(off) * TB, src.begin() + static_cast<size_t>(off + 1) * TB,
                          ent_src.data.get() + e * TB);
                ++e;
            }
            for (int off = 0; off < 343; ++off) used[off] = 0;
        }
    });
    return SFMM_OK;
}

// Blocks of the mixed-precision M2L on the f16 matrix cores: 32 targets per
// wave group (4 / waves-per-group groups per block), offsets in (azimuth
// group, polar class, id) order so that each azimuth group is one run.
int build_dense_blocks(sfmm_plan_s* P, int nc)
{
    if (P->dns[nc].blk_row) return SFMM_OK;
    // targets per block: 16 per group of waves, one wave per component
    const int ncw = nc == 3 ? 4 : nc;
    const int TB = 16 * (P->dense_waves / ncw);
    const auto& cls = P->tables.off_class;
    const auto& grp = P->dense_grp;
    std::vector<int> order;
    for (int off = 0; off < 343; ++off)
        if (cls[off] >= 0) order.push_back(off);
    std::sort(order.begin(), order.end(), [&](int x, int y) {
        return grp[x] != grp[y] ? grp[x] < grp[y] : cls[x] != cls[y] ? cls[x] < cls[y] : x < y;
    });
    HostArray<int32_t> blk_tgt, ent_meta, ent_src;
    std::vector<int64_t> blk_row;
    if (int rc = cut_m2l_blocks(P, TB, order, [&](int off) { return cls[off] | grp[off] << 8; }, blk_tgt, blk_row,
                                ent_meta, ent_src))
        return rc;
    sfmm::DenseArgs A;
    A.n_block = static_cast<int64_t>(blk_row.size()) - 1;
    int rc;
    if ((rc = P->upload(blk_tgt, &A.blk_tgt)) || (rc = P->upload(blk_row, &A.blk_row)) ||
        (rc = P->upload(ent_meta, &A.ent_meta)) || (rc = P->upload(ent_src, &A.ent_src)))
        return rc;
    A.ops = P->dense_ops;
    A.az = P->dense_az;
    A.nodemax = P->dense_nodemax;
    A.zero = P->dense_zero;
    A.op_exp = P->dense_op_exp;
    A.waves = P->dense_waves;
    P->dns[nc] = A;
    return SFMM_OK;
}

// Sets the dense M2L's wavefronts per workgroup; its work lists (cut for the
// previous shape) are rebuilt on next use.  The previous lists stay allocated
// until the plan is destroyed: a batch created earlier may have captured them.
void set_dense_waves(sfmm_plan_s* P, int waves)
{
    if (P->dense_waves == waves) return;
    for (auto& A : P->dns) A = sfmm::DenseArgs{};
    P->dense_waves = waves;
}

// Groups the targets of each level by child octant into blocks of 16/nc and
// lists, per block, the offsets that occur with the source of every target.
int build_m2l_blocks(sfmm_plan_s* P, int nc)
{
    // built into a local and published only when every upload succeeded, so a
    // failed attempt leaves the plan without blocks and the next one retries
    if (P->m2l[nc].blk_row) return SFMM_OK;
    sfmm::M2LArgs A;
    const int TB = 16 / nc;
    // offsets in (polar class, distance class, id) order: consecutive offsets
    // share the rotation operators, and mostly the coaxial one as well
    const auto& cls = P->tables.off_class;
    const auto& d2 = P->tables.off_d2;
    std::vector<int> order;
    for (int off = 0; off < 343; ++off)
        if (cls[off] >= 0) order.push_back(off);
    std::sort(order.begin(), order.end(), [&](int x, int y) {
        return cls[x] != cls[y] ? cls[x] < cls[y] : d2[x] != d2[y] ? d2[x] < d2[y] : x < y;
    });
    HostArray<int32_t> blk_tgt, ent_meta, ent_src;
    std::vector<int64_t> blk_row;
    if (int rc = cut_m2l_blocks(P, TB, order, [&](int off) { return off | cls[off] << 9 | d2[off] << 16; }, blk_tgt,
                                blk_row, ent_meta, ent_src))
        return rc;
    A.n_block = static_cast<int64_t>(blk_row.size()) - 1;
    int rc;
    if ((rc = P->upload(blk_tgt, &A.blk_tgt)) || (rc = P->upload(blk_row, &A.blk_row)) ||
        (rc = P->upload(ent_meta, &A.ent_meta)) || (rc = P->upload(ent_src, &A.ent_src)))
        return rc;
    A.fp32 = P->precision == SFMM_MIXED;
    A.rotA = A.fp32 ? static_cast<const void*>(P->rotA32) : static_cast<const void*>(P->rotA64);
    A.coaxA = A.fp32 ? static_cast<const void*>(P->coaxA32) : static_cast<const void*>(P->coaxA64);
    A.off_cs = P->d.off_cs;
    A.zero = P->zero_cols;
    P->m2l[nc] = A;
    return SFMM_OK;
}

// M2M and L2L work lists of every level: blocks of 16/nc parents, each with one
// entry per child octant present (the children of that octant).  M2M at parent
// level l and L2L into child level l+1 run over the same blocks.
int build_translate_blocks(sfmm_plan_s* P, int nc)
{
    if (P->m2m[nc].blk_row) return SFMM_OK;
    const sfmm::Tree& t = P->tree->t;
    const int TB = 16 / nc;
    std::vector<int64_t> row{0};
    std::vector<int32_t> parent_of_block, oct, child;
    auto& lv_m = P->m2m_lvl[nc];
    auto& lv_l = P->l2l_lvl[nc];
    lv_m.assign(t.depth + 2, 0);
    lv_l.assign(t.depth + 2, 0);
    for (int l = 0; l <= t.depth; ++l) {
        const auto& lv = t.lv[l];
        lv_m[l] = static_cast<int64_t>(row.size()) - 1;
        if (l == t.depth) continue;
        std::vector<int32_t> parents;
        for (int32_t a = 0; a < lv.n_node(); ++a)
            if (lv.nchild[a] > 0) parents.push_back(a);
        for (size_t b = 0; b < parents.size(); b += TB) {
            int32_t ch[8][16];
            for (auto& r : ch) std::fill(r, r + 16, -1);
            bool present[8] = {};
            for (int tt = 0; tt < TB; ++tt) {
                if (b + tt >= parents.size()) {
                    parent_of_block.push_back(-1);
                    continue;
                }
                const int32_t a = parents[b + tt];
                parent_of_block.push_back(static_cast<int32_t>(P->level_first[l] + a));
                const auto& cl = t.lv[l + 1];
                for (int c = 0; c < lv.nchild[a]; ++c) {
                    const int32_t ci = lv.child0[a] + c;
                    const int o = static_cast<int>(cl.code[ci] & 7);
                    ch[o][tt] = static_cast<int32_t>(P->level_first[l + 1] + ci);
                    present[o] = true;
                }
            }
            for (int o = 0; o < 8; ++o) {
                if (!present[o]) continue;
                oct.push_back(o);
                child.insert(child.end(), ch[o], ch[o] + TB);
            }
            row.push_back(static_cast<int64_t>(oct.size()));
        }
    }
    lv_m[t.depth + 1] = static_cast<int64_t>(row.size()) - 1;
    for (int l = 1; l <= t.depth + 1; ++l) lv_l[l] = lv_m[l - 1];  // children at level l: parents at l-1
    sfmm::TransArgs A;
    int rc;
    if ((rc = P->upload(row, &A.blk_row)) || (rc = P->upload(parent_of_block, &A.blk_node)) ||
        (rc = P->upload(oct, &A.ent_oct)) || (rc = P->upload(child, &A.ent_node)))
        return rc;
    sfmm::TransArgs B = A;  // published together once every upload succeeded
    A.frag = P->m2m_frag;
    B.frag = P->l2l_frag;
    P->m2m[nc] = A;
    P->l2l[nc] = B;
    return SFMM_OK;
}

sfmm::LevelSpan level_blocks(const std::vector<int64_t>& first, int l)
{
    return {first[l], first[l + 1] - first[l]};
}

// Checks the arguments of an evaluation and makes every buffer and work list it
// needs (all allocation and upload happens here, none in run_on).
int prepare(sfmm_plan_s* P, int kind, int ncharge, int which, sfmm::Mode& mode, int& qdim, int& nc)
{
    if (int rc = mode_of(kind, ncharge, mode, qdim, nc)) return rc;
    if ((which & ~(SFMM_BOTH | SFMM_SELF_LIST)) != 0 || (which & SFMM_BOTH) == 0)
        return fail(SFMM_ERR_ARG, "which must be SFMM_FAR, SFMM_NEAR or SFMM_BOTH, optionally | SFMM_SELF_LIST");
    if (mode == sfmm::MODE_POT && ncharge > P->max_ncharge)
        return fail(SFMM_ERR_ARG, "ncharge exceeds the plan's max_ncharge");
    HIP_TRY(hipSetDevice(P->device));
    if (int rc = ensure_buffers(P, nc, qdim)) return rc;
    if (P->m2l_kernel == 0) {
        if (int rc = P->dense ? build_dense_blocks(P, nc) : build_m2l_blocks(P, nc)) return rc;
        if (int rc = build_translate_blocks(P, nc)) return rc;
    }
    return SFMM_OK;
}

int run_on(sfmm_plan_s* P, sfmm::Mode mode, int nc, int qdim, const double* dQ, double* dU, int which,
           hipStream_t s, bool timed);

int run(sfmm_plan_s* P, int kind, int ncharge, const double* dQ, double* dU, int which)
{
    sfmm::Mode mode;
    int qdim, nc;
    if (int rc = prepare(P, kind, ncharge, which, mode, qdim, nc)) return rc;
    return run_on(P, mode, nc, qdim, dQ, dU, which, P->stream, true);
}

// The evaluation's launches on stream s; `timed` records the stage events
// (off while a graph is being captured).
int run_on(sfmm_plan_s* P, sfmm::Mode mode, int nc, int qdim, const double* dQ, double* dU, int which,
           hipStream_t s, bool timed)
{
    const sfmm::DevPlan& d = P->d;
    const int depth = P->tree->t.depth;
    hipEvent_t* ev = P->ev;
    if (timed && P->rec_used < P->rec_count) ev = &P->rec_ev[9 * static_cast<size_t>(P->rec_used++)];
#define REC(k) \
    if (timed) HIP_TRY(hipEventRecord(ev[k], s))
    REC(0);
    sfmm::launch_gather_charges(d, qdim, dQ, P->qbuf, s);
    sfmm::launch_fill(P->obuf, static_cast<int64_t>(nc) * d.n, 0.0, s);
    REC(1);
    const bool far = (which & SFMM_FAR) && depth >= 2;  // fmm.hpp:816-820
    // the mixed M2L reads an FP32 copy of the multipoles (levels >= 2), made in
    // one pass before it (P2M and M2M write only the FP64 multipoles)
    float* M32 = P->m2l_kernel == 0 && (P->dense || P->m2l[nc].fp32) ? P->M32 : nullptr;
    if (far) sfmm::launch_p2m(d, mode, nc, P->qbuf, P->M, s);
    REC(2);
    const sfmm::TransArgs& m2m = P->m2m[nc];
    if (far)
        for (int l = depth - 1; l >= 2; --l) {
            if (P->m2l_kernel == 0)
                sfmm::launch_translate(m2m, level_blocks(P->m2m_lvl[nc], l), P->p, nc, false, P->M, s);
            else
                sfmm::launch_m2m(d, {P->level_first[l], P->level_first[l + 1] - P->level_first[l]}, nc, P->M, s);
        }
    REC(3);
    if (far) {
        if (M32) {
            // and, for the dense M2L, the range of every FP32 multipole, which scales the columns
            const int64_t f2 = P->level_first[2];
            const int row = P->dense ? P->dense_kd : P->K;
            sfmm::launch_fp32_copy(P->M + f2 * nc * P->K, P->level_first.back() - f2, nc, P->p, M32 + f2 * nc * row,
                                   P->dense ? P->dense_nodemax + f2 * nc : nullptr, s);
        }
        if (P->dense) {
            sfmm::launch_m2l_dense(P->dns[nc], P->p, nc, M32, P->L, s);
        } else if (P->m2l_kernel == 0) {
            sfmm::M2LArgs& A = P->m2l[nc];
            if (A.fp32) A.M32 = P->M32;
            sfmm::launch_m2l_mfma(A, P->p, nc, P->M, P->L, s);
        }
        else sfmm::launch_m2l(d, nc, P->M, P->L, s);
    }
    REC(4);
    if (far)
        for (int l = 3; l <= depth; ++l) {
            if (P->m2l_kernel == 0)
                sfmm::launch_translate(P->l2l[nc], level_blocks(P->l2l_lvl[nc], l), P->p, nc, true, P->L, s);
            else
                sfmm::launch_l2l(d, {P->level_first[l], P->level_first[l + 1] - P->level_first[l]}, nc, P->L, s);
        }
    REC(5);
    if (far) sfmm::launch_l2p(d, nc, P->L, P->obuf, s);
    REC(6);
    if (which & SFMM_NEAR) {
        sfmm::launch_p2p(d, P->p2p, mode, nc, P->qbuf, P->obuf, P->p2p_partial, s, P->precision == SFMM_MIXED,
                         P->p2p_tpl);
        sfmm::launch_self_correction(d, P->self, mode, nc, (which & SFMM_SELF_LIST) != 0, P->qbuf, P->obuf, s);
    }
    REC(7);
    sfmm::launch_scatter_outputs(d, nc, P->obuf, dU, s);
    REC(8);
#undef REC
    HIP_TRY(hipGetLastError());
    if (timed) {
        P->last_ev = ev;
        P->timed = true;
    }
    return SFMM_OK;
}


constexpr int kBatchMaxStreams = 8;

// Items to streams, longest first (cost ~ particles x coefficients).
std::vector<int> batch_assign(const std::vector<sfmm_batch_item>& items, int ns)
{
    std::vector<int> order(items.size()), out(items.size(), 0);
    std::vector<double> cost(items.size()), load(ns, 0.0);
    for (size_t i = 0; i < items.size(); ++i) {
        const sfmm_plan_s* P = items[i].plan;
        cost[i] = static_cast<double>(P->tree->t.n) * P->K;
        order[i] = static_cast<int>(i);
    }
    std::stable_sort(order.begin(), order.end(), [&](int a, int b) { return cost[a] > cost[b]; });
    for (int i : order) {
        const int k = static_cast<int>(std::min_element(load.begin(), load.end()) - load.begin());
        out[i] = k;
        load[k] += cost[i];
    }
    return out;
}

// Issues the items' evaluations on ns streams (own: each on its plan's stream)
// after the work already queued on `origin`, and makes `origin` wait for all of them.
// plan_order: each item also runs after the work queued on its plan's own stream,
// and that stream's later work after the item (off while capturing: the plans'
// streams are not part of the capture, and a captured launch orders them itself).
int batch_issue(sfmm_batch_s* B, const std::vector<sfmm_batch_item>& items, const std::vector<int>& where, int ns,
                hipStream_t origin, bool own = false, bool plan_order = true)
{
    HIP_TRY(hipEventRecord(B->fork[0], origin));
    if (own) {
        for (size_t i = 0; i < items.size(); ++i) {
            const sfmm_batch_item& it = items[i];
            HIP_TRY(hipStreamWaitEvent(it.plan->stream, B->fork[0], 0));
            sfmm::Mode mode;
            int qdim, nc;
            if (int r = mode_of(it.kind, it.ncharge, mode, qdim, nc)) return r;
            if (int r = run_on(it.plan, mode, nc, qdim, it.dQ, it.dU, it.which, it.plan->stream, false)) return r;
            HIP_TRY(hipEventRecord(B->pjoin[i], it.plan->stream));
            HIP_TRY(hipStreamWaitEvent(origin, B->pjoin[i], 0));
        }
        return SFMM_OK;
    }
    for (int k = 0; k < ns; ++k) HIP_TRY(hipStreamWaitEvent(B->streams[k], B->fork[0], 0));
    for (size_t i = 0; i < items.size(); ++i) {
        const sfmm_batch_item& it = items[i];
        sfmm::Mode mode;
        int qdim, nc;
        if (int r = mode_of(it.kind, it.ncharge, mode, qdim, nc)) return r;
        // the plan's buffers are shared with its own stream: the item runs after the
        // work already queued there, and later work there runs after the item
        hipStream_t bs = B->streams[where[i]];
        if (plan_order) {
            HIP_TRY(hipEventRecord(B->pre[i], it.plan->stream));
            HIP_TRY(hipStreamWaitEvent(bs, B->pre[i], 0));
        }
        if (int r = run_on(it.plan, mode, nc, qdim, it.dQ, it.dU, it.which, bs, false)) return r;
        if (plan_order) {
            HIP_TRY(hipEventRecord(B->post[i], bs));
            HIP_TRY(hipStreamWaitEvent(it.plan->stream, B->post[i], 0));
        }
    }
    for (int k = 0; k < ns; ++k) {
        HIP_TRY(hipEventRecord(B->join[k], B->streams[k]));
        HIP_TRY(hipStreamWaitEvent(origin, B->join[k], 0));
    }
    return SFMM_OK;
}

// Captures the batch's launch sequence (its stream assignment) into a HIP graph
// and instantiates it.  Inside the capture there are only kernel launches and
// event records and waits between the capture stream and the worker streams it
// forks (or, for `own`, the plans' streams, which join through the fork): every
// buffer and work list exists beforehand, the fill kernel clears memory, and no
// event recorded outside the capture is waited for.  The capture is always
// ended, also after a failure; every status is checked.
int batch_capture(sfmm_batch_s* B, const std::vector<sfmm_batch_item>& items, const std::vector<int>& where,
                  int ns, bool own, hipGraph_t* graph, hipGraphExec_t* exec)
{
    *graph = nullptr;
    *exec = nullptr;
    HIP_TRY(hipStreamBeginCapture(B->cap, hipStreamCaptureModeRelaxed));
    const int rc = batch_issue(B, items, where, ns, B->cap, own, false);
    hipGraph_t g = nullptr;
    const hipError_t e = hipStreamEndCapture(B->cap, &g);
    if (rc != SFMM_OK || e != hipSuccess || !g) {
        if (g) (void)hipGraphDestroy(g);
        if (rc != SFMM_OK) return rc;
        return fail(SFMM_ERR_HIP, std::string("batch graph capture: ") + hipGetErrorString(e));
    }
    hipGraphExec_t x = nullptr;
    const hipError_t e2 = hipGraphInstantiate(&x, g, nullptr, nullptr, 0);
    if (e2 != hipSuccess || !x) {
        (void)hipGraphDestroy(g);
        return fail(SFMM_ERR_HIP, std::string("batch graph instantiate: ") + hipGetErrorString(e2));
    }
    *graph = g;
    *exec = x;
    return SFMM_OK;
}

// Times the tile shapes of plan P for an evaluation `it` (its dU is scratch) on
// stream s and keeps the fastest: the near field's targets per lane (1, 2, 4;
// near field alone) when `near`, the dense M2L's wavefronts per workgroup (8, 4;
// far field alone) when `far` and the plan uses it.  Each shape is timed as the
// best of three runs after a warm-up; a shape other than the default (2 and 8)
// is kept only when at least 2% faster.  The plan must be prepared for `it`.
int tune_tiles(sfmm_plan_s* P, const sfmm_batch_item& it, hipStream_t s, bool near, bool far, int* tpl_out,
               int* waves_out)
{
    sfmm::Mode mode;
    int qdim, nc;
    if (int rc = mode_of(it.kind, it.ncharge, mode, qdim, nc)) return rc;
    hipEvent_t t0 = nullptr, t1 = nullptr;
    HIP_TRY(hipEventCreate(&t0));
    HIP_TRY(hipEventCreate(&t1));
    int rc = SFMM_OK;
    auto timed = [&](int which, float* best) {
        *best = -1.0f;
        for (int r = 0; r < 4 && rc == SFMM_OK; ++r) {
            float ms = 0.0f;
            hipError_t e = hipEventRecord(t0, s);
            if (e == hipSuccess) rc = run_on(P, mode, nc, qdim, it.dQ, it.dU, which, s, false);
            if (rc != SFMM_OK) return;
            if (e == hipSuccess) e = hipEventRecord(t1, s);
            if (e == hipSuccess) e = hipEventSynchronize(t1);
            if (e == hipSuccess) e = hipEventElapsedTime(&ms, t0, t1);
            if (e != hipSuccess) {
                rc = fail(SFMM_ERR_HIP, std::string("tile tuning: ") + hipGetErrorString(e));
                return;
            }
            if (r > 0 && (*best < 0.0f || ms < *best)) *best = ms;  // run 0 warms up
        }
    };
    if (near) {
        int best_tpl = 2;
        float best_ms = -1.0f;
        for (int tpl : {2, 1, 4}) {
            P->p2p_tpl = tpl;
            float ms = 0.0f;
            timed(SFMM_NEAR, &ms);
            if (rc != SFMM_OK) break;
            if (best_ms < 0.0f || ms < 0.98f * best_ms) {
                best_ms = ms;
                best_tpl = tpl;
            }
        }
        P->p2p_tpl = best_tpl;
        if (tpl_out) *tpl_out = best_tpl;
    }
    if (far && P->dense && rc == SFMM_OK) {
        int best_w = 8;
        float best_ms = -1.0f;
        for (int w : {8, 4}) {
            set_dense_waves(P, w);
            if ((rc = prepare(P, it.kind, it.ncharge, it.which, mode, qdim, nc)) != SFMM_OK) break;
            float ms = 0.0f;
            timed(SFMM_FAR, &ms);
            if (rc != SFMM_OK) break;
            if (best_ms < 0.0f || ms < 0.98f * best_ms) {
                best_ms = ms;
                best_w = w;
            }
        }
        if (rc == SFMM_OK) {
            set_dense_waves(P, best_w);
            rc = prepare(P, it.kind, it.ncharge, it.which, mode, qdim, nc);
        }
        if (waves_out) *waves_out = best_w;
    }
    (void)hipEventDestroy(t0);
    (void)hipEventDestroy(t1);
    return rc;
}

// Replays a captured batch on `origin`: after the work queued on `origin` and on
// every plan's own stream, and before the later work on both.
int batch_replay(sfmm_batch_s* B, hipGraphExec_t exec, const std::vector<sfmm_batch_item>& items,
                 hipStream_t origin)
{
    for (size_t i = 0; i < items.size(); ++i) {
        HIP_TRY(hipEventRecord(B->pre[i], items[i].plan->stream));
        HIP_TRY(hipStreamWaitEvent(origin, B->pre[i], 0));
    }
    HIP_TRY(hipGraphLaunch(exec, origin));
    HIP_TRY(hipEventRecord(B->done, origin));
    for (const auto& it : items) HIP_TRY(hipStreamWaitEvent(it.plan->stream, B->done, 0));
    return SFMM_OK;
}

}  // namespace

namespace {
struct Fnv {
    uint64_t h = 1469598103934665603ULL;
    void bytes(const void* p, size_t n)
    {
        const unsigned char* c = static_cast<const unsigned char*>(p);
        for (size_t i = 0; i < n; ++i) h = (h ^ c[i]) * 1099511628211ULL;
    }
    template <class T>
    void vec(const std::vector<T>& v)
    {
        const uint64_t n = v.size();
        bytes(&n, sizeof n);
        bytes(v.data(), v.size() * sizeof(T));
    }
};
}  // namespace

extern "C" {

const char* sfmm_last_error(void) { return g_last_error.c_str(); }

int sfmm_box_from_positions(int64_t n, const double* positions, int vertical_axis, double phideg, double out4[4])
{
    if (!positions || !out4 || n <= 0) return fail(SFMM_ERR_ARG, "null argument or empty positions");
    if (vertical_axis < 0 || vertical_axis > 2) return fail(SFMM_ERR_ARG, "vertical_axis must be 0, 1 or 2");
    const sfmm::BoxGeom b = sfmm::make_box(n, positions, vertical_axis, phideg);
    out4[0] = b.width;
    for (int d = 0; d < 3; ++d) out4[1 + d] = b.corner[d];
    return SFMM_OK;
}

int sfmm_tree_create(const sfmm_tree_desc* desc, sfmm_tree* out)
{
    if (!desc || !out || !desc->positions) return fail(SFMM_ERR_ARG, "null argument");
    *out = nullptr;
    if (desc->n <= 0) return fail(SFMM_ERR_ARG, "particle count must be positive");
    if (desc->vertical_axis < 0 || desc->vertical_axis > 2) return fail(SFMM_ERR_ARG, "vertical_axis must be 0, 1 or 2");
    auto T = std::make_unique<sfmm_tree_s>();
    T->pos.assign(desc->positions, desc->positions + 3 * desc->n);
    sfmm::BoxGeom box = desc->box_given
                            ? sfmm::make_box_explicit(desc->corner, desc->width, desc->vertical_axis, desc->phideg)
                            : sfmm::make_box(desc->n, T->pos.data(), desc->vertical_axis, desc->phideg);
    int threads = desc->threads > 0 ? desc->threads : static_cast<int>(std::thread::hardware_concurrency());
    if (threads <= 0) threads = 1;
    std::string err =
        desc->on_device
            ? sfmm::build_tree_device(T->t, desc->n, T->pos.data(), box, desc->max_level, desc->max_per_leaf)
            : sfmm::build_tree(T->t, desc->n, T->pos.data(), box, desc->max_level, desc->max_per_leaf, threads);
    if (!err.empty()) {
        // the device build reports HIP failures as "<call>: <HIP error>"
        const int code = err.find("out of device memory") != std::string::npos ? SFMM_ERR_ALLOC
                         : err.rfind("hip", 0) == 0                            ? SFMM_ERR_HIP
                                                                                : SFMM_ERR_ARG;
        return fail(code, err);
    }
    *out = T.release();
    return SFMM_OK;
}

int sfmm_tree_destroy(sfmm_tree tree)
{
    delete tree;
    return SFMM_OK;
}

int sfmm_tree_digest(sfmm_tree tree, uint64_t* digest)
{
    if (!tree || !digest) return fail(SFMM_ERR_ARG, "null argument");
    const sfmm::Tree& t = tree->t;
    Fnv f;
    const int64_t head[4] = {t.n, t.max_level, t.cap, t.depth};
    f.bytes(head, sizeof head);
    f.vec(t.perm);
    for (const sfmm::TreeLevel& l : t.lv) {
        f.vec(l.code), f.vec(l.pbegin), f.vec(l.pend), f.vec(l.parent), f.vec(l.child0), f.vec(l.nchild);
        f.vec(l.node2leaf), f.vec(l.leaf2node), f.vec(l.nb_row), f.vec(l.nb), f.vec(l.il_row), f.vec(l.il_src);
        f.vec(l.il_off);
    }
    f.vec(t.leaf_level), f.vec(t.leaf_node), f.vec(t.leaf_row), f.vec(t.near_row), f.vec(t.near_begin);
    f.vec(t.near_end);
    *digest = f.h;
    return SFMM_OK;
}

int sfmm_tree_near_sources(sfmm_tree tree, int64_t i, int64_t* out, int64_t capacity, int64_t* count)
{
    if (!tree || !count) return fail(SFMM_ERR_ARG, "null argument");
    const sfmm::Tree& t = tree->t;
    if (i < 0 || i >= t.n) return fail(SFMM_ERR_ARG, "particle index out of range");
    tree->build_maps();
    const int32_t g = tree->slot_leaf[tree->iperm[i]];
    int64_t cnt = 0;
    for (int64_t r = t.near_row[g]; r < t.near_row[g + 1]; ++r)
        for (int64_t sl = t.near_begin[r]; sl < t.near_end[r]; ++sl, ++cnt)
            if (out && cnt < capacity) out[cnt] = t.perm[sl];
    *count = cnt;
    if (out && cnt > capacity) return fail(SFMM_ERR_ARG, "capacity is smaller than the number of near sources");
    return SFMM_OK;
}

int sfmm_tree_get_info(sfmm_tree tree, sfmm_tree_info* info)
{
    if (!tree || !info) return fail(SFMM_ERR_ARG, "null argument");
    const sfmm::TreeStats st = sfmm::tree_stats(tree->t);
    info->depth = tree->t.depth;
    info->n_node = st.n_node;
    info->n_leaf = st.n_leaf;
    info->m2l_pairs = st.m2l_pairs;
    info->near_pairs = st.near_pairs;
    info->width = tree->t.box.width;
    for (int d = 0; d < 3; ++d) info->corner[d] = tree->t.box.corner[d];
    return SFMM_OK;
}

int sfmm_tree_level_counts(sfmm_tree tree, int level, int64_t counts[6])
{
    if (!tree || !counts) return fail(SFMM_ERR_ARG, "null argument");
    if (level < 0 || level > tree->t.depth) return fail(SFMM_ERR_ARG, "level out of range");
    sfmm::RefLists r;
    sfmm::reference_lists(tree->t, level, r);
    const auto& lv = tree->t.lv[level];
    counts[0] = lv.n_node();
    counts[1] = lv.n_leaf();
    counts[2] = static_cast<int64_t>(r.idx_val.size());
    counts[3] = static_cast<int64_t>(r.ilist_val.size() / 4);
    counts[4] = static_cast<int64_t>(r.clist_val.size() / 2);
    counts[5] = static_cast<int64_t>(r.nlist_val.size() / 2);
    return SFMM_OK;
}

int sfmm_tree_level_lists(sfmm_tree tree, int level, int64_t* c_node, int64_t* leaf2node, int64_t* idx_row,
                          int64_t* idx_val, int64_t* ilist_row, int64_t* ilist_val, int64_t* clist_row,
                          int64_t* clist_val, int64_t* plist, int64_t* nlist_row, int64_t* nlist_val)
{
    if (!tree) return fail(SFMM_ERR_ARG, "null argument");
    if (level < 0 || level > tree->t.depth) return fail(SFMM_ERR_ARG, "level out of range");
    sfmm::RefLists r;
    sfmm::reference_lists(tree->t, level, r);
    auto put = [](int64_t* dst, const std::vector<int64_t>& v) {
        if (dst && !v.empty()) std::memcpy(dst, v.data(), v.size() * sizeof(int64_t));
    };
    put(c_node, r.c_node);
    put(leaf2node, r.leaf2node);
    put(idx_row, r.idx_row);
    put(idx_val, r.idx_val);
    put(ilist_row, r.ilist_row);
    put(ilist_val, r.ilist_val);
    put(clist_row, r.clist_row);
    put(clist_val, r.clist_val);
    put(plist, r.plist);
    put(nlist_row, r.nlist_row);
    put(nlist_val, r.nlist_val);
    return SFMM_OK;
}

int sfmm_plan_create(sfmm_tree tree, const sfmm_plan_desc* desc, sfmm_plan* out)
{
    if (!tree || !desc || !out) return fail(SFMM_ERR_ARG, "null argument");
    *out = nullptr;
    if (desc->p < 1 || desc->p > 16) return fail(SFMM_ERR_ARG, "expansion order p must be in 1..16");
    if (desc->precision != SFMM_FP64 && desc->precision != SFMM_MIXED) return fail(SFMM_ERR_ARG, "unknown precision");
    if (desc->m2l_kernel != 0 && desc->m2l_kernel != 1) return fail(SFMM_ERR_ARG, "m2l_kernel must be 0 or 1");
    if (desc->precision == SFMM_MIXED && desc->m2l_kernel != 0)
        return fail(SFMM_ERR_ARG, "mixed precision runs on the MFMA M2L kernel");
    if ((desc->self_row == nullptr) != (desc->self_val == nullptr))
        return fail(SFMM_ERR_ARG, "self_row and self_val must be given together");
    if (desc->max_ncharge < 1 || desc->max_ncharge > 4) return fail(SFMM_ERR_ARG, "max_ncharge must be in 1..4");
    auto P = std::make_unique<sfmm_plan_s>();
    P->tree = tree;
    P->p = desc->p;
    P->K = (desc->p + 1) * (desc->p + 1);
    P->precision = desc->precision;
    P->max_ncharge = desc->max_ncharge;
    P->m2l_kernel = desc->m2l_kernel;
    HIP_TRY(hipGetDevice(&P->device));
    HIP_TRY(hipStreamCreateWithFlags(&P->stream, hipStreamNonBlocking));
    for (auto& e : P->ev) HIP_TRY(hipEventCreate(&e));
    const sfmm::Tree& t = tree->t;
    const int p = P->p, K = P->K;
    P->tables = *tables_for(p);
    const sfmm::Tables& tb = P->tables;

    // ---- nodes, numbered level by level (gnode)
    const int D = t.depth;
    P->level_first.assign(D + 2, 0);
    for (int l = 0; l <= D; ++l) P->level_first[l + 1] = P->level_first[l] + t.lv[l].n_node();
    const int64_t nnode = P->level_first[D + 1];
    std::vector<int8_t> octant(nnode), nchild(nnode);
    std::vector<int32_t> parent(nnode), child0(nnode);
    for (int l = 0; l <= D; ++l) {
        const auto& lv = t.lv[l];
        for (int32_t a = 0; a < lv.n_node(); ++a) {
            const int64_t g = P->level_first[l] + a;
            octant[g] = static_cast<int8_t>(lv.code[a] & 7);
            nchild[g] = lv.nchild[a];
            parent[g] = l > 0 ? static_cast<int32_t>(P->level_first[l - 1] + lv.parent[a]) : -1;
            child0[g] = lv.child0[a] >= 0 ? static_cast<int32_t>(P->level_first[l + 1] + lv.child0[a]) : -1;
        }
    }
    // ---- leaves, and the leaves with far field (levels >= 2) with their centres
    const int64_t n = t.n;
    const int nthreads = host_threads();
    std::vector<int32_t> fl_gnode, fl_level;
    std::vector<int64_t> fl_begin, fl_end, lf_begin, lf_end;
    for (size_t g = 0; g < t.leaf_node.size(); ++g) {
        const int l = t.leaf_level[g];
        const auto& lv = t.lv[l];
        const int32_t a = t.leaf_node[g];
        lf_begin.push_back(lv.pbegin[a]);
        lf_end.push_back(lv.pend[a]);
        if (l < 2) continue;
        fl_gnode.push_back(static_cast<int32_t>(P->level_first[l] + a));
        fl_level.push_back(l);
        fl_begin.push_back(lv.pbegin[a]);
        fl_end.push_back(lv.pend[a]);
    }
    const int64_t nfl = static_cast<int64_t>(fl_gnode.size());
    std::vector<double> fl_c(3 * static_cast<size_t>(nfl)), fl_iw(nfl);
    HostArray<int32_t> slot_fleaf;  // far leaf of each slot, -1 for none
    slot_fleaf.resize(n);
    parallel_ranges(n, nthreads, [&](int64_t s0, int64_t s1, int) {
        std::fill(slot_fleaf.data.get() + s0, slot_fleaf.data.get() + s1, -1);
    });
    parallel_ranges(nfl, nthreads, [&](int64_t f0, int64_t f1, int) {
        for (int64_t f = f0; f < f1; ++f) {
            const int l = fl_level[f];
            const int32_t a = static_cast<int32_t>(fl_gnode[f] - P->level_first[l]);
            t.box.centre(l, t.lv[l].code[a], &fl_c[3 * f]);
            fl_iw[f] = std::ldexp(1.0, l) / t.box.width;
            for (int64_t s = fl_begin[f]; s < fl_end[f]; ++s) slot_fleaf[s] = static_cast<int32_t>(f);
        }
    });
    // ---- particles in sorted order, in one pass: positions, the inverse
    // permutation, and coordinates relative to their far leaf's centre in leaf widths
    HostArray<double> px, py, pz, ux, uy, uz;
    HostArray<int32_t> iperm;
    for (auto* a : {&px, &py, &pz, &ux, &uy, &uz}) a->resize(n);
    iperm.resize(n);
    parallel_ranges(n, nthreads, [&](int64_t s0, int64_t s1, int) {
        for (int64_t s = s0; s < s1; ++s) {
            const double* x = tree->pos.data() + 3 * static_cast<int64_t>(t.perm[s]);
            px[s] = x[0];
            py[s] = x[1];
            pz[s] = x[2];
            iperm[t.perm[s]] = static_cast<int32_t>(s);
            const int32_t f = slot_fleaf[s];
            if (f < 0) {
                ux[s] = uy[s] = uz[s] = 0.0;
                continue;
            }
            double r[3];
            t.box.rotate(x, r);
            const double* c = &fl_c[3 * static_cast<size_t>(f)];
            const double iw = fl_iw[f];
            ux[s] = (r[0] - c[0]) * iw;
            uy[s] = (r[1] - c[1]) * iw;
            uz[s] = (r[2] - c[2]) * iw;
        }
    });
    // ---- self lists (fmm.hpp:398-408): per target, listed sources other than
    // itself, flagged when in its near field (is_neighbor, partition.hpp:484-488)
    std::vector<int64_t> sc_tgt, sc_row{0}, sc_src;
    std::vector<int8_t> sc_near;
    if (desc->self_row) {
        const int64_t* row = desc->self_row;
        const int64_t* val = desc->self_val;
        if (row[0] != 0) return fail(SFMM_ERR_ARG, "self_row[0] must be 0");
        std::vector<int32_t> slot_leaf(n);
        for (size_t g = 0; g < lf_begin.size(); ++g)
            for (int64_t s = lf_begin[g]; s < lf_end[g]; ++s) slot_leaf[s] = static_cast<int32_t>(g);
        P->slist_row.assign(1, 0);
        std::vector<int64_t> near_seen, far_j;
        for (int64_t i = 0; i < n; ++i) {
            if (row[i + 1] < row[i]) return fail(SFMM_ERR_ARG, "self_row must be non-decreasing");
            const int64_t ti = iperm[i];
            const int32_t g = slot_leaf[ti];
            bool any = false;
            near_seen.clear();
            far_j.clear();
            for (int64_t k = row[i]; k < row[i + 1]; ++k) {
                const int64_t j = val[k];
                if (j < 0 || j >= n) return fail(SFMM_ERR_ARG, "self_val entry outside [0, n)");
                if (j == i) continue;  // never part of the near sum
                const int64_t sj = iperm[j];
                bool near = false;
                for (int64_t r = t.near_row[g]; r < t.near_row[g + 1] && !near; ++r)
                    near = t.near_begin[r] <= sj && sj < t.near_end[r];
                // a near source listed twice is still excluded from the near sum once
                // (is_self is a predicate, fmm.hpp:786); a far one listed twice is
                // subtracted twice (_slist keeps every listed entry, fmm.hpp:398-408)
                if (near) {
                    if (std::find(near_seen.begin(), near_seen.end(), sj) != near_seen.end()) continue;
                    near_seen.push_back(sj);
                } else {
                    far_j.push_back(j);
                }
                sc_src.push_back(sj);
                sc_near.push_back(near ? 1 : 0);
                any = true;
            }
            if (any) {
                sc_tgt.push_back(ti);
                sc_row.push_back(static_cast<int64_t>(sc_src.size()));
            }
            std::sort(far_j.begin(), far_j.end());  // _slist rows are sorted (csr.hpp:38-56)
            P->slist_val.insert(P->slist_val.end(), far_j.begin(), far_j.end());
            P->slist_row.push_back(static_cast<int64_t>(P->slist_val.size()));
        }
    }
    // ---- M2L work of the row kernel: all nodes at levels >= 2 (the MFMA kernel
    //      builds its own offset-grouped blocks from the tree's lists)
    std::vector<int32_t> tgt, il_src;
    std::vector<int16_t> il_off;
    std::vector<int64_t> il_row{0};
    for (int l = 2; l <= D && P->m2l_kernel == 1; ++l) {
        const auto& lv = t.lv[l];
        for (int32_t a = 0; a < lv.n_node(); ++a) {
            tgt.push_back(static_cast<int32_t>(P->level_first[l] + a));
            for (int64_t k = lv.il_row[a]; k < lv.il_row[a + 1]; ++k) {
                il_src.push_back(static_cast<int32_t>(P->level_first[l] + lv.il_src[k]));
                il_off.push_back(lv.il_off[k]);
            }
            il_row.push_back(static_cast<int64_t>(il_src.size()));
        }
    }
    // ---- near-field work items: a leaf's concatenated near sources cut into runs
    //      of at most ~2^18 pairs; split leaves go through deterministic partial sums.
    //      Per leaf (host threads): its source count and run length; then the item
    //      and partial-row offsets by prefix sums; then the items in place.
    const int64_t nleaf = static_cast<int64_t>(t.leaf_node.size());
    std::vector<int64_t> near_pref(t.near_begin.size());
    std::vector<int64_t> l_total(nleaf), l_per(nleaf), l_nitem(nleaf), l_item0(nleaf + 1, 0), l_row0(nleaf + 1, 0),
        l_split0(nleaf + 1, 0);
    parallel_ranges(nleaf, nthreads, [&](int64_t g0, int64_t g1, int) {
        for (int64_t g = g0; g < g1; ++g) {
            int64_t total = 0;
            for (int64_t r = t.near_row[g]; r < t.near_row[g + 1]; ++r) {
                near_pref[r] = total;
                total += t.near_end[r] - t.near_begin[r];
            }
            const int64_t nt = lf_end[g] - lf_begin[g];
            const int64_t per = std::max<int64_t>(64, ((int64_t(1) << 18) / std::max<int64_t>(nt, 1)) / 64 * 64);
            l_per[g] = per;
            l_nitem[g] = std::max<int64_t>(1, (total + per - 1) / per);
            l_total[g] = total;
        }
    });
    for (int64_t g = 0; g < nleaf; ++g) {
        const int64_t nt = lf_end[g] - lf_begin[g];
        const bool split = l_nitem[g] > 1;
        l_item0[g + 1] = l_item0[g] + l_nitem[g];
        l_row0[g + 1] = l_row0[g] + (split ? l_nitem[g] * nt : 0);
        l_split0[g + 1] = l_split0[g] + (split ? nt : 0);
    }
    const int64_t nitems = l_item0[nleaf], nsplit = l_split0[nleaf], prow = l_row0[nleaf];
    std::vector<int64_t> it_leaf(nitems), it_begin(nitems), it_count(nitems), it_part(nitems);
    std::vector<int64_t> sp_slot(nsplit), sp_first(nsplit), sp_nitem(nsplit), sp_stride(nsplit);
    parallel_ranges(nleaf, nthreads, [&](int64_t g0, int64_t g1, int) {
        for (int64_t g = g0; g < g1; ++g) {
            const int64_t nt = lf_end[g] - lf_begin[g], per = l_per[g], nitem = l_nitem[g], total = l_total[g];
            for (int64_t i = 0; i < nitem; ++i) {
                const int64_t k = l_item0[g] + i;
                it_leaf[k] = g;
                it_begin[k] = i * per;
                it_count[k] = std::min(per, total - i * per);
                it_part[k] = nitem == 1 ? -1 : l_row0[g] + i * nt;
            }
            if (nitem > 1)
                for (int64_t j = 0; j < nt; ++j) {
                    const int64_t k = l_split0[g] + j;
                    sp_slot[k] = lf_begin[g] + j;
                    sp_first[k] = l_row0[g] + j;
                    sp_nitem[k] = nitem;
                    sp_stride[k] = nt;
                }
        }
    });
    P->p2p.n_item = static_cast<int64_t>(it_leaf.size());
    P->p2p.n_split_targets = static_cast<int64_t>(sp_slot.size());
    P->p2p.n_partial_rows = prow;
    std::vector<double> inv_width(D + 1);
    for (int l = 0; l <= D; ++l) inv_width[l] = std::ldexp(1.0, l) / t.box.width;
    // ---- tables
    std::vector<double> m2mT(tb.m2m.size()), l2lT(tb.l2l.size());
    for (int o = 0; o < 8; ++o)
        for (int r = 0; r < K; ++r)
            for (int c = 0; c < K; ++c) {
                const size_t src = (static_cast<size_t>(o) * K + r) * K + c;
                const size_t dst = (static_cast<size_t>(o) * K + c) * K + r;
                m2mT[dst] = tb.m2m[src];
                l2lT[dst] = tb.l2l[src];
            }
    std::vector<int32_t> om_rows(K);
    for (int r = 0; r < K; ++r) {
        int nn = 0;
        while ((nn + 1) * (nn + 1) <= r) ++nn;
        const int qq = r - nn * nn;
        const int m = (qq + 1) / 2;
        const int part = (qq == 0 || (qq & 1)) ? 0 : 1;
        om_rows[tb.om[r]] = m | (nn << 8) | (part << 16);
    }
    std::vector<double> rec(3 * 21 * 21, 0.0);
    for (int nn = 0; nn <= 20; ++nn)
        for (int m = 0; m < nn; ++m) {
            const double den = std::sqrt(static_cast<double>(nn * nn - m * m));
            rec[nn * 21 + m] = (2 * nn - 1) / den;
            rec[21 * 21 + nn * 21 + m] = std::sqrt(static_cast<double>((nn - 1) * (nn - 1) - m * m)) / den;
        }
    for (int m = 1; m <= 20; ++m) rec[2 * 21 * 21 + m] = std::sqrt((2.0 * m - 1) / (2.0 * m));

    sfmm::DevPlan& d = P->d;
    d.p = p;
    d.K = K;
    d.depth = D;
    d.n = n;
    d.n_far_leaf = static_cast<int64_t>(fl_gnode.size());
    d.n_leaf = static_cast<int64_t>(lf_begin.size());
    d.n_target = static_cast<int64_t>(tgt.size());
    d.rot_size = tb.rot_size;
    d.coax_size = tb.coax_size;
    int rc = SFMM_OK;
#define UP(vec, field) \
    if ((rc = P->upload(vec, &d.field)) != SFMM_OK) return rc
    UP(px, px); UP(py, py); UP(pz, pz); UP(ux, ux); UP(uy, uy); UP(uz, uz);
    UP(t.perm, perm); UP(iperm, iperm);
    P->self.n_row = static_cast<int64_t>(sc_tgt.size());
    if ((rc = P->upload(sc_tgt, &P->self.tgt)) || (rc = P->upload(sc_row, &P->self.row)) ||
        (rc = P->upload(sc_src, &P->self.src)) || (rc = P->upload(sc_near, &P->self.near)))
        return rc;
    UP(fl_gnode, fl_gnode); UP(fl_level, fl_level); UP(fl_begin, fl_begin); UP(fl_end, fl_end);
    UP(slot_fleaf, slot_fleaf);
    UP(lf_begin, lf_begin); UP(lf_end, lf_end);
    UP(t.near_row, near_row); UP(t.near_begin, near_begin); UP(t.near_end, near_end); UP(near_pref, near_pref);
    UP(octant, octant); UP(parent, parent); UP(child0, child0); UP(nchild, nchild);
    UP(tgt, tgt_gnode); UP(il_row, il_row); UP(il_src, il_src); UP(il_off, il_off);
    UP(inv_width, inv_width);
    UP(m2mT, m2mT); UP(l2lT, l2lT); UP(tb.rot, rot); UP(tb.coax, coax); UP(tb.off_cs, off_cs);
    UP(tb.off_class, off_class); UP(tb.off_d2, off_d2); UP(tb.om, om); UP(om_rows, om_rows);
    UP(tb.rot_off, rot_off); UP(tb.coax_off, coax_off); UP(rec, rec);
#undef UP
    if ((rc = P->upload(it_leaf, &P->p2p.leaf)) || (rc = P->upload(it_begin, &P->p2p.src_begin)) ||
        (rc = P->upload(it_count, &P->p2p.src_count)) || (rc = P->upload(it_part, &P->p2p.partial)) ||
        (rc = P->upload(sp_slot, &P->p2p.split_slot)) || (rc = P->upload(sp_first, &P->p2p.split_first)) ||
        (rc = P->upload(sp_nitem, &P->p2p.split_nitem)) || (rc = P->upload(sp_stride, &P->p2p.split_stride)))
        return rc;
    P->dense = P->precision == SFMM_MIXED && P->m2l_kernel == 0 && p <= sfmm::kDenseMaxP;
    if (P->dense) {
        const std::shared_ptr<const DenseHost> dh = dense_for(p, tb);
        const std::vector<uint16_t>& ops = dh->ops;
        const std::vector<float>& az = dh->az;
        P->dense_op_exp = dh->op_exp;
        P->dense_grp = dh->grp;
        P->dense_kd = sfmm::dense_order(p).kd;
        const uint16_t* dops = nullptr;
        const float* dmax = nullptr;
        if ((rc = P->upload(ops, &dops)) || (rc = P->upload(az, &P->dense_az)) ||
            (rc = P->upload(std::vector<float>(P->dense_kd, 0.0f), &P->dense_zero)) ||
            (rc = P->upload(std::vector<float>(static_cast<size_t>(4) * nnode, 0.0f), &dmax)))
            return rc;
        P->dense_ops = dops;
        P->dense_nodemax = const_cast<float*>(dmax);
    }
    if (!P->dense) {
        std::vector<double> rotA, coaxA;
        sfmm::build_m2l_fragments(tb, P->precision == SFMM_MIXED, rotA, coaxA);
        if (P->precision == SFMM_MIXED) {
            std::vector<float> r32(rotA.begin(), rotA.end()), c32(coaxA.begin(), coaxA.end());
            if ((rc = P->upload(r32, &P->rotA32)) || (rc = P->upload(c32, &P->coaxA32))) return rc;
        } else {
            if ((rc = P->upload(rotA, &P->rotA64)) || (rc = P->upload(coaxA, &P->coaxA64))) return rc;
        }
    }
    if ((rc = P->upload(std::vector<double>(4 * P->K, 0.0), &P->zero_cols))) return rc;
    {
        // matrix-core M2M / L2L operators from the row-major per-octant tables
        const int K = P->K;
        for (int up = 0; up < 2; ++up) {
            const std::vector<double>& R = up ? tb.l2l : tb.m2m;
            const int nf = sfmm::translate_nfrag(P->p, up);
            std::vector<double> frag(static_cast<size_t>(8) * nf * 64, 0.0);
            for (int o = 0; o < 8; ++o)
                for (int r = 0; r < K; ++r)
                    for (int c = 0; c < K; ++c) {
                        int lane = 0;
                        const int f = sfmm::translate_frag_index(P->p, up, r, c, &lane);
                        const double v = R[(static_cast<size_t>(o) * K + r) * K + c];
                        if (f >= 0) frag[(static_cast<size_t>(o) * nf + f) * 64 + lane] = v;
                        else if (v != 0.0) return fail(SFMM_ERR_STATE, "translation operator outside its band");
                    }
            if ((rc = P->upload(frag, up ? &P->l2l_frag : &P->m2m_frag))) return rc;
        }
    }
    *out = P.release();
    return SFMM_OK;
}

int sfmm_plan_destroy(sfmm_plan plan)
{
    if (plan) {
        (void)hipSetDevice(plan->device);
        (void)hipStreamSynchronize(plan->stream);
    }
    delete plan;
    return SFMM_OK;
}

int sfmm_plan_slist(sfmm_plan plan, int64_t* row, int64_t* val, int64_t* nnz)
{
    if (!plan || !nnz) return fail(SFMM_ERR_ARG, "null argument");
    const int64_t n = plan->tree->t.n;
    const bool given = !plan->slist_row.empty();
    *nnz = given ? plan->slist_row.back() : 0;
    if (row) {
        if (given) std::memcpy(row, plan->slist_row.data(), sizeof(int64_t) * (n + 1));
        else std::fill(row, row + n + 1, int64_t(0));
    }
    if (val && given && !plan->slist_val.empty())
        std::memcpy(val, plan->slist_val.data(), sizeof(int64_t) * plan->slist_val.size());
    return SFMM_OK;
}

int sfmm_plan_stream(sfmm_plan plan, void** stream)
{
    if (!plan || !stream) return fail(SFMM_ERR_ARG, "null argument");
    *stream = plan->stream;
    return SFMM_OK;
}

int sfmm_synchronize(sfmm_plan plan)
{
    if (!plan) return fail(SFMM_ERR_ARG, "null argument");
    HIP_TRY(hipSetDevice(plan->device));
    HIP_TRY(hipStreamSynchronize(plan->stream));
    return SFMM_OK;
}

int sfmm_evaluate_device(sfmm_plan plan, int kind, int ncharge, const double* dQ, double* dU, int which)
{
    if (!plan || !dQ || !dU) return fail(SFMM_ERR_ARG, "null argument");
    return run(plan, kind, ncharge, dQ, dU, which);
}

int sfmm_plan_tune(sfmm_plan plan, int kind, int ncharge, const double* dQ)
{
    if (!plan || !dQ) return fail(SFMM_ERR_ARG, "null argument");
    sfmm::Mode mode;
    int qdim, nc;
    HIP_TRY(hipSetDevice(plan->device));
    if (int rc = prepare(plan, kind, ncharge, SFMM_BOTH, mode, qdim, nc)) return rc;
    double* scratch = nullptr;
    if (dev_malloc(&scratch, sizeof(double) * nc * plan->tree->t.n) != hipSuccess)
        return fail(SFMM_ERR_ALLOC, "hipMalloc failed (tuning outputs)");
    const sfmm_batch_item it{plan, kind, ncharge, SFMM_BOTH, dQ, scratch};
    int rc = tune_tiles(plan, it, plan->stream, true, true, nullptr, nullptr);
    const hipError_t e = hipStreamSynchronize(plan->stream);
    (void)hipFree(scratch);
    if (rc != SFMM_OK) return rc;
    if (e != hipSuccess) return fail(SFMM_ERR_HIP, std::string("tuning: ") + hipGetErrorString(e));
    return SFMM_OK;
}

int sfmm_plan_set_tiles(sfmm_plan plan, int p2p_targets_per_lane, int m2l_waves)
{
    if (!plan) return fail(SFMM_ERR_ARG, "null argument");
    if (p2p_targets_per_lane != 1 && p2p_targets_per_lane != 2 && p2p_targets_per_lane != 4)
        return fail(SFMM_ERR_ARG, "near-field targets per lane must be 1, 2 or 4");
    if (m2l_waves != 0 && m2l_waves != 4 && m2l_waves != 8)
        return fail(SFMM_ERR_ARG, "M2L wavefronts per workgroup must be 4 or 8 (0: keep)");
    HIP_TRY(hipSetDevice(plan->device));
    HIP_TRY(hipStreamSynchronize(plan->stream));  // work in flight keeps its shapes
    plan->p2p_tpl = p2p_targets_per_lane;
    if (m2l_waves != 0 && plan->dense) set_dense_waves(plan, m2l_waves);
    return SFMM_OK;
}

int sfmm_plan_tiles(sfmm_plan plan, int* p2p_targets_per_lane, int* m2l_waves)
{
    if (!plan || !p2p_targets_per_lane || !m2l_waves) return fail(SFMM_ERR_ARG, "null argument");
    *p2p_targets_per_lane = plan->p2p_tpl;
    *m2l_waves = plan->dense ? plan->dense_waves : 0;
    return SFMM_OK;
}

int sfmm_evaluate(sfmm_plan plan, int kind, int ncharge, const double* Q, double* U, int which)
{
    if (!plan || !Q || !U) return fail(SFMM_ERR_ARG, "null argument");
    sfmm::Mode mode;
    int qdim, nc;
    if (int rc = mode_of(kind, ncharge, mode, qdim, nc)) return rc;
    HIP_TRY(hipSetDevice(plan->device));
    const int64_t n = plan->tree->t.n;
    if (qdim * n > plan->stage_q) {
        (void)hipFree(plan->hQ);
        plan->hQ = nullptr;
        if (dev_malloc(&plan->hQ, sizeof(double) * qdim * n) != hipSuccess) return fail(SFMM_ERR_ALLOC, "hipMalloc failed (staging)");
        plan->stage_q = qdim * n;
    }
    if (nc * n > plan->stage_u) {
        (void)hipFree(plan->hU);
        plan->hU = nullptr;
        if (dev_malloc(&plan->hU, sizeof(double) * nc * n) != hipSuccess) return fail(SFMM_ERR_ALLOC, "hipMalloc failed (staging)");
        plan->stage_u = nc * n;
    }
    HIP_TRY(hipMemcpyAsync(plan->hQ, Q, sizeof(double) * qdim * n, hipMemcpyHostToDevice, plan->stream));
    sfmm::launch_fill(plan->hU, nc * n, 0.0, plan->stream);
    if (int rc = run(plan, kind, ncharge, plan->hQ, plan->hU, which)) return rc;
    std::vector<double> tmp(static_cast<size_t>(nc) * n);
    HIP_TRY(hipMemcpyAsync(tmp.data(), plan->hU, sizeof(double) * nc * n, hipMemcpyDeviceToHost, plan->stream));
    HIP_TRY(hipStreamSynchronize(plan->stream));
    for (size_t i = 0; i < tmp.size(); ++i) U[i] += tmp[i];
    return SFMM_OK;
}

int sfmm_last_stage_ms(sfmm_plan plan, float ms[9])
{
    if (!plan || !ms) return fail(SFMM_ERR_ARG, "null argument");
    if (!plan->timed) return fail(SFMM_ERR_STATE, "no evaluation has run yet");
    HIP_TRY(hipSetDevice(plan->device));
    const hipEvent_t* ev = plan->last_ev;
    HIP_TRY(hipEventSynchronize(ev[8]));
    for (int k = 0; k < 8; ++k) HIP_TRY(hipEventElapsedTime(&ms[k], ev[k], ev[k + 1]));
    HIP_TRY(hipEventElapsedTime(&ms[8], ev[0], ev[8]));
    return SFMM_OK;
}

int sfmm_record_stages(sfmm_plan plan, int count)
{
    if (!plan) return fail(SFMM_ERR_ARG, "null argument");
    HIP_TRY(hipSetDevice(plan->device));
    HIP_TRY(hipStreamSynchronize(plan->stream));
    const size_t need = 9 * static_cast<size_t>(std::max(count, 0));
    if (need > plan->rec_ev.size()) plan->last_ev = plan->ev;  // the vector may reallocate
    plan->rec_ev.reserve(need);
    while (plan->rec_ev.size() < need) {
        hipEvent_t e;
        HIP_TRY(hipEventCreate(&e));
        plan->rec_ev.push_back(e);
    }
    plan->rec_count = std::max(count, 0);
    plan->rec_used = 0;
    return SFMM_OK;
}

int sfmm_recorded_stage_ms(sfmm_plan plan, float* ms, int capacity, int* recorded)
{
    if (!plan || !recorded || (capacity > 0 && !ms)) return fail(SFMM_ERR_ARG, "null argument");
    HIP_TRY(hipSetDevice(plan->device));
    const int nrec = std::min(plan->rec_used, capacity);
    for (int r = 0; r < nrec; ++r) {
        hipEvent_t* ev = &plan->rec_ev[9 * static_cast<size_t>(r)];
        HIP_TRY(hipEventSynchronize(ev[8]));
        for (int k = 0; k < 8; ++k) HIP_TRY(hipEventElapsedTime(&ms[9 * r + k], ev[k], ev[k + 1]));
        HIP_TRY(hipEventElapsedTime(&ms[9 * r + 8], ev[0], ev[8]));
    }
    *recorded = plan->rec_used;
    return SFMM_OK;
}

int sfmm_direct(sfmm_plan plan, int kind, int ncharge, const double* Q, int64_t ntarget, const int64_t* targets,
                double* out)
{
    if (!plan || !Q || !targets || !out) return fail(SFMM_ERR_ARG, "null argument");
    sfmm::Mode mode;
    int qdim, nc;
    if (int rc = mode_of(kind, ncharge, mode, qdim, nc)) return rc;
    const int64_t n = plan->tree->t.n;
    for (int64_t k = 0; k < ntarget; ++k)
        if (targets[k] < 0 || targets[k] >= n) return fail(SFMM_ERR_ARG, "target index out of range");
    if (ntarget <= 0) return SFMM_OK;
    HIP_TRY(hipSetDevice(plan->device));
    if (int rc = ensure_buffers(plan, nc, qdim)) return rc;
    const int nsplit = sfmm::direct_nsplit(ntarget, n);
    double *dQ = nullptr, *dout = nullptr, *part = nullptr;
    int64_t* dt = nullptr;
    auto cleanup = [&]() {
        (void)hipFree(dQ);
        (void)hipFree(dout);
        (void)hipFree(part);
        (void)hipFree(dt);
    };
    if (dev_malloc(&dQ, sizeof(double) * qdim * n) != hipSuccess ||
        dev_malloc(&dout, sizeof(double) * nc * ntarget) != hipSuccess ||
        dev_malloc(&part, sizeof(double) * sfmm::direct_scratch_size(ntarget, nc, nsplit)) != hipSuccess ||
        dev_malloc(&dt, sizeof(int64_t) * ntarget) != hipSuccess) {
        cleanup();
        return fail(SFMM_ERR_ALLOC, "hipMalloc failed (direct sum)");
    }
    hipStream_t s = plan->stream;
    hipError_t e = hipMemcpyAsync(dQ, Q, sizeof(double) * qdim * n, hipMemcpyHostToDevice, s);
    if (e == hipSuccess) e = hipMemcpyAsync(dt, targets, sizeof(int64_t) * ntarget, hipMemcpyHostToDevice, s);
    if (e == hipSuccess) {
        sfmm::launch_gather_charges(plan->d, qdim, dQ, plan->qbuf, s);
        sfmm::launch_direct(plan->d, mode, nc, plan->qbuf, ntarget, dt, dout, part, nsplit, s);
        e = hipGetLastError();
    }
    if (e == hipSuccess) e = hipMemcpyAsync(out, dout, sizeof(double) * nc * ntarget, hipMemcpyDeviceToHost, s);
    if (e == hipSuccess) e = hipStreamSynchronize(s);
    cleanup();
    if (e != hipSuccess) return fail(SFMM_ERR_HIP, std::string("direct sum: ") + hipGetErrorString(e));
    return SFMM_OK;
}


// ------------------------------------------------------------------ direct sum without allocation
// The partial sums of sfmm_direct_device_noalloc: ntarget * nout * nsplit
// doubles, at most 2^19 * 4 (16 MiB) for ntarget <= 2^19.  A static device array,
// so the call allocates nothing; every device holds its own copy, reached only
// through the address hipGetSymbolAddress gives with that device current.
namespace {

constexpr int64_t kNoAllocScratch = 4 * int64_t(SFMM_DIRECT_NOALLOC_MAX_TARGETS);
__device__ __align__(256) double scratch[kNoAllocScratch];

constexpr int kMaxDevices = 64;
std::mutex g_scratch_mu;
double* g_scratch_ptr[kMaxDevices] = {};

// The device address of `scratch` on device `dev` (current), looked up once.
int noalloc_scratch(int dev, double** out)
{
    if (dev < 0 || dev >= kMaxDevices) return fail(SFMM_ERR_ARG, "device index out of range");
    std::lock_guard<std::mutex> lock(g_scratch_mu);
    if (!g_scratch_ptr[dev]) {
        void* ptr = nullptr;
        const hipError_t e = hipGetSymbolAddress(&ptr, HIP_SYMBOL(scratch));
        if (e != hipSuccess || !ptr)
            return fail(SFMM_ERR_HIP, std::string("hipGetSymbolAddress(scratch): ") + hipGetErrorString(e));
        g_scratch_ptr[dev] = static_cast<double*>(ptr);
    }
    *out = g_scratch_ptr[dev];
    return SFMM_OK;
}

}  // namespace

int sfmm_direct_device_noalloc(sfmm_plan plan, int kind, int ncharge, const double* dQ, int64_t ntarget,
                               const int64_t* dtargets, double* dout)
{
    if (!plan || !dQ) return fail(SFMM_ERR_ARG, "null argument");
    sfmm::Mode mode;
    int qdim, nc;
    if (int rc = mode_of(kind, ncharge, mode, qdim, nc)) return rc;
    if (ntarget < 0 || ntarget > SFMM_DIRECT_NOALLOC_MAX_TARGETS)
        return fail(SFMM_ERR_ARG, "sfmm_direct_device_noalloc: ntarget must be in 0.." +
                                      std::to_string(SFMM_DIRECT_NOALLOC_MAX_TARGETS) + " (static scratch)");
    if (ntarget > 0 && (!dtargets || !dout)) return fail(SFMM_ERR_ARG, "null argument");
    if (ntarget == 0) return SFMM_OK;
    HIP_TRY(hipSetDevice(plan->device));
    double* part = nullptr;
    if (int rc = noalloc_scratch(plan->device, &part)) return rc;
    const int64_t n = plan->tree->t.n;
    const int nsplit = sfmm::direct_nsplit(ntarget, n);
    if (sfmm::direct_scratch_size(ntarget, nc, nsplit) > kNoAllocScratch)
        return fail(SFMM_ERR_ARG, "sfmm_direct_device_noalloc: partial sums exceed the static scratch");
    // every partial sum the reduction reads is written first: nothing to clear
    sfmm::launch_direct(plan->d, mode, nc, dQ, ntarget, dtargets, dout, part, nsplit, plan->stream, plan->d.perm);
    HIP_TRY(hipGetLastError());
    return SFMM_OK;
}

int sfmm_batch_create(const sfmm_batch_item* items, int count, int streams, sfmm_batch* out)
{
    if (!items || !out || count <= 0) return fail(SFMM_ERR_ARG, "null argument or empty batch");
    *out = nullptr;
    if (streams < 0 || streams > kBatchMaxStreams) return fail(SFMM_ERR_ARG, "streams must be in 0..8 (0: tuned)");
    auto B = std::make_unique<sfmm_batch_s>();
    B->items.assign(items, items + count);
    std::vector<const sfmm_plan_s*> seen;
    for (const auto& it : B->items) {
        if (!it.plan || !it.dQ || !it.dU) return fail(SFMM_ERR_ARG, "batch item with a null plan or array");
        if (std::find(seen.begin(), seen.end(), it.plan) != seen.end())
            return fail(SFMM_ERR_ARG, "a plan appears twice in the batch (one plan is not reentrant)");
        seen.push_back(it.plan);
        if (it.plan->device != items[0].plan->device) return fail(SFMM_ERR_ARG, "batch plans on different devices");
    }
    B->device = items[0].plan->device;
    HIP_TRY(hipSetDevice(B->device));
    // every buffer and work list exists before the capture
    std::vector<int64_t> nout(count);
    for (int i = 0; i < count; ++i) {
        const auto& it = B->items[i];
        sfmm::Mode mode;
        int qdim, nc;
        if (int rc = prepare(it.plan, it.kind, it.ncharge, it.which, mode, qdim, nc)) return rc;
        HIP_TRY(hipStreamSynchronize(it.plan->stream));
        nout[i] = static_cast<int64_t>(nc) * it.plan->tree->t.n;
    }
    B->streams.resize(kBatchMaxStreams);
    for (auto& st : B->streams) HIP_TRY(hipStreamCreateWithFlags(&st, hipStreamNonBlocking));
    B->fork.resize(1);
    B->join.resize(kBatchMaxStreams);
    for (auto& e : B->fork) HIP_TRY(hipEventCreateWithFlags(&e, hipEventDisableTiming));
    for (auto& e : B->join) HIP_TRY(hipEventCreateWithFlags(&e, hipEventDisableTiming));
    B->pjoin.resize(count);
    for (auto& e : B->pjoin) HIP_TRY(hipEventCreateWithFlags(&e, hipEventDisableTiming));
    B->pre.resize(count);
    B->post.resize(count);
    for (auto& e : B->pre) HIP_TRY(hipEventCreateWithFlags(&e, hipEventDisableTiming));
    for (auto& e : B->post) HIP_TRY(hipEventCreateWithFlags(&e, hipEventDisableTiming));

    HIP_TRY(hipStreamCreateWithFlags(&B->cap, hipStreamNonBlocking));
    HIP_TRY(hipEventCreateWithFlags(&B->done, hipEventDisableTiming));

    int ns = std::min(streams, count);
    if (streams == 0) {
        // tuning, into scratch outputs: the near-field tile shape of each expansion
        // order, then the stream assignment, each candidate timed as captured graphs
        std::vector<double*> scratch(count, nullptr);
        auto release = [&]() {
            for (double* p : scratch) (void)hipFree(p);
        };
        std::vector<sfmm_batch_item> trial = B->items;
        for (int i = 0; i < count; ++i) {
            if (dev_malloc(&scratch[i], sizeof(double) * nout[i]) != hipSuccess) {
                release();
                return fail(SFMM_ERR_ALLOC, "hipMalloc failed (batch tuning outputs)");
            }
            trial[i].dU = scratch[i];
        }
        hipEvent_t t0, t1;
        HIP_TRY(hipEventCreate(&t0));
        HIP_TRY(hipEventCreate(&t1));
        hipStream_t s0 = B->streams[0];
        int rc = SFMM_OK;
        // best of three timed runs after a warm-up
        auto time_best = [&](const std::function<int()>& issue, float* best_ms) -> int {
            *best_ms = -1.0f;
            if (int r = issue()) return r;
            for (int r = 0; r < 3; ++r) {
                float ms = 0.0f;
                HIP_TRY(hipEventRecord(t0, s0));
                if (int r2 = issue()) return r2;
                HIP_TRY(hipEventRecord(t1, s0));
                HIP_TRY(hipEventSynchronize(t1));
                HIP_TRY(hipEventElapsedTime(&ms, t0, t1));
                if (*best_ms < 0.0f || ms < *best_ms) *best_ms = ms;
            }
            return SFMM_OK;
        };
        // tile shapes (near field, dense M2L) of each expansion order, timed on its
        // largest problem and set on every plan of that order
        std::vector<int> orders;
        for (const auto& it : B->items)
            if (std::find(orders.begin(), orders.end(), it.plan->p) == orders.end()) orders.push_back(it.plan->p);
        for (int p : orders) {
            int big = -1, big_dense = -1;
            for (int i = 0; i < count; ++i) {
                const sfmm_plan_s* Q = B->items[i].plan;
                if (Q->p != p) continue;
                if (big < 0 || Q->tree->t.n > B->items[big].plan->tree->t.n) big = i;
                if (Q->dense && (big_dense < 0 || Q->tree->t.n > B->items[big_dense].plan->tree->t.n)) big_dense = i;
            }
            int tpl = 2, waves = 8;
            if ((rc = tune_tiles(B->items[big].plan, trial[big], s0, true, false, &tpl, nullptr)) != SFMM_OK) break;
            if (big_dense >= 0 &&
                (rc = tune_tiles(B->items[big_dense].plan, trial[big_dense], s0, false, true, nullptr, &waves)) != SFMM_OK)
                break;
            for (int i = 0; i < count && rc == SFMM_OK; ++i) {
                sfmm_plan_s* Q = B->items[i].plan;
                if (Q->p != p) continue;
                Q->p2p_tpl = tpl;
                if (!Q->dense) continue;
                set_dense_waves(Q, waves);
                sfmm::Mode mode;
                int qdim, nc;
                rc = prepare(Q, B->items[i].kind, B->items[i].ncharge, B->items[i].which, mode, qdim, nc);
            }
        }
        // stream assignment: 1, 2, 4, 8 streams, then every item on its plan's own
        // stream; more streams are taken only when at least 2% faster (streams beyond
        // the hardware queues share them)
        float best = -1.0f;
        for (int cand = 1; cand <= 2 * std::min(count, kBatchMaxStreams) && rc == SFMM_OK; cand *= 2) {
            const bool own = cand > std::min(count, kBatchMaxStreams);
            const std::vector<int> where = batch_assign(trial, own ? 1 : cand);
            hipGraph_t g = nullptr;
            hipGraphExec_t x = nullptr;
            if ((rc = batch_capture(B.get(), trial, where, cand, own, &g, &x)) != SFMM_OK) break;
            float cbest = -1.0f;
            rc = time_best([&]() { return batch_replay(B.get(), x, trial, s0); }, &cbest);
            (void)hipStreamSynchronize(s0);
            (void)hipGraphExecDestroy(x);
            (void)hipGraphDestroy(g);
            if (rc != SFMM_OK) break;
            if (best < 0.0f || cbest < 0.98f * best) {
                best = cbest;
                ns = own ? 1 : cand;
                B->own = own;
            }
        }
        (void)hipStreamSynchronize(s0);
        (void)hipEventDestroy(t0);
        (void)hipEventDestroy(t1);
        release();
        if (rc != SFMM_OK) return rc;
    }
    B->nstream = ns;
    B->where = batch_assign(B->items, ns);
    if (int rc = batch_capture(B.get(), B->items, B->where, B->own ? static_cast<int>(B->items.size()) : ns, B->own,
                               &B->graph, &B->exec))
        return rc;
    *out = B.release();
    return SFMM_OK;
}

int sfmm_batch_launch(sfmm_batch batch, void* stream)
{
    if (!batch) return fail(SFMM_ERR_ARG, "null argument");
    HIP_TRY(hipSetDevice(batch->device));
    return batch_replay(batch, batch->exec, batch->items, stream ? static_cast<hipStream_t>(stream) : batch->streams[0]);
}

int sfmm_batch_synchronize(sfmm_batch batch)
{
    if (!batch) return fail(SFMM_ERR_ARG, "null argument");
    HIP_TRY(hipSetDevice(batch->device));
    HIP_TRY(hipStreamSynchronize(batch->streams[0]));
    return SFMM_OK;
}

int sfmm_batch_streams(sfmm_batch batch, int* streams)
{
    if (!batch || !streams) return fail(SFMM_ERR_ARG, "null argument");
    *streams = batch->own ? static_cast<int>(batch->items.size()) : batch->nstream;
    return SFMM_OK;
}

int sfmm_batch_destroy(sfmm_batch batch)
{
    if (!batch) return SFMM_OK;
    (void)hipSetDevice(batch->device);
    (void)hipStreamSynchronize(batch->streams[0]);
    delete batch;
    return SFMM_OK;
}

}  // extern "C"
