// Octree construction on the GPU.  It builds the same tree as build_tree (tree.cpp)
// and copies it into the host Tree, array for array (the reference partitioner's
// semantics: box.hpp:57-63, partition.hpp:46-103,168-366, tree.hpp:122-174).
//
//   codes       one thread per particle; cell_index/spread3 are shared with the host
//               build, so the codes are bit-identical
//   sort        stable LSD radix sort of (code, index) pairs
//   levels      per level one pass over the sorted particles flags the first particle
//               of every child octant of a subdividing node; an inclusive scan numbers
//               the children in code order (the host's emission order)
//   leaves      per level a scan over the nodes without children
//   leaf order  a second stable sort, keyed by the leaf's first slot and fed the
//               particles in original order, leaves each leaf ascending by index
//               (partition.hpp:52)
//   lists       per level: count pass, scan, fill pass over the nodes; a node's
//               neighbours and interaction list come from the children of its
//               parent's neighbours, in the same order as the host build
//   near        per leaf: count pass, scan, then the ranges are written, sorted and
//               merged in place and compacted after a second scan
#include <hip/hip_runtime.h>

#include <rocprim/device/device_radix_sort.hpp>
#include <rocprim/device/device_scan.hpp>

#include <algorithm>
#include <cmath>
#include <string>
#include <vector>

#include "morton.hpp"
#include "tree.hpp"

namespace sfmm {

namespace {

constexpr int kBlock = 256;
constexpr int kMaxLevels = 21;

inline unsigned grid_for(int64_t count) { return static_cast<unsigned>((std::max<int64_t>(count, 1) + kBlock - 1) / kBlock); }

__device__ inline int64_t thread_id() { return static_cast<int64_t>(blockIdx.x) * blockDim.x + threadIdx.x; }

struct BoxArg {
    int X, Y, Z;
    double c, s, width, corner[3];
};

__global__ void codes_kernel(int64_t n, const double* __restrict__ pos, BoxArg b, int L, int64_t* __restrict__ key,
                             int32_t* __restrict__ idx, int* __restrict__ bad)
{
    const int64_t i = thread_id();
    if (i >= n) return;
    const double* v = pos + 3 * i;
    double r[3];
    rotate_xy(b.c, b.s, v[b.X], v[b.Y], r[0], r[1]);
    r[2] = v[b.Z];
    int32_t iv[3];
    bool ok = true;
    for (int d = 0; d < 3; ++d) ok = cell_index(r[d], b.corner[d], b.width, L, iv[d]) && ok;
    if (!ok) *bad = 1;
    key[i] = spread3(static_cast<uint32_t>(iv[0])) | (spread3(static_cast<uint32_t>(iv[1])) << 1) |
             (spread3(static_cast<uint32_t>(iv[2])) << 2);
    idx[i] = static_cast<int32_t>(i);
}

template <class T>
__global__ void fill_kernel(int64_t count, T* __restrict__ out, T value)
{
    const int64_t i = thread_id();
    if (i < count) out[i] = value;
}

__global__ void iota_kernel(int64_t count, int32_t* __restrict__ out)
{
    const int64_t i = thread_id();
    if (i < count) out[i] = static_cast<int32_t>(i);
}

// Flags the first particle of each child octant of a subdividing node; particles of a
// node that stays a leaf record the leaf's first slot.
__global__ void split_kernel(int64_t n, const int64_t* __restrict__ key, const int32_t* __restrict__ node_of,
                             const int64_t* __restrict__ pbegin, const int64_t* __restrict__ pend, int cap,
                             int can_split, int sh, int32_t* __restrict__ flag, int64_t* __restrict__ leafbegin)
{
    const int64_t s = thread_id();
    if (s >= n) return;
    const int32_t a = node_of[s];
    int32_t f = 0;
    if (a >= 0) {
        const int64_t b0 = pbegin[a];
        if (can_split && pend[a] - b0 > cap)
            f = (s == b0 || (key[s] >> sh) != (key[s - 1] >> sh)) ? 1 : 0;
        else
            leafbegin[s] = b0;
    }
    flag[s] = f;
}

__global__ void emit_kernel(int64_t n, const int64_t* __restrict__ key, const int32_t* __restrict__ node_of,
                            const int64_t* __restrict__ pbegin, const int64_t* __restrict__ pend, int cap, int sh,
                            const int32_t* __restrict__ flag, const int32_t* __restrict__ incl,
                            int32_t* __restrict__ node_next, int64_t* __restrict__ ncode, int64_t* __restrict__ npbegin,
                            int32_t* __restrict__ nparent)
{
    const int64_t s = thread_id();
    if (s >= n) return;
    const int32_t a = node_of[s];
    int32_t c = -1;
    if (a >= 0 && pend[a] - pbegin[a] > cap) {
        c = incl[s] - 1;
        if (flag[s]) {
            ncode[c] = key[s] >> sh;
            npbegin[c] = s;
            nparent[c] = a;
        }
    }
    node_next[s] = c;
}

// Child ranges end where the next sibling starts (or at the parent's end); the parent
// learns its first child and its number of children.
__global__ void link_kernel(int32_t nnext, const int32_t* __restrict__ nparent, const int64_t* __restrict__ npbegin,
                            const int64_t* __restrict__ pend, int64_t* __restrict__ npend, int32_t* __restrict__ child0,
                            int8_t* __restrict__ nchild)
{
    const int64_t c = thread_id();
    if (c >= nnext) return;
    const int32_t P = nparent[c];
    const bool first = c == 0 || nparent[c - 1] != P;
    const bool last = c + 1 == nnext || nparent[c + 1] != P;
    npend[c] = last ? pend[P] : npbegin[c + 1];
    if (first) child0[P] = static_cast<int32_t>(c);
    if (last) {
        int64_t f = c;
        while (f > 0 && nparent[f - 1] == P) --f;
        nchild[P] = static_cast<int8_t>(c - f + 1);
    }
}

__global__ void leaf_flag_kernel(int32_t nn, const int8_t* __restrict__ nchild, int32_t* __restrict__ flag)
{
    const int64_t a = thread_id();
    if (a < nn) flag[a] = nchild[a] == 0 ? 1 : 0;
}

__global__ void leaf_map_kernel(int32_t nn, const int8_t* __restrict__ nchild, const int32_t* __restrict__ incl,
                                int32_t* __restrict__ node2leaf, int32_t* __restrict__ leaf2node,
                                int32_t* __restrict__ g_level, int32_t* __restrict__ g_node, int64_t leaf_off, int level)
{
    const int64_t a = thread_id();
    if (a >= nn) return;
    if (nchild[a] == 0) {
        const int32_t f = incl[a] - 1;
        node2leaf[a] = f;
        leaf2node[f] = static_cast<int32_t>(a);
        g_level[leaf_off + f] = level;
        g_node[leaf_off + f] = static_cast<int32_t>(a);
    } else {
        node2leaf[a] = -1;
    }
}

__global__ void leaf_key_kernel(int64_t n, const int32_t* __restrict__ perm, const int64_t* __restrict__ leafbegin,
                                uint32_t* __restrict__ leafkey)
{
    const int64_t s = thread_id();
    if (s < n) leafkey[perm[s]] = static_cast<uint32_t>(leafbegin[s]);
}

// Neighbours (|d| <= 1 per axis) and interaction list (levels >= 2) of each node from
// the children of its parent's neighbours; WRITE = false counts, true fills.
template <bool WRITE>
__global__ void lists_kernel(int32_t nn, const int64_t* __restrict__ code, const int32_t* __restrict__ parent,
                             const int64_t* __restrict__ up_nb_row, const int32_t* __restrict__ up_nb,
                             const int32_t* __restrict__ up_child0, const int8_t* __restrict__ up_nchild, int far,
                             int64_t* __restrict__ cnb, int64_t* __restrict__ cil, const int64_t* __restrict__ nb_row,
                             int32_t* __restrict__ nb, const int64_t* __restrict__ il_row, int32_t* __restrict__ il_src,
                             int16_t* __restrict__ il_off)
{
    const int64_t a = thread_id();
    if (a > nn) return;
    if (a == nn) {
        if (!WRITE) cnb[a] = cil[a] = 0;
        return;
    }
    const int64_t ca = code[a];
    const int32_t ai = code_axis(ca, 0), aj = code_axis(ca, 1), ak = code_axis(ca, 2);
    const int32_t P = parent[a];
    int64_t knb = WRITE ? nb_row[a] : 0, kil = WRITE ? il_row[a] : 0;
    for (int64_t k = up_nb_row[P]; k < up_nb_row[P + 1]; ++k) {
        const int32_t Y = up_nb[k];
        const int32_t z0 = up_child0[Y];
        const int32_t z1 = z0 + up_nchild[Y];
        for (int32_t z = z0; z >= 0 && z < z1; ++z) {
            const int64_t cz = code[z];
            const int di = code_axis(cz, 0) - ai, dj = code_axis(cz, 1) - aj, dk = code_axis(cz, 2) - ak;
            if (di >= -1 && di <= 1 && dj >= -1 && dj <= 1 && dk >= -1 && dk <= 1) {
                if (WRITE) nb[knb] = z;
                ++knb;
            } else if (far) {
                if (WRITE) {
                    il_src[kil] = z;
                    il_off[kil] = static_cast<int16_t>(offset_id(di, dj, dk));
                }
                ++kil;
            }
        }
    }
    if (!WRITE) {
        cnb[a] = knb;
        cil[a] = kil;
    }
}

struct LevelPtr {
    const int64_t* pbegin;
    const int64_t* pend;
    const int32_t* parent;
    const int8_t* nchild;
    const int64_t* nb_row;
    const int32_t* nb;
};

// Near-field particle ranges of each leaf: the subtrees of its same-level neighbours
// and the leaves adjacent to its ancestors (partition.hpp:324-366).  MODE 0 counts the
// raw ranges; MODE 1 writes them at `off`, sorts them by start and merges touching
// ranges in place, leaving the merged count in `cnt`.
template <int MODE>
__global__ void near_kernel(int64_t nleaves, const int32_t* __restrict__ g_level, const int32_t* __restrict__ g_node,
                            const LevelPtr* __restrict__ lvp, int64_t* __restrict__ cnt, const int64_t* __restrict__ off,
                            int64_t* __restrict__ rb, int64_t* __restrict__ re)
{
    const int64_t g = thread_id();
    if (g > nleaves) return;
    if (g == nleaves) {
        cnt[g] = 0;
        return;
    }
    const int l = g_level[g];
    int32_t a = g_node[g];
    const int64_t base = MODE ? off[g] : 0;
    int64_t m = 0;
    {
        const LevelPtr me = lvp[l];
        for (int64_t k = me.nb_row[a]; k < me.nb_row[a + 1]; ++k) {
            if (MODE) {
                const int32_t Y = me.nb[k];
                rb[base + m] = me.pbegin[Y];
                re[base + m] = me.pend[Y];
            }
            ++m;
        }
    }
    for (int la = l - 1; la >= 0; --la) {
        a = lvp[la + 1].parent[a];
        const LevelPtr A = lvp[la];
        for (int64_t k = A.nb_row[a]; k < A.nb_row[a + 1]; ++k) {
            const int32_t Y = A.nb[k];
            if (A.nchild[Y] != 0) continue;
            if (MODE) {
                rb[base + m] = A.pbegin[Y];
                re[base + m] = A.pend[Y];
            }
            ++m;
        }
    }
    if (!MODE) {
        cnt[g] = m;
        return;
    }
    // ranges are disjoint and non-empty, so their starts are distinct
    for (int64_t i = 1; i < m; ++i) {
        const int64_t b = rb[base + i], e = re[base + i];
        int64_t j = i - 1;
        while (j >= 0 && rb[base + j] > b) {
            rb[base + j + 1] = rb[base + j];
            re[base + j + 1] = re[base + j];
            --j;
        }
        rb[base + j + 1] = b;
        re[base + j + 1] = e;
    }
    int64_t w = 0;
    for (int64_t i = 0; i < m; ++i) {
        if (w > 0 && re[base + w - 1] == rb[base + i]) {
            re[base + w - 1] = re[base + i];
        } else {
            rb[base + w] = rb[base + i];
            re[base + w] = re[base + i];
            ++w;
        }
    }
    cnt[g] = w;
}

__global__ void near_compact_kernel(int64_t nleaves, const int64_t* __restrict__ off, const int64_t* __restrict__ row,
                                    const int64_t* __restrict__ rb, const int64_t* __restrict__ re,
                                    int64_t* __restrict__ ob, int64_t* __restrict__ oe)
{
    const int64_t g = thread_id();
    if (g >= nleaves) return;
    const int64_t src = off[g], dst = row[g], m = row[g + 1] - row[g];
    for (int64_t k = 0; k < m; ++k) {
        ob[dst + k] = rb[src + k];
        oe[dst + k] = re[src + k];
    }
}

// Device allocations of one build, released together.
struct Arena {
    std::vector<void*> ptrs;
    ~Arena()
    {
        for (void* p : ptrs) (void)hipFree(p);
    }
    template <class T>
    T* get(int64_t count)
    {
        void* p = nullptr;
        if (hipMalloc(&p, static_cast<size_t>(std::max<int64_t>(count, 1)) * sizeof(T)) != hipSuccess) return nullptr;
        ptrs.push_back(p);
        return static_cast<T*>(p);
    }
};

// Scan/sort temporary storage, grown on demand.
struct Temp {
    void* p = nullptr;
    size_t bytes = 0;
    ~Temp()
    {
        if (p) (void)hipFree(p);
    }
    hipError_t reserve(size_t need)
    {
        if (p && need <= bytes) return hipSuccess;  // never null: a null pointer asks rocPRIM for the size
        if (p) {
            hipError_t e = hipFree(p);
            if (e != hipSuccess) return e;
            p = nullptr;
            bytes = 0;
        }
        need = std::max<size_t>(need, 256);
        hipError_t e = hipMalloc(&p, need);
        if (e == hipSuccess) bytes = need;
        return e;
    }
};

struct StreamGuard {
    hipStream_t s = nullptr;
    ~StreamGuard()
    {
        if (s) (void)hipStreamDestroy(s);
    }
};

struct DevLevel {
    int32_t nn = 0, nleaf = 0;
    int64_t *code = nullptr, *pbegin = nullptr, *pend = nullptr;
    int32_t *parent = nullptr, *child0 = nullptr;
    int8_t* nchild = nullptr;
    int32_t *node2leaf = nullptr, *leaf2node = nullptr;
    int64_t *nb_row = nullptr, *il_row = nullptr;
    int32_t *nb = nullptr, *il_src = nullptr;
    int16_t* il_off = nullptr;
    int64_t n_nb = 0, n_il = 0;
};

template <class T>
hipError_t download(std::vector<T>& v, const T* d, int64_t count, hipStream_t s)
{
    v.resize(static_cast<size_t>(count));
    if (count == 0) return hipSuccess;
    return hipMemcpyAsync(v.data(), d, static_cast<size_t>(count) * sizeof(T), hipMemcpyDeviceToHost, s);
}

}  // namespace

#define TRY(expr)                                                                         \
    do {                                                                                  \
        hipError_t e_ = (expr);                                                           \
        if (e_ != hipSuccess) return std::string(#expr) + ": " + hipGetErrorString(e_); \
    } while (0)
#define ALLOC(ptr, T, count)                                                      \
    do {                                                                          \
        ptr = mem.get<T>(count);                                                  \
        if (!ptr) return "device tree build: hipMalloc failed (out of device memory)"; \
    } while (0)

std::string build_tree_device(Tree& t, int64_t n, const double* pos, const BoxGeom& box, int max_level, int cap)
{
    if (n <= 0) return "particle count must be positive";
    if (n > 0x7fffffffLL) return "particle count exceeds 2^31-1";
    if (max_level < 0 || max_level > 20) return "max_level must be in [0, 20] (20-bit Morton fields)";
    if (cap < 0) return "max_particles_per_node must be non-negative";
    if (!(box.width > 0) || !std::isfinite(box.width)) return "box width must be positive and finite";
    const int L = max_level;

    StreamGuard sg;
    TRY(hipStreamCreateWithFlags(&sg.s, hipStreamNonBlocking));
    const hipStream_t s = sg.s;
    Arena mem;
    Temp tmp;

    auto incl_scan32 = [&](const int32_t* in, int32_t* out, int64_t count) -> hipError_t {
        size_t bytes = 0;
        hipError_t e = rocprim::inclusive_scan(nullptr, bytes, in, out, static_cast<size_t>(count),
                                               rocprim::plus<int32_t>(), s);
        if (e == hipSuccess) e = tmp.reserve(bytes);
        if (e == hipSuccess)
            e = rocprim::inclusive_scan(tmp.p, bytes, in, out, static_cast<size_t>(count), rocprim::plus<int32_t>(), s);
        return e;
    };
    auto excl_scan64 = [&](const int64_t* in, int64_t* out, int64_t count) -> hipError_t {
        size_t bytes = 0;
        hipError_t e = rocprim::exclusive_scan(nullptr, bytes, in, out, int64_t(0), static_cast<size_t>(count),
                                               rocprim::plus<int64_t>(), s);
        if (e == hipSuccess) e = tmp.reserve(bytes);
        if (e == hipSuccess)
            e = rocprim::exclusive_scan(tmp.p, bytes, in, out, int64_t(0), static_cast<size_t>(count),
                                        rocprim::plus<int64_t>(), s);
        return e;
    };
    auto read = [&](auto* host, const auto* dev) -> hipError_t {
        hipError_t e = hipMemcpyAsync(host, dev, sizeof(*host), hipMemcpyDeviceToHost, s);
        if (e == hipSuccess) e = hipStreamSynchronize(s);
        return e;
    };

    // 1. codes
    double* d_pos;
    int64_t *key, *key_s;
    int32_t *idx, *perm;
    int* bad;
    ALLOC(d_pos, double, 3 * n);
    ALLOC(key, int64_t, n);
    ALLOC(key_s, int64_t, n);
    ALLOC(idx, int32_t, n);
    ALLOC(perm, int32_t, n);
    ALLOC(bad, int, 1);
    TRY(hipMemcpyAsync(d_pos, pos, sizeof(double) * 3 * n, hipMemcpyHostToDevice, s));
    fill_kernel<int><<<1, kBlock, 0, s>>>(1, bad, 0);
    BoxArg ba{box.X, box.Y, box.Z, box.cphi, box.sphi, box.width, {box.corner[0], box.corner[1], box.corner[2]}};
    codes_kernel<<<grid_for(n), kBlock, 0, s>>>(n, d_pos, ba, L, key, idx, bad);
    TRY(hipGetLastError());
    int h_bad = 0;
    TRY(read(&h_bad, bad));
    if (h_bad) return "a particle lies outside the box (or is not finite)";

    // 2. one stable sort by code
    if (L > 0) {
        // codes are non-negative below bit 60: sort them as unsigned keys
        const uint64_t* uk = reinterpret_cast<const uint64_t*>(key);
        uint64_t* uks = reinterpret_cast<uint64_t*>(key_s);
        const unsigned bits = static_cast<unsigned>(3 * L);
        size_t bytes = 0;
        TRY(rocprim::radix_sort_pairs(nullptr, bytes, uk, uks, idx, perm, n, 0u, bits, s));
        TRY(tmp.reserve(bytes));
        TRY(rocprim::radix_sort_pairs(tmp.p, bytes, uk, uks, idx, perm, n, 0u, bits, s));
    } else {
        TRY(hipMemcpyAsync(key_s, key, sizeof(int64_t) * n, hipMemcpyDeviceToDevice, s));
        TRY(hipMemcpyAsync(perm, idx, sizeof(int32_t) * n, hipMemcpyDeviceToDevice, s));
    }

    // 3. levels top-down
    int32_t *node_of, *node_next, *flag, *incl;
    int64_t* leafbegin;
    ALLOC(node_of, int32_t, n);
    ALLOC(node_next, int32_t, n);
    ALLOC(flag, int32_t, n);
    ALLOC(incl, int32_t, n);
    ALLOC(leafbegin, int64_t, n);
    fill_kernel<int32_t><<<grid_for(n), kBlock, 0, s>>>(n, node_of, 0);
    std::vector<DevLevel> lv(1);
    {
        DevLevel& r = lv[0];
        r.nn = 1;
        ALLOC(r.code, int64_t, 1);
        ALLOC(r.pbegin, int64_t, 1);
        ALLOC(r.pend, int64_t, 1);
        ALLOC(r.parent, int32_t, 1);
        fill_kernel<int64_t><<<1, kBlock, 0, s>>>(1, r.code, 0);
        fill_kernel<int64_t><<<1, kBlock, 0, s>>>(1, r.pbegin, 0);
        fill_kernel<int64_t><<<1, kBlock, 0, s>>>(1, r.pend, n);
        fill_kernel<int32_t><<<1, kBlock, 0, s>>>(1, r.parent, -1);
    }
    for (int l = 0;; ++l) {
        DevLevel& cur = lv[l];
        ALLOC(cur.child0, int32_t, cur.nn);
        ALLOC(cur.nchild, int8_t, cur.nn);
        fill_kernel<int32_t><<<grid_for(cur.nn), kBlock, 0, s>>>(cur.nn, cur.child0, -1);
        fill_kernel<int8_t><<<grid_for(cur.nn), kBlock, 0, s>>>(cur.nn, cur.nchild, 0);
        const int can = l < L ? 1 : 0;
        const int sh = can ? 3 * (L - l - 1) : 0;
        split_kernel<<<grid_for(n), kBlock, 0, s>>>(n, key_s, node_of, cur.pbegin, cur.pend, cap, can, sh, flag,
                                                    leafbegin);
        TRY(hipGetLastError());
        if (!can) break;
        TRY(incl_scan32(flag, incl, n));
        int32_t nnext = 0;
        TRY(read(&nnext, incl + (n - 1)));
        if (nnext == 0) break;
        if (static_cast<int>(lv.size()) >= kMaxLevels) return "device tree build: more than 21 levels";
        DevLevel nx;
        nx.nn = nnext;
        ALLOC(nx.code, int64_t, nnext);
        ALLOC(nx.pbegin, int64_t, nnext);
        ALLOC(nx.pend, int64_t, nnext);
        ALLOC(nx.parent, int32_t, nnext);
        emit_kernel<<<grid_for(n), kBlock, 0, s>>>(n, key_s, node_of, cur.pbegin, cur.pend, cap, sh, flag, incl,
                                                   node_next, nx.code, nx.pbegin, nx.parent);
        link_kernel<<<grid_for(nnext), kBlock, 0, s>>>(nnext, nx.parent, nx.pbegin, cur.pend, nx.pend, cur.child0,
                                                       cur.nchild);
        TRY(hipGetLastError());
        std::swap(node_of, node_next);
        lv.push_back(nx);
    }
    const int depth = static_cast<int>(lv.size()) - 1;

    // 4. leaves per level, then the global (level, leaf) order
    int64_t total_nodes = 0;
    for (const DevLevel& d : lv) total_nodes += d.nn;
    int32_t *g_level, *g_node;
    ALLOC(g_level, int32_t, total_nodes);
    ALLOC(g_node, int32_t, total_nodes);
    std::vector<int64_t> leaf_row(depth + 2, 0);
    for (int l = 0; l <= depth; ++l) {
        DevLevel& d = lv[l];
        leaf_flag_kernel<<<grid_for(d.nn), kBlock, 0, s>>>(d.nn, d.nchild, flag);
        TRY(incl_scan32(flag, incl, d.nn));
        TRY(read(&d.nleaf, incl + (d.nn - 1)));
        ALLOC(d.node2leaf, int32_t, d.nn);
        ALLOC(d.leaf2node, int32_t, d.nleaf);
        leaf_map_kernel<<<grid_for(d.nn), kBlock, 0, s>>>(d.nn, d.nchild, incl, d.node2leaf, d.leaf2node, g_level,
                                                          g_node, leaf_row[l], l);
        TRY(hipGetLastError());
        leaf_row[l + 1] = leaf_row[l] + d.nleaf;
    }
    const int64_t nleaves = leaf_row[depth + 1];

    // particles of a leaf ascending by original index
    uint32_t *leafkey, *leafkey_s;
    int32_t* perm_final;
    ALLOC(leafkey, uint32_t, n);
    ALLOC(leafkey_s, uint32_t, n);
    ALLOC(perm_final, int32_t, n);
    leaf_key_kernel<<<grid_for(n), kBlock, 0, s>>>(n, perm, leafbegin, leafkey);
    iota_kernel<<<grid_for(n), kBlock, 0, s>>>(n, idx);
    TRY(hipGetLastError());
    {
        unsigned bits = 1;
        while (bits < 31 && (int64_t(1) << bits) < n) ++bits;
        size_t bytes = 0;
        TRY(rocprim::radix_sort_pairs(nullptr, bytes, leafkey, leafkey_s, idx, perm_final, n, 0u, bits, s));
        TRY(tmp.reserve(bytes));
        TRY(rocprim::radix_sort_pairs(tmp.p, bytes, leafkey, leafkey_s, idx, perm_final, n, 0u, bits, s));
    }

    // 5. same-level neighbours and interaction lists, top-down
    {
        DevLevel& r = lv[0];
        ALLOC(r.nb_row, int64_t, 2);
        ALLOC(r.nb, int32_t, 1);
        fill_kernel<int64_t><<<1, kBlock, 0, s>>>(1, r.nb_row, 0);
        fill_kernel<int64_t><<<1, kBlock, 0, s>>>(1, r.nb_row + 1, 1);
        fill_kernel<int32_t><<<1, kBlock, 0, s>>>(1, r.nb, 0);
        r.n_nb = 1;
    }
    int64_t *cnb = nullptr, *cil = nullptr;
    {
        int64_t max_nn = 0;
        for (const DevLevel& d : lv) max_nn = std::max<int64_t>(max_nn, d.nn);
        ALLOC(cnb, int64_t, max_nn + 1);
        ALLOC(cil, int64_t, max_nn + 1);
    }
    for (int l = 1; l <= depth; ++l) {
        DevLevel& d = lv[l];
        const DevLevel& up = lv[l - 1];
        const int far = l >= 2 ? 1 : 0;
        lists_kernel<false><<<grid_for(int64_t(d.nn) + 1), kBlock, 0, s>>>(
            d.nn, d.code, d.parent, up.nb_row, up.nb, up.child0, up.nchild, far, cnb, cil, nullptr, nullptr, nullptr,
            nullptr, nullptr);
        TRY(hipGetLastError());
        ALLOC(d.nb_row, int64_t, int64_t(d.nn) + 1);
        ALLOC(d.il_row, int64_t, int64_t(d.nn) + 1);
        TRY(excl_scan64(cnb, d.nb_row, int64_t(d.nn) + 1));
        TRY(excl_scan64(cil, d.il_row, int64_t(d.nn) + 1));
        TRY(read(&d.n_nb, d.nb_row + d.nn));
        TRY(read(&d.n_il, d.il_row + d.nn));
        ALLOC(d.nb, int32_t, d.n_nb);
        ALLOC(d.il_src, int32_t, d.n_il);
        ALLOC(d.il_off, int16_t, d.n_il);
        lists_kernel<true><<<grid_for(int64_t(d.nn) + 1), kBlock, 0, s>>>(d.nn, d.code, d.parent, up.nb_row, up.nb,
                                                                           up.child0, up.nchild, far, nullptr, nullptr,
                                                                           d.nb_row, d.nb, d.il_row, d.il_src, d.il_off);
        TRY(hipGetLastError());
    }

    // 6. near-field ranges of each leaf
    std::vector<LevelPtr> hp(depth + 1);
    for (int l = 0; l <= depth; ++l)
        hp[l] = {lv[l].pbegin, lv[l].pend, lv[l].parent, lv[l].nchild, lv[l].nb_row, lv[l].nb};
    LevelPtr* d_lvp;
    ALLOC(d_lvp, LevelPtr, depth + 1);
    TRY(hipMemcpyAsync(d_lvp, hp.data(), sizeof(LevelPtr) * hp.size(), hipMemcpyHostToDevice, s));
    int64_t *ncnt, *noff, *mcnt, *nrow;
    ALLOC(ncnt, int64_t, nleaves + 1);
    ALLOC(noff, int64_t, nleaves + 1);
    ALLOC(mcnt, int64_t, nleaves + 1);
    ALLOC(nrow, int64_t, nleaves + 1);
    near_kernel<0><<<grid_for(nleaves + 1), kBlock, 0, s>>>(nleaves, g_level, g_node, d_lvp, ncnt, nullptr, nullptr,
                                                            nullptr);
    TRY(hipGetLastError());
    TRY(excl_scan64(ncnt, noff, nleaves + 1));
    int64_t nraw = 0;
    TRY(read(&nraw, noff + nleaves));
    int64_t *rb, *re;
    ALLOC(rb, int64_t, nraw);
    ALLOC(re, int64_t, nraw);
    near_kernel<1><<<grid_for(nleaves + 1), kBlock, 0, s>>>(nleaves, g_level, g_node, d_lvp, mcnt, noff, rb, re);
    TRY(hipGetLastError());
    TRY(excl_scan64(mcnt, nrow, nleaves + 1));
    int64_t nnear = 0;
    TRY(read(&nnear, nrow + nleaves));
    int64_t *nbeg, *nend;
    ALLOC(nbeg, int64_t, nnear);
    ALLOC(nend, int64_t, nnear);
    near_compact_kernel<<<grid_for(nleaves), kBlock, 0, s>>>(nleaves, noff, nrow, rb, re, nbeg, nend);
    TRY(hipGetLastError());

    // 7. copy into the host tree
    t = Tree{};
    t.n = n;
    t.max_level = max_level;
    t.cap = cap;
    t.box = box;
    t.depth = depth;
    t.lv.resize(depth + 1);
    TRY(download(t.perm, perm_final, n, s));
    for (int l = 0; l <= depth; ++l) {
        const DevLevel& d = lv[l];
        TreeLevel& h = t.lv[l];
        TRY(download(h.code, d.code, d.nn, s));
        TRY(download(h.pbegin, d.pbegin, d.nn, s));
        TRY(download(h.pend, d.pend, d.nn, s));
        TRY(download(h.parent, d.parent, d.nn, s));
        TRY(download(h.child0, d.child0, d.nn, s));
        TRY(download(h.nchild, d.nchild, d.nn, s));
        TRY(download(h.node2leaf, d.node2leaf, d.nn, s));
        TRY(download(h.leaf2node, d.leaf2node, d.nleaf, s));
        TRY(download(h.nb_row, d.nb_row, int64_t(d.nn) + 1, s));
        TRY(download(h.nb, d.nb, d.n_nb, s));
        if (l == 0) {
            h.il_row = {0};
        } else {
            TRY(download(h.il_row, d.il_row, int64_t(d.nn) + 1, s));
            TRY(download(h.il_src, d.il_src, d.n_il, s));
            TRY(download(h.il_off, d.il_off, d.n_il, s));
        }
    }
    TRY(download(t.leaf_level, g_level, nleaves, s));
    TRY(download(t.leaf_node, g_node, nleaves, s));
    t.leaf_row = leaf_row;
    TRY(download(t.near_row, nrow, nleaves + 1, s));
    TRY(download(t.near_begin, nbeg, nnear, s));
    TRY(download(t.near_end, nend, nnear, s));
    TRY(hipStreamSynchronize(s));
    return {};
}

}  // namespace sfmm
