// Tiling of the offset-grouped M2L onto 16x16x4 MFMA tiles, shared by the host
// (which lays out the A-operand fragments) and the kernel (which consumes them).
//
// For one interaction offset the rotate-translate-rotate operator is three
// block-diagonal products applied to 16 stacked columns (targets x components):
//   stage 1  Y = P X   per degree n: cos block rows a=0..n, cols b=0..n;
//                      sin block rows a=0..n (row 0 zero), cols b=1..n
//   stage 2  Z = C Y   per order k: rows j=k..p, cols n=k..p (cos and sin columns
//                      share the block)
//   stage 3  L = Q Z   per degree j, same block shapes as stage 1
// The blocks of each stage are packed block-diagonally into 16x4 MFMA tiles
// (pack_stage, pack_stage3 below); one "fragment" is the 16x4 A operand of one
// MFMA, stored as 64 values in lane order (lane l holds A[l % 16][l / 16]).
//
// Column-stacked buffers live in LDS in two row orders:
//   ps(n, b, part): degree-major, per degree the n+1 cos rows then the n sin rows
//   om(k, n, part): order-major, per order k the cos rows n=k..p then the sin rows
#pragma once

#if !defined(__HIPCC__) && !defined(__host__)
#define __host__
#define __device__
#endif

namespace sfmm {

__host__ __device__ constexpr int ceil_div(int a, int b) { return (a + b - 1) / b; }
__host__ __device__ constexpr int ps_row(int n, int b, int part) { return part == 0 ? n * n + b : n * n + n + b; }
__host__ __device__ constexpr int om_base(int P, int k) { return k == 0 ? 0 : (P + 1) + 2 * ((k - 1) * (P + 1) - (k - 1) * k / 2); }
__host__ __device__ constexpr int om_row(int P, int k, int n, int part)
{
    return om_base(P, k) + (part == 0 ? 0 : (P + 1 - k)) + (n - k);
}

// Fragments per class and direction of one tile set per rotation block (16-row
// m-tiles, 4-column k-steps): the stride of the rotation fragment table, which
// bounds every packed layout.
__host__ __device__ constexpr int rot_nks(int n, int part) { return part == 0 ? ceil_div(n + 1, 4) : ceil_div(n, 4); }
__host__ __device__ constexpr int rot_mts(int n) { return ceil_div(n + 1, 16); }
__host__ __device__ constexpr int rot_nfrag(int P)
{
    int f = 0;
    for (int n = 0; n <= P; ++n) f += rot_mts(n) * (rot_nks(n, 0) + rot_nks(n, 1));
    return f;
}

// Packed tiles of stages 1 and 2.  The blocks of a stage (stage 1: cos block
// of degree n, rows a = 0..n, cols b = 0..n, and sin block, rows a = 1..n,
// cols b = 1..n; stage 2: block of order k, rows j = k..P, cols n = k..P, once
// per part) are cut into pieces of at most 16 rows and packed first-fit
// decreasing into groups of at most 16 rows.  A group is one MFMA tile: its
// pieces sit block-diagonally, rows stacked in insertion order and columns
// concatenated, and it runs ceil(columns / 4) k-steps (a k-step may take
// columns of two pieces; the A operand is zero across pieces).  Against one
// tile set per block this removes the k-padding of the small blocks: at P = 12,
// 43 instead of 52 MFMAs per stage.
constexpr int kMaxPieces = 8;
constexpr int kMaxGroups = 40;
struct Piece {
    int n = 0, part = 0, r0 = 0, rows = 0, cols = 0;  // n: degree (stage 1) or order k (stage 2)
};
struct Group {
    int np = 0, rows = 0, cols = 0;
    Piece pc[kMaxPieces] = {};
    int row0[kMaxPieces] = {}, col0[kMaxPieces] = {};
    // stage 3 only: the sin tile of the group (same rows, columns b = 1..j per piece)
    int cols_s = 0;
    int col0s[kMaxPieces] = {};
    __host__ __device__ constexpr int nks() const { return ceil_div(cols, 4); }
    __host__ __device__ constexpr int nks_s() const { return ceil_div(cols_s, 4); }
};
struct Packing {
    int ng = 0;
    Group g[kMaxGroups] = {};
    __host__ __device__ constexpr int frag(int gi) const
    {
        int f = 0;
        for (int i = 0; i < gi; ++i) f += g[i].nks();
        return f;
    }
    __host__ __device__ constexpr int nfrag() const { return frag(ng); }
    // piece holding packed row i (column f), or -1
    __host__ __device__ constexpr int piece_of_row(int gi, int i) const
    {
        for (int q = 0; q < g[gi].np; ++q)
            if (i >= g[gi].row0[q] && i < g[gi].row0[q] + g[gi].pc[q].rows) return q;
        return -1;
    }
    __host__ __device__ constexpr int piece_of_col(int gi, int f) const
    {
        for (int q = 0; q < g[gi].np; ++q)
            if (f >= g[gi].col0[q] && f < g[gi].col0[q] + g[gi].pc[q].cols) return q;
        return -1;
    }
    __host__ __device__ constexpr int piece_of_col_s(int gi, int f) const
    {
        for (int q = 0; q < g[gi].np; ++q)
            if (f >= g[gi].col0s[q] && f < g[gi].col0s[q] + g[gi].pc[q].n) return q;
        return -1;
    }
    // stage 3: fragments of group gi (cos k-steps, then sin k-steps)
    __host__ __device__ constexpr int frag3(int gi) const
    {
        int f = 0;
        for (int i = 0; i < gi; ++i) f += g[i].nks() + g[i].nks_s();
        return f;
    }
};

__host__ __device__ constexpr Packing pack_stage(int P, int stage)
{
    Piece pcs[4 * 18] = {};
    int np = 0;
    for (int n = 0; n <= P; ++n)
        for (int part = 0; part < 2; ++part) {
            int size = 0;
            if (stage == 1) size = part == 0 ? n + 1 : n;
            else size = (part == 1 && n == 0) ? 0 : P + 1 - n;  // n plays the order k
            for (int r0 = 0; r0 < size; r0 += 16) {
                Piece q;
                q.n = n;
                q.part = part;
                q.r0 = r0;
                q.rows = size - r0 < 16 ? size - r0 : 16;
                q.cols = size;
                pcs[np++] = q;
            }
        }
    // stable sort by rows, descending
    for (int i = 1; i < np; ++i) {
        const Piece x = pcs[i];
        int j = i - 1;
        while (j >= 0 && pcs[j].rows < x.rows) {
            pcs[j + 1] = pcs[j];
            --j;
        }
        pcs[j + 1] = x;
    }
    Packing pk;
    for (int i = 0; i < np; ++i) {
        int gi = 0;
        while (gi < pk.ng && (pk.g[gi].rows + pcs[i].rows > 16 || pk.g[gi].np == kMaxPieces)) ++gi;
        if (gi == pk.ng) ++pk.ng;
        Group& G = pk.g[gi];
        G.pc[G.np] = pcs[i];
        G.row0[G.np] = G.rows;
        G.col0[G.np] = G.cols;
        G.rows += pcs[i].rows;
        G.cols += pcs[i].cols;
        ++G.np;
    }
    return pk;
}

// Stage 3 (back rotation of degree j, then the azimuth rotation that pairs the
// cos and sin rows of each order a): a piece is rows a = r0.. of degree j; the
// group's cos tile and sin tile share its row layout (the sin row a = 0 is
// zero) and have their own column layouts (cos b = 0..j, sin b = 1..j).
__host__ __device__ constexpr Packing pack_stage3(int P)
{
    Piece pcs[2 * 18] = {};
    int np = 0;
    for (int j = 0; j <= P; ++j)
        for (int r0 = 0; r0 <= j; r0 += 16) {
            Piece q;
            q.n = j;
            q.r0 = r0;
            q.rows = j + 1 - r0 < 16 ? j + 1 - r0 : 16;
            q.cols = j + 1;
            pcs[np++] = q;
        }
    for (int i = 1; i < np; ++i) {
        const Piece x = pcs[i];
        int k = i - 1;
        while (k >= 0 && pcs[k].rows < x.rows) {
            pcs[k + 1] = pcs[k];
            --k;
        }
        pcs[k + 1] = x;
    }
    Packing pk;
    for (int i = 0; i < np; ++i) {
        int gi = 0;
        while (gi < pk.ng && (pk.g[gi].rows + pcs[i].rows > 16 || pk.g[gi].np == kMaxPieces)) ++gi;
        if (gi == pk.ng) ++pk.ng;
        Group& G = pk.g[gi];
        G.pc[G.np] = pcs[i];
        G.row0[G.np] = G.rows;
        G.col0[G.np] = G.cols;
        G.col0s[G.np] = G.cols_s;
        G.rows += pcs[i].rows;
        G.cols += pcs[i].cols;
        G.cols_s += pcs[i].n;
        ++G.np;
    }
    return pk;
}

template <int P>
struct Packed {
    static constexpr Packing s1 = pack_stage(P, 1);
    static constexpr Packing s2 = pack_stage(P, 2);
    static constexpr Packing s3 = pack_stage3(P);
};

// Static work split over the NW wavefronts of a workgroup: longest tiles
// first, each to the least loaded wave.  PIPE: the three stages run in the same
// phase (pipelined kernel), so the split balances each wave's total over all
// stages; otherwise each stage is balanced on its own (barrier between stages).
constexpr int kMaxWaves = 16;
struct WaveTiles {
    int count[kMaxWaves] = {};
    int item[kMaxWaves][48] = {};  // packed (index << 8) | (part << 4) | mt
};

template <int P, int NW, bool PIPE>
struct M2LSchedule {
    WaveTiles s1, s2, s3;  // items: packed group index (stage 3: cos and sin tiles of the group)

    constexpr M2LSchedule()
    {
        // all tiles: (stage, packed item, cost)
        int st[3 * 64] = {}, it[3 * 64] = {}, cost[3 * 64] = {};
        int nt = 0;
        const Packing p1 = pack_stage(P, 1), p2 = pack_stage(P, 2);
        for (int g = 0; g < p1.ng; ++g) {
            st[nt] = 1;
            it[nt] = g;
            cost[nt++] = p1.g[g].nks();
        }
        for (int g = 0; g < p2.ng; ++g) {
            st[nt] = 2;
            it[nt] = g;
            cost[nt++] = p2.g[g].nks();
        }
        const Packing p3 = pack_stage3(P);
        for (int g = 0; g < p3.ng; ++g) {
            st[nt] = 3;
            it[nt] = g;
            cost[nt++] = p3.g[g].nks() + p3.g[g].nks_s();
        }
        int load[3][kMaxWaves] = {};
        bool done[3 * 64] = {};
        for (int rep = 0; rep < nt; ++rep) {
            int b = -1;
            for (int t = 0; t < nt; ++t)
                if (!done[t] && (b < 0 || cost[t] > cost[b])) b = t;
            done[b] = true;
            int* ld = load[PIPE ? 0 : st[b] - 1];
            int w = 0;
            for (int v = 1; v < NW; ++v)
                if (ld[v] < ld[w]) w = v;
            ld[w] += cost[b];
            WaveTiles& T = st[b] == 1 ? s1 : st[b] == 2 ? s2 : s3;
            T.item[w][T.count[w]++] = it[b];
        }
    }
};

}  // namespace sfmm
