// Dispatch and fragment layout of the matrix-core M2M / L2L (translate_mfma.hpp).
#include "translate_mfma.hpp"

namespace sfmm {

bool translate_lower(const TransArgs& a, LevelSpan b, int p, int nc, double* io, hipStream_t s);
bool translate_upper(const TransArgs& a, LevelSpan b, int p, int nc, double* io, hipStream_t s);

void launch_translate(const TransArgs& a, LevelSpan blocks, int p, int nc, bool upper, double* io, hipStream_t s)
{
    if (blocks.count == 0) return;
    if (upper) translate_upper(a, blocks, p, nc, io, s);
    else translate_lower(a, blocks, p, nc, io, s);
}

int translate_nfrag(int p, bool upper) { return tr_nfrag(p, upper); }

// Fragment (m-tile mt, k-step ks) holds A[16 mt + l % 16][4 ks + l / 16] in lane l.
int translate_frag_index(int p, bool upper, int row, int col, int* lane)
{
    const int K = tr_K(p);
    if (row < 0 || row >= K || col < 0 || col >= K) return -1;
    const int mt = row / 16, ks = col / 4;
    if (ks < tr_kb(p, upper, mt) || ks >= tr_ke(p, upper, mt)) return -1;
    *lane = (row % 16) + 16 * (col % 4);
    return tr_frag(p, upper, mt) + ks - tr_kb(p, upper, mt);
}

}  // namespace sfmm
