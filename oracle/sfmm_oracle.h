/*
 * CPU oracle for the Laplace FMM evaluation path — TEST INFRASTRUCTURE ONLY.
 *
 * A plain C++17 restatement of the reference algorithm (icyspirit/SimpleFMM:
 * box.hpp, zorder.hpp, partition.hpp, tree.hpp, fmm.hpp), used as the checker
 * for the HIP product path.  Only tests/, __graft_entry__.smoke() and the
 * cpu_baseline leg of bench.py may load it; the product never does.
 *
 * Every entry point is C-ABI so that Python (ctypes) can drive it.
 */
#ifndef SFMM_ORACLE_H
#define SFMM_ORACLE_H

#include <stdint.h>

#ifdef __cplusplus
extern "C" {
#endif

/* Evaluation modes (what a charge is and what is written per particle). */
enum {
    ORACLE_POTGRAD = 0,   /* scalar q; out[4i..4i+3] = phi, dphi/dx, dphi/dy, dphi/dz */
    ORACLE_VECPOT = 1,    /* 3-vector Q; out[3i..] = sum Q_j/|r|           (rinv)        */
    ORACLE_CURL = 2,      /* 3-vector Q; out[3i..] = sum Q_j x R/|R|^3      (rinv<true>)  */
    ORACLE_POT = 3        /* scalar q; out[i] = phi                                       */
};

/* Which part of the sum: far field (rinv_nonear), near field (N2N) or both. */
enum { ORACLE_FAR = 1, ORACLE_NEAR = 2, ORACLE_BOTH = 3 };

/* Tree: Box(positions, vertical_axis, phideg) + OctreePartitioner::refine(max_level, cap). */
void* oracle_tree_build(int64_t n, const double* pos, int max_level, int cap, int vertical_axis, double phideg);
void oracle_tree_free(void* tree);
int oracle_tree_depth(const void* tree);
/* OctreePartitioner::is_neighbor(i, j) (partition.hpp:484-488): j's leaf is in the near list of i's leaf. */
int oracle_is_neighbor(const void* tree, int64_t i, int64_t j);
/* box: width, corner x, y, z */
void oracle_tree_box(const void* tree, double* out4);
/* counts: n_node, n_leaf, nnz(indices), nnz(ilist), nnz(clist), nnz(nlist) */
void oracle_tree_counts(const void* tree, int level, int64_t* out6);
/* Copies one level's lists.  Row arrays have nrow+1 entries.  ilist values are
 * (src, di, dj, dk); clist values (child, child code); plist (parent, code);
 * nlist (level, leaf).  Any pointer may be NULL. */
void oracle_tree_copy(const void* tree, int level, int64_t* c_node, int64_t* leaf2node, int64_t* idx_row,
                      int64_t* idx_val, int64_t* ilist_row, int64_t* ilist_val, int64_t* clist_row,
                      int64_t* clist_val, int64_t* plist, int64_t* nlist_row, int64_t* nlist_val);

/* FMM evaluation with expansion order p.  out is accumulated into (+=), as the
 * reference's U.  m2l_dense selects the dense O(p^4) M2L (FMM_M2L_DIRECT) instead
 * of rotate-translate-rotate.  Returns 0 on success. */
int oracle_fmm(const void* tree, int p, int mode, const double* q, double* out, int which, int m2l_dense);

/* Same, with user self-exclusions: for particle i the pairs (i, j) for j in
 * self_val[self_row[i]..self_row[i+1]) are excluded instead of (i, i)
 * (FMM3D's self_generator constructor). */
int oracle_fmm_self(const void* tree, int p, int mode, const double* q, double* out, int which, int m2l_dense,
                    const int64_t* self_row, const int64_t* self_val);

/* Direct O(N * ntarget) sum over all sources j != target (main.cpp:70-96). */
int oracle_direct(int64_t n, const double* pos, int mode, const double* q, int64_t ntarget,
                  const int64_t* targets, double* out);

/* Building blocks, exposed for unit tests. */
double oracle_Z(int n, int m, double theta);                       /* sqrt(4pi/(2n+1)) sph_legendre(n,|m|,theta) */
double oracle_A(int n, int m);                                     /* 1/sqrt((n-m)!(n+m)!) */
double oracle_wigner_R(int n, int a, int b, double theta);         /* sgn(a)sgn(b) d^n_ab(-theta), long double */
/* Multipole of one unit source at offset (x,y,z) from the centre, p order,
 * full m range: out[2*(n(n+1)+m)] = Re, +1 = Im.  grad != 0 gives d/dx_d of it
 * for d = grad-1 (0..2). */
void oracle_harmonic(int p, double x, double y, double z, int grad, double* out);

#ifdef __cplusplus
}
#endif

#endif
