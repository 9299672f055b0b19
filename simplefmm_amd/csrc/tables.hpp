// Operator tables for the FMM plan, computed on the host in FP64 (rotation
// matrices from long double) and uploaded once per plan.
//
// Storage convention ("real form"): an expansion with complex coefficients
// c(n,m), |m| <= n <= p, and c(n,-m) = conj(c(n,m)) (true for real charges) is
// stored as (p+1)^2 reals, degree n occupying [n^2, (n+1)^2):
//     [ Re c(n,0), Re c(n,1), Im c(n,1), ..., Re c(n,n), Im c(n,n) ].
// Coefficients are also scaled by the cell width w_l = width/2^l of their level:
// multipoles store c(n,m)/w_l^n, locals store c(n,m)*w_l^(n+1).  With that
// scaling every translation operator is the same at every level.
#pragma once

#include <cstdint>
#include <vector>

namespace sfmm {

inline int rf_cos(int n, int m) { return m == 0 ? n * n : n * n + 2 * m - 1; }  // Re c(n,m)
inline int rf_sin(int n, int m) { return n * n + 2 * m; }                         // Im c(n,m), m >= 1

struct Tables {
    int p = 0;
    int K = 0;  // (p+1)^2
    // child -> parent (M2M) and parent -> child (L2L) for the 8 child octants,
    // dense K x K row-major, row = output real-form index
    std::vector<double> m2m, l2l;
    // rotate-translate-rotate M2L pieces
    //   polar class c (0..69, from (di^2+dj^2, dk)): per degree n the forward
    //   rotation cos block (n+1)x(n+1) then sin block n x n, then the backward
    //   rotation blocks in the same shape; rot_off(n) gives each degree's start
    //   inside one direction's blocks, rot_size the size of one direction.
    std::vector<double> rot;    // [class][dir 0=fwd,1=back][rot_size]
    int rot_size = 0;
    std::vector<int> rot_off;   // per degree
    //   coaxial translation for |d|^2 = 0..27: per order k >= 0 a (p+1-k)^2 block,
    //   entry (j,n) = tz_k(j,n) |d|^-(j+n+1), j,n = k..p
    std::vector<double> coax;   // [d2][coax_size]
    int coax_size = 0;
    std::vector<int> coax_off;  // per order k
    //   per offset id ((dk+3)*7+(dj+3))*7+(di+3): polar class, |d|^2, azimuth
    //   table cos(m phi), sin(m phi) for m = 0..p
    std::vector<int32_t> off_class, off_d2;
    std::vector<double> off_cs;  // [343][p+1][2]
    // order-major index of the coaxial stage: (k, n, part) with part 0 = cos
    std::vector<int32_t> om;     // om[rf index] -> order-major position
};

// MFMA A-operand fragments of the M2L blocks (layout in m2l_sched.hpp):
// rotA [70 classes][2 directions][nfrag][64], coaxA [28 |d|^2][nfrag][64].
// Lane l of fragment (mt, ks) holds operator element (16 mt + i', 4 ks + l / 16)
// with i' = l % 16, or, with interleave_rows (FP32 MFMA, whose output register r
// covers MFMA rows 4 (l / 16) + r), i' = 4 (i % 4) + i / 4 for i = l % 16, so
// that register r holds operator rows [4r, 4r + 4) as in the FP64 layout.
void build_m2l_fragments(const Tables& t, bool interleave_rows, std::vector<double>& rotA,
                         std::vector<double>& coaxA);

// Operators of the mixed-precision M2L on the f16 matrix cores (m2l_dense.hpp),
// p <= 12.  Per polar class c, the azimuth-aligned operator D_c = R_c^T T R_c
// (back rotation, coaxial translation, forward rotation) in its cos and sin
// blocks, rows and columns in dense_order(p), scaled by 2^op_exp (largest entry
// below 2^14) and split as hi = f16(x), lo = f16(x - hi); stored as the A
// fragments of 16x16x32 MFMA tiles, ops[class][fragment][hi, lo][64 lanes][8]
// with lane l, element j of fragment (mt, ks) holding entry
// (16 mt + l % 16, 32 ks + 8 (l / 16) + j).  az[group][m] = (cos m phi, sin m phi)
// of the azimuth groups and off_grp the group of each offset id (-1 for none).
void build_dense_m2l(const Tables& t, std::vector<uint16_t>& ops, int& op_exp, std::vector<float>& az,
                     std::vector<int32_t>& off_grp);

// IEEE binary16 bits of x rounded to nearest even (|x| < 65520), and back.
uint16_t f16_bits(double x);
double f16_value(uint16_t b);

// Builds all tables for expansion order p (1 <= p <= 20).
void build_tables(int p, Tables& t);

// Polar rotation R_n[a][b] = sgn(a) sgn(b) d^n_{ab}(-theta), a, b = -n..n, as in
// the reference (fmm.hpp:114-125,339-377); sgn(m) = (-1)^m for m < 0 else 1.
void wigner_rotation(int n, long double theta, std::vector<double>& R);

// Dense complex -> real-form conversion of a linear map y(j,k) = sum T x(n,m)
// over full m ranges, for inputs with conjugate symmetry.  Tre/Tim are indexed
// [(j(j+1)+k) * K + n(n+1)+m] with K = (p+1)^2; out is K x K real form.
void complex_to_real_form(int p, const std::vector<double>& Tre, const std::vector<double>& Tim,
                          std::vector<double>& out);

}  // namespace sfmm
