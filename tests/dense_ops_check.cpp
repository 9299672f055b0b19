// CPU check of the dense M2L operators of the mixed-precision kernel
// (tables.cpp: build_dense_m2l, layout in m2l_dense.hpp) against the
// rotate-translate-rotate stages applied one after the other, as the row-form
// M2L kernel does (fmm.hpp:620-688 in real form): for random multipole
// vectors and every interaction offset, the dense f16 hi/lo operator with the
// azimuth rotations must reproduce the staged result to ~2^-20 of its norm.
// Prints one line "max_rel <value> offsets <count>" and exits 0.
#include "../simplefmm_amd/csrc/m2l_dense.hpp"
#include "../simplefmm_amd/csrc/tables.hpp"

#include <cmath>
#include <cstdio>
#include <cstdlib>
#include <random>
#include <vector>

using namespace sfmm;

int main(int argc, char** argv)
{
    const int p = argc > 1 ? std::atoi(argv[1]) : 12;
    Tables t;
    build_tables(p, t);
    std::vector<uint16_t> ops;
    std::vector<float> az;
    std::vector<int32_t> grp;
    int op_exp = 0;
    build_dense_m2l(t, ops, op_exp, az, grp);
    const DenseOrder O = dense_order(p);
    const int K = (p + 1) * (p + 1);
    std::mt19937_64 rng(7);
    std::uniform_real_distribution<double> U(-1, 1);
    double worst = 0.0;
    int count = 0;
    for (int id = 0; id < 343; ++id) {
        const int cls = t.off_class[id];
        if (cls < 0) continue;
        const int di = id % 7 - 3, dj = (id / 7) % 7 - 3, dk = id / 49 - 3;
        if (std::abs(di) <= 1 && std::abs(dj) <= 1 && std::abs(dk) <= 1) continue;  // near offsets
        ++count;
        std::vector<double> M(K);
        for (auto& v : M) v = U(rng);
        const double* cs = t.off_cs.data() + static_cast<size_t>(id) * (p + 1) * 2;
        // staged: azimuth, forward rotation, coaxial translation, back rotation, azimuth
        std::vector<double> X(K), Y(K, 0.0), Z(K, 0.0), Lref(K, 0.0);
        for (int n = 0; n <= p; ++n) {
            X[rf_cos(n, 0)] = M[rf_cos(n, 0)];
            for (int m = 1; m <= n; ++m) {
                const double xr = M[rf_cos(n, m)], xi = M[rf_sin(n, m)], c = cs[2 * m], s = cs[2 * m + 1];
                X[rf_cos(n, m)] = xr * c - xi * s;
                X[rf_sin(n, m)] = xr * s + xi * c;
            }
        }
        const double* fwd = t.rot.data() + static_cast<size_t>(cls) * 2 * t.rot_size;
        const double* bwd = fwd + t.rot_size;
        for (int n = 0; n <= p; ++n)
            for (int a = 0; a <= n; ++a) {
                double sc = 0, ss = 0;
                for (int b = 0; b <= n; ++b) sc += fwd[t.rot_off[n] + a * (n + 1) + b] * X[rf_cos(n, b)];
                Y[rf_cos(n, a)] = sc;
                if (a == 0) continue;
                for (int b = 1; b <= n; ++b) ss += fwd[t.rot_off[n] + (n + 1) * (n + 1) + (a - 1) * n + b - 1] * X[rf_sin(n, b)];
                Y[rf_sin(n, a)] = ss;
            }
        const double* C = t.coax.data() + static_cast<size_t>(t.off_d2[id]) * t.coax_size;
        for (int k = 0; k <= p; ++k) {
            const int w = p + 1 - k;
            for (int j = k; j <= p; ++j) {
                double sc = 0, ss = 0;
                for (int n = k; n <= p; ++n) {
                    const double c = C[t.coax_off[k] + (j - k) * w + (n - k)];
                    sc += c * Y[rf_cos(n, k)];
                    if (k > 0) ss += c * Y[rf_sin(n, k)];
                }
                Z[rf_cos(j, k)] = sc;
                if (k > 0) Z[rf_sin(j, k)] = ss;
            }
        }
        for (int j = 0; j <= p; ++j)
            for (int a = 0; a <= j; ++a) {
                double lc = 0, ls = 0;
                for (int b = 0; b <= j; ++b) lc += bwd[t.rot_off[j] + a * (j + 1) + b] * Z[rf_cos(j, b)];
                if (a > 0)
                    for (int b = 1; b <= j; ++b) ls += bwd[t.rot_off[j] + (j + 1) * (j + 1) + (a - 1) * j + b - 1] * Z[rf_sin(j, b)];
                const double c = cs[2 * a], s = cs[2 * a + 1];
                Lref[rf_cos(j, a)] = lc * c + ls * s;
                if (a > 0) Lref[rf_sin(j, a)] = ls * c - lc * s;
            }
        // dense: the same azimuth (from the FP32 group table), operator decoded from the fragments
        const int g = grp[id];
        if (g < 0) { std::printf("offset %d without an azimuth group\n", id); return 1; }
        const float* ag = az.data() + static_cast<size_t>(g) * (p + 1) * 2;
        std::vector<double> Bc(O.kc, 0.0), Bs(O.ks, 0.0);
        for (int r = 0; r < K; ++r) {
            const int pos = dense_pos(O, r);
            if (pos < 0) { std::printf("coefficient %d has no position\n", r); return 1; }
        }
        for (int n = 0; n <= p; ++n)
            for (int m = 0; m <= n; ++m) {
                const double xr = M[rf_cos(n, m)], xi = m ? M[rf_sin(n, m)] : 0.0;
                const double c = ag[2 * m], s = ag[2 * m + 1];
                Bc[dense_pos(O, rf_cos(n, m))] = xr * c - xi * s;
                if (m) Bs[dense_pos(O, rf_sin(n, m)) - O.kc] = xr * s + xi * c;
            }
        const size_t per_class = static_cast<size_t>(O.chunks) * 512;
        auto entry = [&](int part, int row, int col) {
            const int mt = row / 32, ks = col / 16, l = row % 32 + 32 * ((col % 16) / 8), jj = col % 8;
            const int f = part == 0 ? mt * O.ksc + ks : O.mtc * O.ksc + mt * O.kss + ks;
            const uint16_t* hi = ops.data() + per_class * cls + static_cast<size_t>(f) * 1024;
            return std::ldexp(f16_value(hi[l * 8 + jj]) + f16_value(hi[512 + l * 8 + jj]), -op_exp);
        };
        std::vector<double> Lc(O.kc, 0.0), Ls(O.ks, 0.0);
        for (int r = 0; r < O.kc; ++r)
            for (int c = 0; c < O.kc; ++c) Lc[r] += entry(0, r, c) * Bc[c];
        for (int r = 0; r < O.ks; ++r)
            for (int c = 0; c < O.ks; ++c) Ls[r] += entry(1, r, c) * Bs[c];
        std::vector<double> Ld(K, 0.0);
        for (int j = 0; j <= p; ++j)
            for (int a = 0; a <= j; ++a) {
                const double lc = Lc[dense_pos(O, rf_cos(j, a))], ls = a ? Ls[dense_pos(O, rf_sin(j, a)) - O.kc] : 0.0;
                const double c = ag[2 * a], s = ag[2 * a + 1];
                Ld[rf_cos(j, a)] = lc * c + ls * s;
                if (a > 0) Ld[rf_sin(j, a)] = ls * c - lc * s;
            }
        double num = 0, den = 0;
        for (int r = 0; r < K; ++r) {
            num += (Ld[r] - Lref[r]) * (Ld[r] - Lref[r]);
            den += Lref[r] * Lref[r];
        }
        worst = std::max(worst, std::sqrt(num / den));
    }
    std::printf("max_rel %.3e offsets %d\n", worst, count);
    return 0;
}
