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
        // dense, as the kernel computes it: the multipole in the FP32 copy's order,
        // columns rotated with the slot's azimuth factor (1 for order 0), the
        // operator decoded from the fragments by the MFMA lane layout, rows rotated
        // back by their order and written to their real-form index
        const int g = grp[id];
        if (g < 0) { std::printf("offset %d without an azimuth group\n", id); return 1; }
        const float* ag = az.data() + static_cast<size_t>(g) * (p + 1) * 2;
        std::vector<double> X32(O.kd, 0.0);
        for (int r = 0; r < K; ++r) {
            const int pos = dense_pos(O, r);
            if (pos < 0 || pos >= O.kd) { std::printf("coefficient %d has no position\n", r); return 1; }
            X32[pos] = M[r];
        }
        std::vector<double> Bc(O.kc, 0.0), Bs(O.kc, 0.0);
        for (int k = 0; k < O.kc; ++k) {
            const int s = dense_kslot(k), q = dense_kq(k);
            const int m = dense_slot_m(O, s);
            const bool zero = (dense_slot_zero(O, s) >> q) & 1;
            const double c = zero ? 1.0 : ag[2 * m], sn = zero ? 0.0 : ag[2 * m + 1];
            Bc[k] = X32[k] * c - X32[O.kc + k] * sn;
            Bs[k] = X32[k] * sn + X32[O.kc + k] * c;
        }
        const size_t per_class = static_cast<size_t>(O.chunks) * 512;
        auto entry = [&](int part, int row, int col) {
            const int mt = row / 16, ks = col / 32, l = row % 16 + 16 * ((col % 32) / 8), jj = col % 8;
            const int f = (part == 0 ? mt : O.mtc + mt) * O.nks + ks;
            const uint16_t* hi = ops.data() + per_class * cls + static_cast<size_t>(f) * 1024;
            return std::ldexp(f16_value(hi[l * 8 + jj]) + f16_value(hi[512 + l * 8 + jj]), -op_exp);
        };
        std::vector<double> Ld(K, 0.0);
        std::vector<int> seen(K, 0);
        for (int row = 0; row < 16 * O.mtc; ++row) {
            double lc = 0, ls = 0;
            for (int c = 0; c < O.kc; ++c) lc += entry(0, row, c) * Bc[c];
            if (row < 16 * O.mts)
                for (int c = 0; c < O.kc; ++c) ls += entry(1, row, c) * Bs[c];
            const int ec = dense_row_elem(O, 0, row), es = dense_row_elem(O, 1, row);
            const int m = ec < 0 ? 0 : ec % 32;
            const double c = ag[2 * m], sn = ag[2 * m + 1];
            if (ec >= 0) {
                Ld[dense_rf(ec, 0)] = lc * c + ls * sn;
                ++seen[dense_rf(ec, 0)];
            }
            if (es >= 0) {
                Ld[dense_rf(es, 1)] = ls * c - lc * sn;
                ++seen[dense_rf(es, 1)];
            }
        }
        for (int r = 0; r < K; ++r)
            if (seen[r] != 1) { std::printf("coefficient %d written %d times\n", r, seen[r]); return 1; }
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
