// Operator tables for the FMM plan (see tables.hpp for the storage convention).
//
// The operators are those of the reference (icyspirit/SimpleFMM fmm.hpp):
//   M2M  fmm.hpp:558-576   L2L  fmm.hpp:714-733
//   M2L  rotate - coaxial translate - rotate, fmm.hpp:620-688, tables :339-396
// restated here in the level-scaled real form.
#include "tables.hpp"

#include "m2l_dense.hpp"
#include "m2l_sched.hpp"

#include <algorithm>
#include <atomic>
#include <cmath>
#include <complex>
#include <thread>

namespace sfmm {

namespace {

// f(k) for k in [0, n) on up to 16 host threads (each k is independent).
template <class F>
void par_for(int n, F&& f)
{
    const unsigned hw = std::max(1u, std::thread::hardware_concurrency());
    const int nt = std::max(1, std::min(n, static_cast<int>(std::min(16u, hw))));
    std::atomic<int> next{0};
    auto work = [&]() {
        for (int k = next++; k < n; k = next++) f(k);
    };
    std::vector<std::thread> pool;
    for (int i = 1; i < nt; ++i) pool.emplace_back(work);
    work();
    for (auto& th : pool) th.join();
}

}  // namespace

namespace {

constexpr double kPi = 3.14159265358979323846;
using cd = std::complex<double>;

inline int iabs(int x) { return x < 0 ? -x : x; }
inline double neg1(int e) { return (e % 2 == 0) ? 1.0 : -1.0; }
inline int fkm(int k, int m) { return (iabs(k) - iabs(m) - iabs(k - m)) >> 1; }  // fmm.hpp:128-131
inline int ci(int n, int m) { return n * (n + 1) + m; }                           // complex index

double Acoef(int n, int m) { return 1 / std::sqrt(std::tgamma(n - m + 1) * std::tgamma(n + m + 1)); }

double Zval(int n, int m, double theta)  // fmm.hpp:107-111
{
    return std::sqrt(4 * kPi / (2 * n + 1)) * std::sph_legendre(n, iabs(m), theta);
}

long double fact(int k)
{
    long double f = 1;
    for (int i = 2; i <= k; ++i) f *= i;
    return f;
}

// child octant geometry (fmm.hpp:444-478): azimuth of (child - parent) and of
// (parent - child); the polar angle is acos(1/sqrt 3) or its supplement.
double phi_child(int c)
{
    static const double v[4] = {-0.75 * kPi, -0.25 * kPi, 0.75 * kPi, 0.25 * kPi};
    return v[c % 4];
}
double phi_parent(int c)
{
    static const double v[4] = {0.25 * kPi, 0.75 * kPi, -0.25 * kPi, -0.75 * kPi};
    return v[c % 4];
}

}  // namespace

void wigner_rotation(int n, long double theta, std::vector<double>& R)
{
    // Wigner's explicit sum for d^n_{ab}(beta) at beta = -theta
    const int w = 2 * n + 1;
    R.assign(static_cast<size_t>(w) * w, 0.0);
    const long double c = std::cos(-theta / 2), s = std::sin(-theta / 2);
    std::vector<long double> cp(2 * n + 1), sp(2 * n + 1);
    cp[0] = sp[0] = 1;
    for (int e = 1; e <= 2 * n; ++e) {
        cp[e] = cp[e - 1] * c;
        sp[e] = sp[e - 1] * s;
    }
    for (int a = -n; a <= n; ++a)
        for (int b = -n; b <= n; ++b) {
            const long double pref = std::sqrt(fact(n + a) * fact(n - a) * fact(n + b) * fact(n - b));
            long double sum = 0;
            for (int t = std::max(0, b - a); t <= std::min(n + b, n - a); ++t) {
                const long double term =
                    pref / (fact(n + b - t) * fact(t) * fact(a - b + t) * fact(n - a - t)) * cp[2 * n + b - a - 2 * t] *
                    sp[a - b + 2 * t];
                sum += ((a - b + t) % 2 == 0) ? term : -term;
            }
            const double sa = a < 0 ? neg1(a) : 1.0, sb = b < 0 ? neg1(b) : 1.0;
            R[static_cast<size_t>(a + n) * w + (b + n)] = sa * sb * static_cast<double>(sum);
        }
}

void complex_to_real_form(int p, const std::vector<double>& Tre, const std::vector<double>& Tim,
                          std::vector<double>& out)
{
    const int K = (p + 1) * (p + 1);
    out.assign(static_cast<size_t>(K) * K, 0.0);
    auto T = [&](int j, int k, int n, int m) {
        const size_t idx = static_cast<size_t>(ci(j, k)) * K + ci(n, m);
        return cd(Tre[idx], Tim[idx]);
    };
    for (int j = 0; j <= p; ++j)
        for (int k = 0; k <= j; ++k) {
            const size_t ro = static_cast<size_t>(rf_cos(j, k)) * K;
            const size_t rs = static_cast<size_t>(k > 0 ? rf_sin(j, k) : 0) * K;
            for (int n = 0; n <= p; ++n) {
                const cd t0 = T(j, k, n, 0);
                out[ro + rf_cos(n, 0)] = t0.real();
                if (k > 0) out[rs + rf_cos(n, 0)] = t0.imag();
                for (int m = 1; m <= n; ++m) {
                    const cd S = T(j, k, n, m) + T(j, k, n, -m);
                    const cd D = T(j, k, n, m) - T(j, k, n, -m);
                    out[ro + rf_cos(n, m)] = S.real();
                    out[ro + rf_sin(n, m)] = -D.imag();
                    if (k > 0) {
                        out[rs + rf_cos(n, m)] = S.imag();
                        out[rs + rf_sin(n, m)] = D.real();
                    }
                }
            }
        }
}

void build_m2l_fragments(const Tables& t, bool interleave_rows, std::vector<double>& rotA,
                         std::vector<double>& coaxA)
{
    auto mrow = [&](int l) { return interleave_rows ? 4 * (l % 4) + (l % 16) / 4 : l % 16; };
    const int P = t.p;
    const Packing p1 = pack_stage(P, 1), p2 = pack_stage(P, 2), p3 = pack_stage3(P);
    const int nrf = rot_nfrag(P), ncf = p2.nfrag();
    rotA.assign(static_cast<size_t>(70) * 2 * nrf * 64, 0.0);
    coaxA.assign(static_cast<size_t>(28) * ncf * 64, 0.0);
    for (int cls = 0; cls < 70; ++cls) {
        // forward rotation (stage 1): packed groups; lane l holds A[l % 16][l / 16]
        {
            const double* blocks = t.rot.data() + static_cast<size_t>(cls) * 2 * t.rot_size;
            double* out = rotA.data() + static_cast<size_t>(cls) * 2 * nrf * 64;
            for (int g = 0; g < p1.ng; ++g)
                for (int ks = 0; ks < p1.g[g].nks(); ++ks) {
                    double* f = out + static_cast<size_t>(p1.frag(g) + ks) * 64;
                    for (int l = 0; l < 64; ++l) {
                        const int i = mrow(l), fc = 4 * ks + l / 16;
                        const int qr = p1.piece_of_row(g, i), qc = p1.piece_of_col(g, fc);
                        double v = 0.0;
                        if (qr >= 0 && qr == qc) {
                            const Piece& pc = p1.g[g].pc[qr];
                            const int n = pc.n;
                            const int a = pc.r0 + i - p1.g[g].row0[qr], b = fc - p1.g[g].col0[qr];
                            v = pc.part == 0 ? blocks[t.rot_off[n] + a * (n + 1) + b]
                                             : blocks[t.rot_off[n] + (n + 1) * (n + 1) + a * n + b];
                        }
                        f[l] = v;
                    }
                }
        }
        // back rotation (stage 3): packed groups, the cos tile's k-steps then the sin tile's
        {
            const double* blocks = t.rot.data() + (static_cast<size_t>(cls) * 2 + 1) * t.rot_size;
            double* out = rotA.data() + (static_cast<size_t>(cls) * 2 + 1) * nrf * 64;
            for (int g = 0; g < p3.ng; ++g) {
                const int nkc = p3.g[g].nks();
                for (int part = 0; part < 2; ++part)
                    for (int ks = 0; ks < (part == 0 ? nkc : p3.g[g].nks_s()); ++ks) {
                        double* f = out + static_cast<size_t>(p3.frag3(g) + part * nkc + ks) * 64;
                        for (int l = 0; l < 64; ++l) {
                            const int i = mrow(l), fc = 4 * ks + l / 16;
                            const int qr = p3.piece_of_row(g, i);
                            const int qc = part == 0 ? p3.piece_of_col(g, fc) : p3.piece_of_col_s(g, fc);
                            double v = 0.0;
                            if (qr >= 0 && qr == qc) {
                                const Piece& pc = p3.g[g].pc[qr];
                                const int j = pc.n, a = pc.r0 + i - p3.g[g].row0[qr];
                                if (part == 0)
                                    v = blocks[t.rot_off[j] + a * (j + 1) + fc - p3.g[g].col0[qr]];
                                else if (a >= 1)
                                    v = blocks[t.rot_off[j] + (j + 1) * (j + 1) + (a - 1) * j + fc - p3.g[g].col0s[qr]];
                            }
                            f[l] = v;
                        }
                    }
            }
        }
    }
    // coaxial translation (stage 2): packed groups, the block of order k once per part
    for (int d2 = 0; d2 < 28; ++d2) {
        const double* C = t.coax.data() + static_cast<size_t>(d2) * t.coax_size;
        double* out = coaxA.data() + static_cast<size_t>(d2) * ncf * 64;
        for (int g = 0; g < p2.ng; ++g)
            for (int ks = 0; ks < p2.g[g].nks(); ++ks) {
                double* f = out + static_cast<size_t>(p2.frag(g) + ks) * 64;
                for (int l = 0; l < 64; ++l) {
                    const int i = mrow(l), fc = 4 * ks + l / 16;
                    const int qr = p2.piece_of_row(g, i), qc = p2.piece_of_col(g, fc);
                    double v = 0.0;
                    if (qr >= 0 && qr == qc) {
                        const Piece& pc = p2.g[g].pc[qr];
                        const int k = pc.n, w = P + 1 - k;
                        const int row = pc.r0 + i - p2.g[g].row0[qr], kk = fc - p2.g[g].col0[qr];
                        v = C[t.coax_off[k] + row * w + kk];
                    }
                    f[l] = v;
                }
            }
    }
}

void build_tables(int p, Tables& t)
{
    t = Tables{};
    t.p = p;
    const int K = (p + 1) * (p + 1);
    t.K = K;
    const double th_diag = std::acos(1 / std::sqrt(3.0));
    std::vector<double> zdiag(K);
    for (int n = 0; n <= p; ++n)
        for (int m = 0; m <= n; ++m) zdiag[ci(n, m)] = Zval(n, m, th_diag);

    // ---- M2M, scaled: M~p(j,k) = sum (-1)^fkm(k,m) A(n,m)A(j-n,k-m)/A(j,k)
    //      (sqrt .75)^n 2^-j Zc(n,-m) e^{-i m phi_c} M~c(j-n,k-m)
    t.m2m.resize(static_cast<size_t>(8) * K * K);
    t.l2l.resize(static_cast<size_t>(8) * K * K);
    par_for(8, [&](int cc) {
        std::vector<double> Tre(static_cast<size_t>(K) * K, 0.0), Tim(static_cast<size_t>(K) * K, 0.0);
        for (int j = 0; j <= p; ++j)
            for (int k = -j; k <= j; ++k)
                for (int n = 0; n <= j; ++n)
                    for (int m = std::max(-n, k - (j - n)); m <= std::min(n, k + (j - n)); ++m) {
                        const double zc = (((cc >> 2) % 2 == 0) ? neg1(n - iabs(m)) : 1.0) * zdiag[ci(n, iabs(m))];
                        const double v = neg1(fkm(k, m)) * Acoef(n, m) * Acoef(j - n, k - m) / Acoef(j, k) *
                                         std::pow(std::sqrt(0.75), n) * std::ldexp(1.0, -j) * zc;
                        const cd e = std::polar(1.0, -m * phi_child(cc));
                        const size_t idx = static_cast<size_t>(ci(j, k)) * K + ci(j - n, k - m);
                        Tre[idx] += v * e.real();
                        Tim[idx] += v * e.imag();
                    }
        std::vector<double> R;
        complex_to_real_form(p, Tre, Tim, R);
        std::copy(R.begin(), R.end(), t.m2m.begin() + static_cast<size_t>(cc) * K * K);

        // ---- L2L, scaled: L~c(j,k) = sum_{n>=j} (-1)^(fkm(m,m-k)+n+j) A(n-j,m-k)A(j,k)/A(n,m)
        //      (sqrt .75)^(n-j) 2^-(n+1) Zp(n-j,m-k) e^{i(m-k) phi_p} L~p(n,m)
        std::fill(Tre.begin(), Tre.end(), 0.0);
        std::fill(Tim.begin(), Tim.end(), 0.0);
        for (int j = 0; j <= p; ++j)
            for (int k = -j; k <= j; ++k)
                for (int n = j; n <= p; ++n)
                    for (int m = std::max(-n, k - (n - j)); m <= std::min(n, k + (n - j)); ++m) {
                        const int nn = n - j, mm = m - k;
                        const double zp = (((cc >> 2) % 2 == 0) ? 1.0 : neg1(nn - iabs(mm))) * zdiag[ci(nn, iabs(mm))];
                        const double v = neg1(fkm(m, m - k) + n + j) * Acoef(nn, mm) * Acoef(j, k) / Acoef(n, m) *
                                         std::pow(std::sqrt(0.75), nn) * std::ldexp(1.0, -(n + 1)) * zp;
                        const cd e = std::polar(1.0, mm * phi_parent(cc));
                        const size_t idx = static_cast<size_t>(ci(j, k)) * K + ci(n, m);
                        Tre[idx] += v * e.real();
                        Tim[idx] += v * e.imag();
                    }
        complex_to_real_form(p, Tre, Tim, R);
        std::copy(R.begin(), R.end(), t.l2l.begin() + static_cast<size_t>(cc) * K * K);
    });

    // ---- rotations for the 70 polar classes (t = di^2+dj^2 in tvals, dk in [-3,3])
    static const int tvals[10] = {0, 1, 2, 4, 5, 8, 9, 10, 13, 18};
    t.rot_off.resize(p + 2);
    t.rot_off[0] = 0;
    for (int n = 0; n <= p; ++n) t.rot_off[n + 1] = t.rot_off[n] + (n + 1) * (n + 1) + n * n;
    t.rot_size = t.rot_off[p + 1];
    t.rot.assign(static_cast<size_t>(70) * 2 * t.rot_size, 0.0);
    par_for(70, [&](int cls) {
        const int ti = cls / 7, dk = cls % 7 - 3;
        {
            if (tvals[ti] == 0 && dk == 0) return;
            const long double theta = std::acos(dk / std::sqrt(static_cast<long double>(tvals[ti] + dk * dk)));
            double* fwd = t.rot.data() + static_cast<size_t>(7 * ti + dk + 3) * 2 * t.rot_size;
            double* bwd = fwd + t.rot_size;
            std::vector<double> R;
            for (int n = 0; n <= p; ++n) {
                wigner_rotation(n, theta, R);
                const int w = 2 * n + 1;
                auto r = [&](int a, int b) { return R[static_cast<size_t>(a + n) * w + (b + n)]; };
                double* fc = fwd + t.rot_off[n];
                double* fs = fc + (n + 1) * (n + 1);
                double* bc = bwd + t.rot_off[n];
                double* bs = bc + (n + 1) * (n + 1);
                for (int a = 0; a <= n; ++a) {
                    fc[a * (n + 1) + 0] = r(a, 0);
                    bc[a * (n + 1) + 0] = r(0, a);
                    for (int b = 1; b <= n; ++b) {
                        fc[a * (n + 1) + b] = r(a, b) + r(a, -b);
                        bc[a * (n + 1) + b] = r(b, a) + r(-b, a);
                    }
                }
                for (int a = 1; a <= n; ++a)
                    for (int b = 1; b <= n; ++b) {
                        fs[(a - 1) * n + (b - 1)] = r(a, b) - r(a, -b);
                        bs[(a - 1) * n + (b - 1)] = r(b, a) - r(-b, a);
                    }
            }
        }
    });

    // ---- coaxial translation blocks for |d|^2 = 0..27
    t.coax_off.resize(p + 2);
    t.coax_off[0] = 0;
    for (int k = 0; k <= p; ++k) t.coax_off[k + 1] = t.coax_off[k] + (p + 1 - k) * (p + 1 - k);
    t.coax_size = t.coax_off[p + 1];
    t.coax.assign(static_cast<size_t>(28) * t.coax_size, 0.0);
    for (int d2 = 1; d2 <= 27; ++d2) {
        const double d = std::sqrt(static_cast<double>(d2));
        double* C = t.coax.data() + static_cast<size_t>(d2) * t.coax_size;
        for (int k = 0; k <= p; ++k) {
            const int w = p + 1 - k;
            for (int j = k; j <= p; ++j)
                for (int n = k; n <= p; ++n)
                    C[t.coax_off[k] + (j - k) * w + (n - k)] =
                        neg1(n + k) * Acoef(n, k) * Acoef(j, k) / Acoef(j + n, 0) * std::pow(d, -(j + n + 1));
        }
    }

    // ---- per offset: polar class, |d|^2, azimuth table
    t.off_class.assign(343, -1);
    t.off_d2.assign(343, 0);
    t.off_cs.assign(static_cast<size_t>(343) * (p + 1) * 2, 0.0);
    for (int dk = -3; dk <= 3; ++dk)
        for (int dj = -3; dj <= 3; ++dj)
            for (int di = -3; di <= 3; ++di) {
                const int id = ((dk + 3) * 7 + (dj + 3)) * 7 + (di + 3);
                const int tt = di * di + dj * dj;
                int ti = -1;
                for (int q = 0; q < 10; ++q)
                    if (tvals[q] == tt) ti = q;
                if (ti < 0 || (tt == 0 && dk == 0)) continue;
                t.off_class[id] = 7 * ti + dk + 3;
                t.off_d2[id] = tt + dk * dk;
                const double ph = std::atan2(static_cast<double>(dj), static_cast<double>(di));
                for (int m = 0; m <= p; ++m) {
                    const cd e = std::polar(1.0, m * ph);
                    t.off_cs[(static_cast<size_t>(id) * (p + 1) + m) * 2] = e.real();
                    t.off_cs[(static_cast<size_t>(id) * (p + 1) + m) * 2 + 1] = e.imag();
                }
            }

    // ---- order-major positions for the coaxial stage
    t.om.assign(K, 0);
    std::vector<int> ooff(p + 2, 0);
    ooff[0] = 0;
    ooff[1] = p + 1;
    for (int k = 1; k <= p; ++k) ooff[k + 1] = ooff[k] + 2 * (p + 1 - k);
    for (int n = 0; n <= p; ++n) {
        t.om[rf_cos(n, 0)] = n;
        for (int k = 1; k <= n; ++k) {
            t.om[rf_cos(n, k)] = ooff[k] + (n - k);
            t.om[rf_sin(n, k)] = ooff[k] + (p + 1 - k) + (n - k);
        }
    }
}

uint16_t f16_bits(double x)
{
    const uint16_t sign = std::signbit(x) ? 0x8000 : 0;
    const double a = std::fabs(x);
    if (a == 0.0) return sign;
    int e;
    (void)std::frexp(a, &e);  // a in [2^(e-1), 2^e)
    int E = std::max(e - 1, -14);  // below 2^-14: subnormal, quantum 2^-24
    int64_t m = static_cast<int64_t>(std::nearbyint(std::ldexp(a, 10 - E)));
    if (m == 2048) {
        ++E;
        m = 1024;
    }
    if (m < 1024) return static_cast<uint16_t>(sign | m);
    return static_cast<uint16_t>(sign | ((E + 15) << 10) | (m - 1024));
}

double f16_value(uint16_t b)
{
    const int E = (b >> 10) & 31, m = b & 1023;
    const double v = E == 0 ? std::ldexp(static_cast<double>(m), -24) : std::ldexp(1024.0 + m, E - 25);
    return (b & 0x8000) ? -v : v;
}

void build_dense_m2l(const Tables& t, std::vector<uint16_t>& ops, int& op_exp, std::vector<float>& az,
                     std::vector<int32_t>& off_grp)
{
    const int P = t.p;
    const DenseOrder O = dense_order(P);
    // the polar classes of interaction offsets (not adjacent: some |d_i| >= 2) and
    // their |d|^2; a class holds only adjacent or only non-adjacent offsets, and
    // the adjacent ones (|d| < 2, never translated) would set the scale
    std::vector<int> cls_d2(70, -1);
    for (int id = 0; id < 343; ++id) {
        const int di = id % 7 - 3, dj = (id / 7) % 7 - 3, dk = id / 49 - 3;
        if (t.off_class[id] >= 0 && std::max({std::abs(di), std::abs(dj), std::abs(dk)}) >= 2)
            cls_d2[t.off_class[id]] = t.off_d2[id];
    }
    // dense blocks per class and part, [row][column] in the rows' and the
    // columns' order (16 mtc rows by kc columns each)
    const int nrow = 16 * O.mtc, ncol = O.kc;
    const size_t per_part = static_cast<size_t>(nrow) * ncol;
    std::vector<double> D(70 * 2 * per_part, 0.0);
    std::vector<double> cmax(70, 0.0);
    par_for(70, [&](int c) {
        if (cls_d2[c] < 1) return;
        const double* fwd = t.rot.data() + static_cast<size_t>(c) * 2 * t.rot_size;
        const double* bwd = fwd + t.rot_size;
        const double* C = t.coax.data() + static_cast<size_t>(cls_d2[c]) * t.coax_size;
        for (int part = 0; part < 2; ++part) {
            double* Dp = D.data() + (static_cast<size_t>(c) * 2 + part) * per_part;
            for (int kr = 0; kr < nrow; ++kr) {
                const int er = dense_row_elem(O, part, kr);
                if (er < 0) continue;
                const int j = er / 32, a = er % 32;
                for (int kk = 0; kk < ncol; ++kk) {
                    const int ec = dense_elem(O, part, kk);
                    if (ec < 0) continue;
                    const int n = ec / 32, b = ec % 32;
                    double v = 0.0;
                    for (int q = part; q <= std::min(j, n); ++q) {
                        const int w = P + 1 - q;
                        const double tq = C[t.coax_off[q] + (j - q) * w + (n - q)];
                        if (part == 0)
                            v += bwd[t.rot_off[j] + a * (j + 1) + q] * tq * fwd[t.rot_off[n] + q * (n + 1) + b];
                        else
                            v += bwd[t.rot_off[j] + (j + 1) * (j + 1) + (a - 1) * j + (q - 1)] * tq *
                                 fwd[t.rot_off[n] + (n + 1) * (n + 1) + (q - 1) * n + (b - 1)];
                    }
                    Dp[static_cast<size_t>(kr) * ncol + kk] = v;
                    cmax[c] = std::max(cmax[c], std::fabs(v));
                }
            }
        }
    });
    const double dmax = *std::max_element(cmax.begin(), cmax.end());
    int e = 0;
    if (dmax > 0.0) (void)std::frexp(dmax, &e);  // dmax < 2^e
    op_exp = 14 - e;
    // fragments: (part, 16-row tile mt, 32-wide K-step ks), lane l element jj =
    // A[16 mt + l % 16][32 ks + 8 (l / 16) + jj] (the A operand of a 16x16x32 MFMA);
    // cos fragment mt * nks + ks, sin fragment (mtc + mt) * nks + ks
    const size_t per_class = static_cast<size_t>(O.chunks) * 512;  // padded to whole copy chunks
    ops.assign(per_class * 70, 0);
    par_for(70, [&](int c) {
        for (int part = 0; part < 2; ++part) {
            const double* Dp = D.data() + (static_cast<size_t>(c) * 2 + part) * per_part;
            const int mts = part == 0 ? O.mtc : O.mts;
            for (int mt = 0; mt < mts; ++mt)
                for (int ks = 0; ks < O.nks; ++ks) {
                    const int f = (part == 0 ? mt : O.mtc + mt) * O.nks + ks;
                    uint16_t* hi = ops.data() + per_class * c + static_cast<size_t>(f) * 1024;
                    uint16_t* lo = hi + 512;
                    for (int l = 0; l < 64; ++l)
                        for (int jj = 0; jj < 8; ++jj) {
                            const int row = 16 * mt + l % 16, col = 32 * ks + 8 * (l / 16) + jj;
                            const double x = std::ldexp(Dp[static_cast<size_t>(row) * ncol + col], op_exp);
                            const uint16_t bh = f16_bits(x);
                            hi[l * 8 + jj] = bh;
                            lo[l * 8 + jj] = f16_bits(x - f16_value(bh));
                        }
                }
        }
    });
    // azimuth groups: the primitive directions (a, b) of (di, dj), (0, 0) with (1, 0)
    std::vector<std::pair<int, int>> dirs;
    off_grp.assign(343, -1);
    az.assign(static_cast<size_t>(kAzGroups) * (P + 1) * 2, 0.0f);
    for (int id = 0; id < 343; ++id) {
        if (t.off_class[id] < 0) continue;
        int di = id % 7 - 3, dj = (id / 7) % 7 - 3;
        if (di == 0 && dj == 0) di = 1;
        int g0 = std::abs(di), g1 = std::abs(dj);
        while (g1) {
            const int r = g0 % g1;
            g0 = g1;
            g1 = r;
        }
        const std::pair<int, int> dir{di / g0, dj / g0};
        int g = static_cast<int>(std::find(dirs.begin(), dirs.end(), dir) - dirs.begin());
        if (g == static_cast<int>(dirs.size())) {
            dirs.push_back(dir);
            for (int m = 0; m <= P; ++m) {
                az[(static_cast<size_t>(g) * (P + 1) + m) * 2] =
                    static_cast<float>(t.off_cs[(static_cast<size_t>(id) * (P + 1) + m) * 2]);
                az[(static_cast<size_t>(g) * (P + 1) + m) * 2 + 1] =
                    static_cast<float>(t.off_cs[(static_cast<size_t>(id) * (P + 1) + m) * 2 + 1]);
            }
        }
        off_grp[id] = g;
    }
}

}  // namespace sfmm
