// CPU oracle for the Laplace FMM evaluation path — TEST INFRASTRUCTURE ONLY.
//
// A plain restatement of the reference algorithm, written for clarity rather
// than speed, used to check the HIP product path.  Only tests/, the smoke test
// and bench.py's cpu_baseline leg may load it.  Each function names the
// reference lines (icyspirit/SimpleFMM) whose behaviour it restates.
//
// Conventions (fmm.hpp:69-131):
//   coefficient index nm(n,m) = n(n+1)+m, |m| <= n <= p, (p+1)^2 per expansion
//   Z(n,m,theta) = sqrt(4pi/(2n+1)) * std::sph_legendre(n,|m|,theta)
//   A(n,m) = 1/sqrt((n-m)!(n+m)!)
//   multipole of a unit source at (rho,theta,phi) from the centre:
//        rho^n Z(n,m,theta) e^{-i m phi}
//   local evaluation: sum rho^n Z(n,m,theta) Re(L_n^m e^{i m phi})
#include "sfmm_oracle.h"

#include <algorithm>
#include <array>
#include <cmath>
#include <complex>
#include <cstring>
#include <memory>
#include <utility>
#include <vector>

namespace {

using cd = std::complex<double>;
constexpr double kPi = 3.14159265358979323846;

inline int iabs(int m) { return m < 0 ? -m : m; }
inline double neg1(int e) { return (e % 2 == 0) ? 1.0 : -1.0; }  // (-1)^e, fmm.hpp:93-96
inline int nmi(int n, int m) { return n * (n + 1) + m; }           // fmm.hpp:69-72
inline int fkm(int k, int m) { return (iabs(k) - iabs(m) - iabs(k - m)) >> 1; }  // fmm.hpp:128-131

double Zf(int n, int m, double theta)  // fmm.hpp:107-111
{
    return std::sqrt(4 * kPi / (2 * n + 1)) * std::sph_legendre(n, iabs(m), theta);
}

double Af(int n, int m)  // fmm.hpp:99-104
{
    return 1 / std::sqrt(std::tgamma(n - m + 1) * std::tgamma(n + m + 1));
}

// ---------------------------------------------------------------- Morton codes
// zorder.hpp:26-58: bit b of axis index a goes to bit 3b+a; 20 bits per axis.
int64_t spread_axis(int32_t v, int axis)
{
    int64_t r = 0;
    for (int b = 0; b < 20; ++b) {
        if ((v >> b) & 1) r |= int64_t(1) << (3 * b + axis);
    }
    return r;
}

int32_t gather_axis(int64_t z, int axis)
{
    int32_t r = 0;
    for (int b = 0; b < 20; ++b) {
        if ((z >> (3 * b + axis)) & 1) r |= int32_t(1) << b;
    }
    return r;
}

struct I3 {
    int i, j, k;
};

I3 from_code(int64_t z) { return {gather_axis(z, 0), gather_axis(z, 1), gather_axis(z, 2)}; }
int64_t to_code(const I3& v) { return spread_axis(v.i, 0) | spread_axis(v.j, 1) | spread_axis(v.k, 2); }

// ---------------------------------------------------------------- Box (box.hpp)
struct Box {
    int X = 0, Y = 1, Z = 2;
    double phi = 0;
    double width = 1;
    double corner[3] = {0, 0, 0};

    void rotate(const double* v, double* r) const  // box.hpp:152-157
    {
        r[0] = std::cos(phi) * v[X] - std::sin(phi) * v[Y];
        r[1] = std::sin(phi) * v[X] + std::cos(phi) * v[Y];
        r[2] = v[Z];
    }

    int64_t code(int level, const double* v) const  // box.hpp:57-63,119-122
    {
        double r[3];
        rotate(v, r);
        const int32_t imax = (int32_t(1) << level) - 1;
        int64_t c = 0;
        for (int d = 0; d < 3; ++d) {
            const double u = (r[d] - corner[d]) / width;
            int32_t iv = static_cast<int32_t>(u * (1 << level));
            iv = std::min(std::max(iv, int32_t(0)), imax);
            c |= spread_axis(iv, d);
        }
        return c;
    }

    void centre(int level, int64_t c, double* out) const  // box.hpp:66-70,147-150
    {
        for (int d = 0; d < 3; ++d) {
            out[d] = corner[d] + width * ((gather_axis(c, d) + 0.5) / (1 << level));
        }
    }

    Box(int64_t n, const double* pos, int vertical_axis, double phideg)  // box.hpp:84-107
    {
        X = (vertical_axis + 1) % 3;
        Y = (vertical_axis + 2) % 3;
        Z = vertical_axis;
        phi = phideg * (kPi / 180);
        double lo[3], hi[3];
        for (int64_t i = 0; i < n; ++i) {
            double r[3];
            rotate(pos + 3 * i, r);
            for (int d = 0; d < 3; ++d) {
                if (i == 0 || r[d] < lo[d]) lo[d] = r[d];
                if (i == 0 || r[d] > hi[d]) hi[d] = r[d];
            }
        }
        double dl[3], c[3];
        for (int d = 0; d < 3; ++d) {
            dl[d] = hi[d] - lo[d];
            c[d] = (hi[d] + lo[d]) / 2;
        }
        const double dmax = std::max(dl[0], std::max(dl[1], dl[2]));
        width = (1 + 2 * 1e-3) * dl[2];
        while (width > 0 && width < dmax) {
            c[2] += width / 2;
            width *= 2;
        }
        for (int d = 0; d < 3; ++d) corner[d] = c[d] - width / 2;
    }
};

// ---------------------------------------------------------------- octree (partition.hpp)
struct Node {
    int l;
    int64_t c;
    std::vector<int> idx;
    std::vector<Node> ch;
};

struct Level {
    std::vector<int64_t> code;     // c_node, ascending
    std::vector<int> leaf2node;    // _i_leaf
    std::vector<int> node2leaf;    // -1 if internal
    std::vector<int> idx_row{0}, idx_val;
    std::vector<int> il_row{0}, il_src;
    std::vector<I3> il_d;
    std::vector<int> cl_row{0}, cl_child;
    std::vector<int64_t> cl_code;
    std::vector<int> pl_parent;
    std::vector<int> nl_row{0}, nl_level, nl_leaf;

    int find(int64_t c) const
    {
        auto it = std::lower_bound(code.begin(), code.end(), c);
        return (it != code.end() && *it == c) ? int(it - code.begin()) : -1;
    }
};

struct Tree {
    int64_t n = 0;
    const double* pos = nullptr;
    std::vector<double> pos_copy;
    Box box;
    int depth = 0;
    std::vector<Level> lv;
    std::vector<std::pair<int, int>> part;  // particle -> (level, leaf)

    Tree(int64_t n_, const double* p, int vertical_axis, double phideg) :
        n(n_), pos_copy(p, p + 3 * n_), box(n_, p, vertical_axis, phideg)
    {
        pos = pos_copy.data();
    }
};

void refine(Node& nd, const Tree& t, int max_level, int cap)  // partition.hpp:46-61,92-103
{
    if (!(nd.l < max_level && static_cast<int>(nd.idx.size()) > cap)) return;
    std::vector<std::pair<int64_t, int>> ci;
    ci.reserve(nd.idx.size());
    for (int i : nd.idx) ci.push_back({t.box.code(nd.l + 1, t.pos + 3 * i), i});
    std::sort(ci.begin(), ci.end());
    size_t a = 0;
    while (a < ci.size()) {
        size_t b = a;
        while (b < ci.size() && ci[b].first == ci[a].first) ++b;
        Node child{nd.l + 1, ci[a].first, {}, {}};
        for (size_t k = a; k < b; ++k) child.idx.push_back(ci[k].second);
        nd.ch.push_back(std::move(child));
        a = b;
    }
    nd.idx.clear();
    nd.idx.shrink_to_fit();
    for (auto& c : nd.ch) refine(c, t, max_level, cap);
}

void collect(const Node& nd, std::vector<Level>& lv)  // DFS, partition.hpp:168-211
{
    Level& L = lv[nd.l];
    L.code.push_back(nd.c);
    if (!nd.idx.empty()) {
        L.leaf2node.push_back(static_cast<int>(L.code.size()) - 1);
        for (int i : nd.idx) L.idx_val.push_back(i);
        L.idx_row.push_back(static_cast<int>(L.idx_val.size()));
    }
    for (const auto& c : nd.ch) collect(c, lv);
}

int max_depth(const Node& nd)
{
    int d = nd.l;
    for (const auto& c : nd.ch) d = std::max(d, max_depth(c));
    return d;
}

// Neighbour window of cell v at level l: [max(x-1,0), min(x+2, 2^l)) per axis,
// visited with i fastest, then j, then k (tree.hpp:66-119).
template <class F>
void for_neighbours(int l, const I3& v, F&& f)
{
    const int imax = 1 << l;
    const int i0 = std::max(v.i - 1, 0), i1 = std::min(v.i + 2, imax);
    const int j0 = std::max(v.j - 1, 0), j1 = std::min(v.j + 2, imax);
    const int k0 = std::max(v.k - 1, 0), k1 = std::min(v.k + 2, imax);
    for (int k = k0; k < k1; ++k)
        for (int j = j0; j < j1; ++j)
            for (int i = i0; i < i1; ++i) f(I3{i, j, k});
}

void build_lists(Tree& t)
{
    const int D = t.depth;
    for (int l = 0; l <= D; ++l) {
        Level& L = t.lv[l];
        L.node2leaf.assign(L.code.size(), -1);
        for (size_t k = 0; k < L.leaf2node.size(); ++k) L.node2leaf[L.leaf2node[k]] = static_cast<int>(k);
    }
    // ilist, partition.hpp:283-297 with tree.hpp:122-174: children of the
    // parent's neighbours that are not adjacent to the node, rows sorted by src.
    for (int l = 2; l <= D; ++l) {
        Level& L = t.lv[l];
        for (size_t a = 0; a < L.code.size(); ++a) {
            const I3 me = from_code(L.code[a]);
            const I3 par = from_code(L.code[a] >> 3);
            std::vector<std::pair<int, I3>> row;
            for_neighbours(l - 1, par, [&](const I3& nb) {
                const int64_t zc = to_code(nb) << 3;
                for (int o = 0; o < 8; ++o) {
                    const I3 c = from_code(zc + o);
                    if (iabs(c.i - me.i) <= 1 && iabs(c.j - me.j) <= 1 && iabs(c.k - me.k) <= 1) continue;
                    const int s = L.find(zc + o);
                    if (s >= 0) row.push_back({s, I3{c.i - me.i, c.j - me.j, c.k - me.k}});
                }
            });
            std::sort(row.begin(), row.end(), [](const auto& x, const auto& y) { return x.first < y.first; });
            for (const auto& e : row) {
                L.il_src.push_back(e.first);
                L.il_d.push_back(e.second);
            }
            L.il_row.push_back(static_cast<int>(L.il_src.size()));
        }
    }
    // clist, partition.hpp:299-313
    for (int l = 0; l < D; ++l) {
        Level& L = t.lv[l];
        const Level& C = t.lv[l + 1];
        for (size_t a = 0; a < L.code.size(); ++a) {
            for (int o = 0; o < 8; ++o) {
                const int64_t z = (L.code[a] << 3) + o;
                const int ci = C.find(z);
                if (ci >= 0) {
                    L.cl_child.push_back(ci);
                    L.cl_code.push_back(z);
                }
            }
            L.cl_row.push_back(static_cast<int>(L.cl_child.size()));
        }
    }
    // plist, partition.hpp:315-322
    for (int l = 1; l <= D; ++l) {
        Level& L = t.lv[l];
        for (int64_t c : L.code) L.pl_parent.push_back(t.lv[l - 1].find(c >> 3));
    }
    // nlist, partition.hpp:324-366
    for (int l = 0; l <= D; ++l) {
        Level& L = t.lv[l];
        for (int node : L.leaf2node) {
            std::vector<std::pair<int, int>> row;
            const int64_t c = L.code[node];
            for (int la = 0; la < l; ++la) {
                const Level& A = t.lv[la];
                for_neighbours(la, from_code(c >> (3 * (l - la))), [&](const I3& nb) {
                    const int s = A.find(to_code(nb));
                    if (s >= 0 && A.node2leaf[s] >= 0) row.push_back({la, A.node2leaf[s]});
                });
            }
            std::vector<std::pair<int, int>> stack;
            for_neighbours(l, from_code(c), [&](const I3& nb) {
                const int s = L.find(to_code(nb));
                if (s >= 0) stack.push_back({l, s});
            });
            // descendants, depth first in child order (order is irrelevant: rows are sorted)
            while (!stack.empty()) {
                auto [ll, s] = stack.back();
                stack.pop_back();
                const Level& Q = t.lv[ll];
                if (Q.node2leaf[s] >= 0) {
                    row.push_back({ll, Q.node2leaf[s]});
                } else {
                    for (int k = Q.cl_row[s]; k < Q.cl_row[s + 1]; ++k) stack.push_back({ll + 1, Q.cl_child[k]});
                }
            }
            std::sort(row.begin(), row.end());
            for (const auto& e : row) {
                L.nl_level.push_back(e.first);
                L.nl_leaf.push_back(e.second);
            }
            L.nl_row.push_back(static_cast<int>(L.nl_level.size()));
        }
    }
    // partitions, partition.hpp:436-445
    t.part.assign(t.n, {0, 0});
    for (int l = 0; l <= D; ++l) {
        const Level& L = t.lv[l];
        for (size_t f = 0; f < L.leaf2node.size(); ++f)
            for (int k = L.idx_row[f]; k < L.idx_row[f + 1]; ++k) t.part[L.idx_val[k]] = {l, static_cast<int>(f)};
    }
}

// ---------------------------------------------------------------- harmonics
struct Sph {
    double rho, theta, phi;
};

Sph to_sph(const double* d)  // fmm.hpp:221-225
{
    const double rho = std::sqrt(d[0] * d[0] + d[1] * d[1] + d[2] * d[2]);
    return {rho, std::acos(d[2] / rho), std::atan2(d[1], d[0])};
}

// rho^n Z(n,m,theta) e^{-i m phi} for all (n,m), n<=p (fmm.hpp:226-232,537-548)
void regular(int p, const Sph& s, cd* out)
{
    std::vector<cd> E(p + 1);
    const cd step{std::cos(s.phi), -std::sin(s.phi)};
    E[0] = 1;
    for (int m = 1; m <= p; ++m) E[m] = E[m - 1] * step;
    double rn = 1;
    for (int n = 0; n <= p; ++n) {
        for (int m = -n; m <= n; ++m) {
            const double v = rn * Zf(n, m, s.theta);
            out[nmi(n, m)] = v * (m >= 0 ? E[m] : std::conj(E[-m]));
        }
        rn *= s.rho;
    }
}

// Cartesian gradient (w.r.t. the source point) of rho^n Z(n,m,theta) e^{-i m phi},
// from the spherical components r, theta, phi (fmm.hpp:234-253).
void regular_grad(int p, const Sph& s, std::array<cd, 3>* out)
{
    const double st = std::sin(s.theta), ct = std::cos(s.theta);
    const double sp = std::sin(s.phi), cp = std::cos(s.phi);
    const double rhat[3] = {st * cp, st * sp, ct};
    const double that[3] = {ct * cp, ct * sp, -st};
    const double phat[3] = {-sp, cp, 0};
    out[0] = {cd(0), cd(0), cd(0)};
    double rnm1 = 1;
    for (int n = 1; n <= p; ++n) {
        for (int m = -n; m <= n; ++m) {
            const double zn = Zf(n, m, s.theta);
            const double a_r = n * rnm1 * zn;
            const double a_t = rnm1 / st * (std::sqrt(double((n + 1) * (n + 1) - m * m)) * Zf(n + 1, m, s.theta) -
                                            (n + 1) * ct * zn);
            const double a_p = m * rnm1 / st * zn;
            const cd vr(a_r * std::cos(-m * s.phi), a_r * std::sin(-m * s.phi));
            const cd vt(a_t * std::cos(-m * s.phi), a_t * std::sin(-m * s.phi));
            const double ang = -(m * s.phi + kPi / 2);
            const cd vp(a_p * std::cos(ang), a_p * std::sin(ang));
            for (int d = 0; d < 3; ++d) out[nmi(n, m)][d] = rhat[d] * vr + that[d] * vt + phat[d] * vp;
        }
        rnm1 *= s.rho;
    }
}

// ---------------------------------------------------------------- operators
long double wigner_d_ld(int n, int a, int b, long double beta)
{
    // Wigner's explicit sum for d^n_{ab}(beta)
    auto fact = [](int k) {
        long double f = 1;
        for (int i = 2; i <= k; ++i) f *= i;
        return f;
    };
    const long double c = std::cos(beta / 2), s = std::sin(beta / 2);
    const long double pref = std::sqrt(fact(n + a) * fact(n - a) * fact(n + b) * fact(n - b));
    long double sum = 0;
    for (int t = std::max(0, b - a); t <= std::min(n + b, n - a); ++t) {
        const long double term = pref / (fact(n + b - t) * fact(t) * fact(a - b + t) * fact(n - a - t)) *
                                 std::pow(c, 2 * n + b - a - 2 * t) * std::pow(s, a - b + 2 * t);
        sum += ((a - b + t) % 2 == 0) ? term : -term;
    }
    return sum;
}

struct Ops {
    int p;
    int K;                    // (p+1)^2
    std::vector<double> A;    // A(n,m), n <= 2p
    std::vector<double> zdiag;  // Z(n,|m|,acos(1/sqrt3)) at nmi(n,|m|), fmm.hpp:264-268
    std::vector<int> i2m;     // order of coefficient index
    // per interaction offset d in [-3,3]^3 (index (d.k+3)*49+(d.j+3)*7+(d.i+3)):
    std::vector<std::vector<double>> rot;   // R_n[a][b] blocks, fmm.hpp:339-377
    std::vector<std::vector<double>> zfar;  // Z(N,|M|,acos(|dk|/|d|)), N<=2p, fmm.hpp:270-283

    double getA(int n, int m) const { return A[n * (4 * p + 1) + (m + 2 * p)]; }
    static int rot_off(int n) { return n == 0 ? 0 : (2 * n - 1) * n * (2 * n + 1) / 3; }  // sum_{k<n}(2k+1)^2
    static int didx(const I3& d) { return (d.k + 3) * 49 + (d.j + 3) * 7 + (d.i + 3); }

    explicit Ops(int p_) : p(p_), K((p_ + 1) * (p_ + 1))
    {
        A.assign((2 * p + 1) * (4 * p + 1), 0.0);
        for (int n = 0; n <= 2 * p; ++n)
            for (int m = -n; m <= n; ++m) A[n * (4 * p + 1) + (m + 2 * p)] = Af(n, m);
        zdiag.assign(K, 0.0);
        const double th = std::acos(1 / std::sqrt(3.0));
        for (int n = 0; n <= p; ++n)
            for (int m = 0; m <= n; ++m) zdiag[nmi(n, m)] = Zf(n, m, th);
        i2m.resize(K);
        for (int n = 0; n <= p; ++n)
            for (int m = -n; m <= n; ++m) i2m[nmi(n, m)] = m;
        rot.resize(343);
        zfar.resize(343);
        for (int dk = -3; dk <= 3; ++dk)
            for (int dj = -3; dj <= 3; ++dj)
                for (int di = -3; di <= 3; ++di) {
                    if (iabs(di) <= 1 && iabs(dj) <= 1 && iabs(dk) <= 1) continue;
                    const I3 d{di, dj, dk};
                    const int d2 = di * di + dj * dj + dk * dk;
                    const long double th_l = std::acos(dk / std::sqrt(static_cast<long double>(d2)));
                    std::vector<double>& R = rot[didx(d)];
                    R.assign(rot_off(p + 1), 0.0);
                    for (int n = 0; n <= p; ++n)
                        for (int a = -n; a <= n; ++a)
                            for (int b = -n; b <= n; ++b) {
                                const double sa = a < 0 ? neg1(a) : 1.0, sb = b < 0 ? neg1(b) : 1.0;
                                R[rot_off(n) + (a + n) * (2 * n + 1) + (b + n)] =
                                    sa * sb * static_cast<double>(wigner_d_ld(n, a, b, -th_l));
                            }
                    std::vector<double>& Zt = zfar[didx(d)];
                    const double thf = std::acos(iabs(dk) / std::sqrt(double(d2)));
                    Zt.assign((2 * p + 1) * (2 * p + 1), 0.0);
                    for (int N = 0; N <= 2 * p; ++N)
                        for (int M = 0; M <= N; ++M) Zt[N * (2 * p + 1) + M] = Zf(N, M, thf);
                }
    }

    // fmm.hpp:444-460 (child) and :462-478 (parent)
    static double phi_child(int64_t c)
    {
        switch (c % 4) {
        case 0: return -0.75 * kPi;
        case 1: return -0.25 * kPi;
        case 2: return 0.75 * kPi;
        default: return 0.25 * kPi;
        }
    }
    static double phi_parent(int64_t c)
    {
        switch (c % 4) {
        case 0: return 0.25 * kPi;
        case 1: return 0.75 * kPi;
        case 2: return -0.25 * kPi;
        default: return -0.75 * kPi;
        }
    }
    double z_child(int64_t c, int n, int m) const
    {
        return (((c >> 2) % 2 == 0) ? neg1(n - iabs(m)) : 1.0) * zdiag[nmi(n, iabs(m))];
    }
    double z_parent(int64_t c, int n, int m) const
    {
        return (((c >> 2) % 2 == 0) ? 1.0 : neg1(n - iabs(m))) * zdiag[nmi(n, iabs(m))];
    }

    // child -> parent, fmm.hpp:558-576; rho = sqrt(0.75) * (child cell width)
    void m2m(double wc, int64_t cc, const cd* Mc, cd* Mp) const
    {
        const double rho = std::sqrt(0.75) * wc;
        const double ph = phi_child(cc);
        for (int j = 0; j <= p; ++j)
            for (int k = -j; k <= j; ++k) {
                cd acc = 0;
                double rn = 1;
                for (int n = 0; n <= j; ++n) {
                    for (int m = std::max(-n, k - (j - n)); m <= std::min(n, k + (j - n)); ++m) {
                        const double v = neg1(fkm(k, m)) * getA(n, m) * getA(j - n, k - m) / getA(j, k) * rn *
                                         z_child(cc, n, -m);
                        acc += v * std::polar(1.0, -m * ph) * Mc[nmi(j - n, k - m)];
                    }
                    rn *= rho;
                }
                Mp[nmi(j, k)] += acc;
            }
    }

    // parent -> child, fmm.hpp:714-733
    void l2l(double wc, int64_t cc, const cd* Lp, cd* Lc) const
    {
        const double rho = std::sqrt(0.75) * wc;
        const double ph = phi_parent(cc);
        for (int j = 0; j <= p; ++j)
            for (int k = -j; k <= j; ++k) {
                cd acc = 0;
                double rn = 1;
                for (int n = j; n <= p; ++n) {
                    for (int m = std::max(-n, k - (n - j)); m <= std::min(n, k + (n - j)); ++m) {
                        const double v = neg1(fkm(m, m - k) + n + j) * getA(n - j, m - k) * getA(j, k) / getA(n, m) *
                                         rn * z_parent(cc, n - j, m - k);
                        acc += v * std::polar(1.0, (m - k) * ph) * Lp[nmi(n, m)];
                    }
                    rn *= rho;
                }
                Lc[nmi(j, k)] += acc;
            }
    }

    // dense M2L, fmm.hpp:592-618 with tables :270-283,317-337; d = source - target
    void m2l_dense(double w, const I3& d, const cd* Ms, cd* Lt) const
    {
        const int d2 = d.i * d.i + d.j * d.j + d.k * d.k;
        const double rho = w * std::sqrt(double(d2));
        const double ph = std::atan2(double(d.j), double(d.i));
        const std::vector<double>& Zt = zfar[didx(d)];
        std::vector<double> pw(2 * p + 2);
        pw[0] = 1 / rho;
        for (int e = 0; e < 2 * p; ++e) pw[e + 1] = pw[e] / rho;
        for (int j = 0; j <= p; ++j)
            for (int k = -j; k <= j; ++k) {
                cd acc = 0;
                for (int n = 0; n <= p; ++n)
                    for (int m = -n; m <= n; ++m) {
                        double c = neg1(fkm(k - m, k) + n) * getA(n, m) * getA(j, k) / getA(j + n, m - k);
                        if (d.k < 0) c *= neg1((j + n) - (m - k));
                        const double v = c * pw[j + n] * Zt[(j + n) * (2 * p + 1) + iabs(m - k)];
                        acc += v * std::polar(1.0, (m - k) * ph) * Ms[nmi(n, m)];
                    }
                Lt[nmi(j, k)] += acc;
            }
    }

    // rotate - coaxial translate - rotate, fmm.hpp:620-688 with tables :339-396
    void m2l_rtr(double w, const I3& d, const cd* Ms, cd* Lt) const
    {
        const int d2 = d.i * d.i + d.j * d.j + d.k * d.k;
        const double rho = w * std::sqrt(double(d2));
        const double ph = std::atan2(double(d.j), double(d.i));
        const std::vector<double>& Rt = rot[didx(d)];
        std::vector<double> pw(2 * p + 2);
        pw[0] = 1 / rho;
        for (int e = 0; e < 2 * p; ++e) pw[e + 1] = pw[e] / rho;
        std::vector<cd> Mt(K), W(K), Lz(K);
        for (int n = 0; n <= p; ++n)
            for (int m = -n; m <= n; ++m) Mt[nmi(n, m)] = std::polar(1.0, m * ph) * Ms[nmi(n, m)];
        auto R = [&](int n, int a, int b) { return Rt[rot_off(n) + (a + n) * (2 * n + 1) + (b + n)]; };
        for (int n = 0; n <= p; ++n)
            for (int a = -n; a <= n; ++a) {
                cd acc = 0;
                for (int b = -n; b <= n; ++b) acc += R(n, a, b) * Mt[nmi(n, b)];
                W[nmi(n, a)] = acc;
            }
        for (int k = -p; k <= p; ++k) {
            const int ak = iabs(k);
            for (int j = ak; j <= p; ++j) {
                cd acc = 0;
                for (int n = ak; n <= p; ++n) {
                    const double tz = neg1(n + ak) * getA(n, ak) * getA(j, ak) / getA(j + n, 0);
                    acc += tz * pw[j + n] * W[nmi(n, k)];
                }
                Lz[nmi(j, k)] = acc;
            }
        }
        for (int j = 0; j <= p; ++j)
            for (int a = -j; a <= j; ++a) {
                cd acc = 0;
                for (int b = -j; b <= j; ++b) acc += R(j, b, a) * Lz[nmi(j, b)];
                Lt[nmi(j, a)] += std::polar(1.0, -a * ph) * acc;
            }
    }
};

// ---------------------------------------------------------------- evaluation
int ncomp_of(int mode)
{
    switch (mode) {
    case ORACLE_POTGRAD: return 4;
    case ORACLE_VECPOT: return 3;
    case ORACLE_CURL: return 3;
    case ORACLE_POT: return 1;
    default: return -1;
    }
}

int qdim_of(int mode) { return (mode == ORACLE_VECPOT || mode == ORACLE_CURL) ? 3 : 1; }

// near-field kernel for target i, source j (fmm.hpp:786-806, main.cpp:78-91)
inline void kernel(int mode, const double* xi, const double* xj, const double* qj, double* u, double sign)
{
    const double R[3] = {xi[0] - xj[0], xi[1] - xj[1], xi[2] - xj[2]};
    const double r = std::sqrt(R[0] * R[0] + R[1] * R[1] + R[2] * R[2]);
    switch (mode) {
    case ORACLE_POT: u[0] += sign * qj[0] / r; break;
    case ORACLE_VECPOT:
        for (int c = 0; c < 3; ++c) u[c] += sign * qj[c] / r;
        break;
    case ORACLE_CURL: {
        const double r3 = std::pow(r, 3.0);
        u[0] += sign * (qj[1] * R[2] - qj[2] * R[1]) / r3;
        u[1] += sign * (qj[2] * R[0] - qj[0] * R[2]) / r3;
        u[2] += sign * (qj[0] * R[1] - qj[1] * R[0]) / r3;
        break;
    }
    case ORACLE_POTGRAD: {
        const double r3 = std::pow(r, 3.0);
        u[0] += sign * qj[0] / r;
        for (int c = 0; c < 3; ++c) u[1 + c] -= sign * qj[0] * R[c] / r3;
        break;
    }
    default: break;
    }
}

bool nlist_contains(const Tree& t, int i, int j)  // partition.hpp:484-488
{
    const auto [l, f] = t.part[i];
    const Level& L = t.lv[l];
    const auto pj = t.part[j];
    for (int k = L.nl_row[f]; k < L.nl_row[f + 1]; ++k)
        if (L.nl_level[k] == pj.first && L.nl_leaf[k] == pj.second) return true;
    return false;
}

int evaluate(const Tree& t, int p, int mode, const double* q, double* out, int which, bool dense,
             const int64_t* self_row, const int64_t* self_val)
{
    const int nc = ncomp_of(mode);
    if (nc < 0 || p < 0 || p > 20) return 1;
    const int qd = qdim_of(mode);
    const int nout = nc;  // outputs per particle
    const int D = t.depth;
    const int64_t n = t.n;

    if ((which & ORACLE_FAR) && D >= 2) {  // rinv_nonear, fmm.hpp:811-862
        const Ops ops(p);
        const int K = ops.K;
        std::vector<std::vector<cd>> M(D + 1), L(D + 1);
        for (int l = 0; l <= D; ++l) {
            M[l].assign(t.lv[l].code.size() * nc * K, cd(0));
            L[l].assign(t.lv[l].code.size() * nc * K, cd(0));
        }
        auto p2m = [&](int l) {  // N2M, fmm.hpp:524-556
            const Level& Lv = t.lv[l];
#pragma omp parallel for schedule(dynamic, 4)
            for (size_t f = 0; f < Lv.leaf2node.size(); ++f) {
                const int node = Lv.leaf2node[f];
                double ctr[3];
                t.box.centre(l, Lv.code[node], ctr);
                std::vector<cd> R(K);
                std::vector<std::array<cd, 3>> G(K);
                cd* Mn = &M[l][static_cast<size_t>(node) * nc * K];
                for (int k = Lv.idx_row[f]; k < Lv.idx_row[f + 1]; ++k) {
                    const int i = Lv.idx_val[k];
                    double x[3];
                    t.box.rotate(t.pos + 3 * i, x);
                    const double dlt[3] = {x[0] - ctr[0], x[1] - ctr[1], x[2] - ctr[2]};
                    const Sph s = to_sph(dlt);
                    const double* qi = q + qd * i;
                    if (mode == ORACLE_POT || mode == ORACLE_VECPOT || mode == ORACLE_POTGRAD) regular(p, s, R.data());
                    if (mode == ORACLE_CURL || mode == ORACLE_POTGRAD) regular_grad(p, s, G.data());
                    for (int a = 0; a < K; ++a) {
                        switch (mode) {
                        case ORACLE_POT: Mn[a] += qi[0] * R[a]; break;
                        case ORACLE_VECPOT:
                            for (int c = 0; c < 3; ++c) Mn[c * K + a] += qi[c] * R[a];
                            break;
                        case ORACLE_CURL:
                            Mn[0 * K + a] += qi[1] * G[a][2] - qi[2] * G[a][1];
                            Mn[1 * K + a] += qi[2] * G[a][0] - qi[0] * G[a][2];
                            Mn[2 * K + a] += qi[0] * G[a][1] - qi[1] * G[a][0];
                            break;
                        case ORACLE_POTGRAD:
                            Mn[a] += qi[0] * R[a];
                            for (int c = 0; c < 3; ++c) Mn[(1 + c) * K + a] -= qi[0] * G[a][c];
                            break;
                        }
                    }
                }
            }
        };
        p2m(D);
        for (int l = D - 1; l >= 2; --l) {  // M2M, fmm.hpp:578-590
            const Level& Lv = t.lv[l];
            const double wc = t.box.width / (1 << (l + 1));
#pragma omp parallel for schedule(dynamic, 16)
            for (size_t a = 0; a < Lv.code.size(); ++a)
                for (int k = Lv.cl_row[a]; k < Lv.cl_row[a + 1]; ++k)
                    for (int c = 0; c < nc; ++c)
                        ops.m2m(wc, Lv.cl_code[k], &M[l + 1][(static_cast<size_t>(Lv.cl_child[k]) * nc + c) * K],
                                &M[l][(a * nc + c) * K]);
            p2m(l);
        }
        for (int l = 2; l <= D; ++l) {  // M2L, fmm.hpp:700-712
            const Level& Lv = t.lv[l];
            const double w = t.box.width / (1 << l);
#pragma omp parallel for schedule(dynamic, 4)
            for (size_t a = 0; a < Lv.code.size(); ++a)
                for (int k = Lv.il_row[a]; k < Lv.il_row[a + 1]; ++k)
                    for (int c = 0; c < nc; ++c) {
                        const cd* Ms = &M[l][(static_cast<size_t>(Lv.il_src[k]) * nc + c) * K];
                        cd* Lt = &L[l][(a * nc + c) * K];
                        if (dense) ops.m2l_dense(w, Lv.il_d[k], Ms, Lt);
                        else ops.m2l_rtr(w, Lv.il_d[k], Ms, Lt);
                    }
        }
        auto l2p = [&](int l) {  // L2N, fmm.hpp:747-771
            const Level& Lv = t.lv[l];
#pragma omp parallel for schedule(dynamic, 4)
            for (size_t f = 0; f < Lv.leaf2node.size(); ++f) {
                const int node = Lv.leaf2node[f];
                double ctr[3];
                t.box.centre(l, Lv.code[node], ctr);
                const cd* Ln = &L[l][static_cast<size_t>(node) * nc * K];
                std::vector<double> zr(K);
                std::vector<cd> F(p + 1);
                for (int k = Lv.idx_row[f]; k < Lv.idx_row[f + 1]; ++k) {
                    const int i = Lv.idx_val[k];
                    double x[3];
                    t.box.rotate(t.pos + 3 * i, x);
                    const double dlt[3] = {x[0] - ctr[0], x[1] - ctr[1], x[2] - ctr[2]};
                    const Sph s = to_sph(dlt);
                    double rn = 1;
                    for (int nn = 0; nn <= p; ++nn) {
                        for (int m = -nn; m <= nn; ++m) zr[nmi(nn, m)] = rn * Zf(nn, m, s.theta);
                        rn *= s.rho;
                    }
                    const cd step{std::cos(s.phi), std::sin(s.phi)};
                    F[0] = 1;
                    for (int m = 1; m <= p; ++m) F[m] = F[m - 1] * step;
                    for (int c = 0; c < nc; ++c) {
                        double acc = 0;
                        for (int a = 0; a < K; ++a) {
                            const int m = ops.i2m[a];
                            const cd e = m >= 0 ? F[m] : std::conj(F[-m]);
                            const cd lv = Ln[c * K + a];
                            acc += zr[a] * (lv.real() * e.real() - lv.imag() * e.imag());
                        }
                        out[static_cast<size_t>(i) * nout + c] += acc;
                    }
                }
            }
        };
        for (int l = 2; l < D; ++l) {  // fmm.hpp:850-856
            l2p(l);
            const Level& Lv = t.lv[l + 1];
            const double wc = t.box.width / (1 << (l + 1));
#pragma omp parallel for schedule(dynamic, 16)
            for (size_t a = 0; a < Lv.code.size(); ++a)
                for (int c = 0; c < nc; ++c)
                    ops.l2l(wc, Lv.code[a], &L[l][(static_cast<size_t>(Lv.pl_parent[a]) * nc + c) * K],
                            &L[l + 1][(a * nc + c) * K]);
        }
        l2p(D);
    }

    if (which & ORACLE_NEAR) {  // N2N, fmm.hpp:773-809
#pragma omp parallel for schedule(dynamic, 64)
        for (int64_t i = 0; i < n; ++i) {
            const auto [li, fi] = t.part[i];
            const Level& Li = t.lv[li];
            const double* xi = t.pos + 3 * i;
            double* u = out + i * nout;
            auto excluded = [&](int64_t j) {
                if (!self_row) return j == i;
                for (int64_t k = self_row[i]; k < self_row[i + 1]; ++k)
                    if (self_val[k] == j) return true;
                return false;
            };
            for (int k = Li.nl_row[fi]; k < Li.nl_row[fi + 1]; ++k) {
                const Level& Lj = t.lv[Li.nl_level[k]];
                const int fj = Li.nl_leaf[k];
                for (int kk = Lj.idx_row[fj]; kk < Lj.idx_row[fj + 1]; ++kk) {
                    const int j = Lj.idx_val[kk];
                    if (!excluded(j)) kernel(mode, xi, t.pos + 3 * j, q + qd * j, u, 1.0);
                }
            }
            if (self_row) {  // _slist: self pairs that fell in the far field, fmm.hpp:398-408,798-807
                for (int64_t k = self_row[i]; k < self_row[i + 1]; ++k) {
                    const int64_t j = self_val[k];
                    if (!nlist_contains(t, static_cast<int>(i), static_cast<int>(j)))
                        kernel(mode, xi, t.pos + 3 * j, q + qd * j, u, -1.0);
                }
            }
        }
    }
    return 0;
}

}  // namespace

extern "C" {

void* oracle_tree_build(int64_t n, const double* pos, int max_level, int cap, int vertical_axis, double phideg)
{
    if (n <= 0 || !pos || max_level < 0 || max_level > 20) return nullptr;
    auto* t = new Tree(n, pos, vertical_axis, phideg);
    if (!(t->box.width > 0)) {  // zero vertical extent: the reference cannot build this box
        delete t;
        return nullptr;
    }
    Node root{0, 0, {}, {}};
    root.idx.resize(n);
    for (int64_t i = 0; i < n; ++i) root.idx[i] = static_cast<int>(i);
    refine(root, *t, max_level, cap);
    t->depth = max_depth(root);
    t->lv.assign(t->depth + 1, Level{});
    collect(root, t->lv);
    build_lists(*t);
    return t;
}

void oracle_tree_free(void* tree) { delete static_cast<Tree*>(tree); }

int oracle_tree_depth(const void* tree) { return static_cast<const Tree*>(tree)->depth; }

int oracle_is_neighbor(const void* tree, int64_t i, int64_t j)
{
    return nlist_contains(*static_cast<const Tree*>(tree), static_cast<int>(i), static_cast<int>(j)) ? 1 : 0;
}

void oracle_tree_box(const void* tree, double* out4)
{
    const Tree* t = static_cast<const Tree*>(tree);
    out4[0] = t->box.width;
    for (int d = 0; d < 3; ++d) out4[1 + d] = t->box.corner[d];
}

void oracle_tree_counts(const void* tree, int level, int64_t* out6)
{
    const Level& L = static_cast<const Tree*>(tree)->lv[level];
    out6[0] = static_cast<int64_t>(L.code.size());
    out6[1] = static_cast<int64_t>(L.leaf2node.size());
    out6[2] = static_cast<int64_t>(L.idx_val.size());
    out6[3] = static_cast<int64_t>(L.il_src.size());
    out6[4] = static_cast<int64_t>(L.cl_child.size());
    out6[5] = static_cast<int64_t>(L.nl_level.size());
}

void oracle_tree_copy(const void* tree, int level, int64_t* c_node, int64_t* leaf2node, int64_t* idx_row,
                      int64_t* idx_val, int64_t* ilist_row, int64_t* ilist_val, int64_t* clist_row,
                      int64_t* clist_val, int64_t* plist, int64_t* nlist_row, int64_t* nlist_val)
{
    const Level& L = static_cast<const Tree*>(tree)->lv[level];
    if (c_node) std::copy(L.code.begin(), L.code.end(), c_node);
    if (leaf2node) std::copy(L.leaf2node.begin(), L.leaf2node.end(), leaf2node);
    if (idx_row) std::copy(L.idx_row.begin(), L.idx_row.end(), idx_row);
    if (idx_val) std::copy(L.idx_val.begin(), L.idx_val.end(), idx_val);
    if (ilist_row) std::copy(L.il_row.begin(), L.il_row.end(), ilist_row);
    if (ilist_val)
        for (size_t k = 0; k < L.il_src.size(); ++k) {
            ilist_val[4 * k] = L.il_src[k];
            ilist_val[4 * k + 1] = L.il_d[k].i;
            ilist_val[4 * k + 2] = L.il_d[k].j;
            ilist_val[4 * k + 3] = L.il_d[k].k;
        }
    if (clist_row) std::copy(L.cl_row.begin(), L.cl_row.end(), clist_row);
    if (clist_val)
        for (size_t k = 0; k < L.cl_child.size(); ++k) {
            clist_val[2 * k] = L.cl_child[k];
            clist_val[2 * k + 1] = L.cl_code[k];
        }
    if (plist)
        for (size_t k = 0; k < L.pl_parent.size(); ++k) {
            plist[2 * k] = L.pl_parent[k];
            plist[2 * k + 1] = L.code[k];
        }
    if (nlist_row) std::copy(L.nl_row.begin(), L.nl_row.end(), nlist_row);
    if (nlist_val)
        for (size_t k = 0; k < L.nl_level.size(); ++k) {
            nlist_val[2 * k] = L.nl_level[k];
            nlist_val[2 * k + 1] = L.nl_leaf[k];
        }
}

int oracle_fmm(const void* tree, int p, int mode, const double* q, double* out, int which, int m2l_dense)
{
    if (!tree || !q || !out) return 1;
    return evaluate(*static_cast<const Tree*>(tree), p, mode, q, out, which, m2l_dense != 0, nullptr, nullptr);
}

int oracle_fmm_self(const void* tree, int p, int mode, const double* q, double* out, int which, int m2l_dense,
                    const int64_t* self_row, const int64_t* self_val)
{
    if (!tree || !q || !out || !self_row || !self_val) return 1;
    return evaluate(*static_cast<const Tree*>(tree), p, mode, q, out, which, m2l_dense != 0, self_row, self_val);
}

int oracle_direct(int64_t n, const double* pos, int mode, const double* q, int64_t ntarget, const int64_t* targets,
                  double* out)
{
    const int nc = ncomp_of(mode);
    if (nc < 0) return 1;
    const int qd = qdim_of(mode);
#pragma omp parallel for schedule(dynamic, 8)
    for (int64_t t = 0; t < ntarget; ++t) {
        const int64_t i = targets[t];
        double* u = out + t * nc;
        for (int64_t j = 0; j < n; ++j)
            if (j != i) kernel(mode, pos + 3 * i, pos + 3 * j, q + qd * j, u, 1.0);
    }
    return 0;
}

double oracle_Z(int n, int m, double theta) { return Zf(n, m, theta); }

double oracle_A(int n, int m) { return Af(n, m); }

double oracle_wigner_R(int n, int a, int b, double theta)
{
    const double sa = a < 0 ? neg1(a) : 1.0, sb = b < 0 ? neg1(b) : 1.0;
    return sa * sb * static_cast<double>(wigner_d_ld(n, a, b, -static_cast<long double>(theta)));
}

void oracle_harmonic(int p, double x, double y, double z, int grad, double* out)
{
    const double d[3] = {x, y, z};
    const Sph s = to_sph(d);
    const int K = (p + 1) * (p + 1);
    if (grad == 0) {
        std::vector<cd> R(K);
        regular(p, s, R.data());
        for (int a = 0; a < K; ++a) {
            out[2 * a] = R[a].real();
            out[2 * a + 1] = R[a].imag();
        }
    } else {
        std::vector<std::array<cd, 3>> G(K);
        regular_grad(p, s, G.data());
        for (int a = 0; a < K; ++a) {
            out[2 * a] = G[a][grad - 1].real();
            out[2 * a + 1] = G[a][grad - 1].imag();
        }
    }
}

}  // extern "C"
