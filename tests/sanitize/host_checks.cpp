// Host code of the plan (octree, Morton order, interaction and near lists,
// operator tables) exercised under AddressSanitizer and UndefinedBehavior-
// Sanitizer: degenerate and clustered point sets, one and several threads.
// Built with g++ only (no device code); exits non-zero on any failed check.
#include <cmath>
#include <cstdio>
#include <cstdlib>
#include <random>
#include <string>
#include <vector>

#include "../../simplefmm_amd/csrc/tables.hpp"
#include "../../simplefmm_amd/csrc/tree.hpp"

namespace {

int failures = 0;

void check(bool ok, const std::string& what)
{
    if (!ok) {
        std::printf("FAIL %s\n", what.c_str());
        ++failures;
    }
}

std::vector<double> points(const std::string& kind, int64_t n, unsigned seed)
{
    std::mt19937_64 rng(seed);
    std::uniform_real_distribution<double> u(0.0, 1.0);
    std::normal_distribution<double> g(0.0, 1.0);
    std::vector<double> p(3 * n);
    for (int64_t i = 0; i < n; ++i) {
        double x = u(rng), y = u(rng), z = u(rng);
        if (kind == "cluster") {  // a dense core and a sparse halo
            const double s = i % 10 == 0 ? 1.0 : 1e-3;
            x = 0.5 + s * g(rng);
            y = 0.5 + s * g(rng);
            z = 0.5 + s * g(rng);
        } else if (kind == "same") {
            x = y = z = 0.25;
        } else if (kind == "plane") {
            z = 0.5;
        } else if (kind == "line") {
            y = 0.3;
            z = 0.7;
        }
        p[3 * i] = x;
        p[3 * i + 1] = y;
        p[3 * i + 2] = z;
    }
    return p;
}

// explicit_box: the unit cube (the box rule of the reference gives a zero width
// when the points do not spread along the vertical axis; build_tree reports that)
void tree_case(const std::string& kind, int64_t n, int max_level, int cap, int threads, int axis,
               bool explicit_box = false)
{
    const std::string tag = kind + " n=" + std::to_string(n) + " cap=" + std::to_string(cap) +
                            " threads=" + std::to_string(threads) + " axis=" + std::to_string(axis) +
                            (explicit_box ? " unit box" : "");
    const std::vector<double> pos = points(kind, n, 17u + static_cast<unsigned>(n));
    const double corner[3] = {0.0, 0.0, 0.0};
    const sfmm::BoxGeom box =
        explicit_box ? sfmm::make_box_explicit(corner, 1.0, axis, 0.0) : sfmm::make_box(n, pos.data(), axis, 30.0);
    sfmm::Tree t;
    const std::string err = sfmm::build_tree(t, n, pos.data(), box, max_level, cap, threads);
    if (!err.empty()) {
        std::printf("tree %s: %s\n", tag.c_str(), err.c_str());
        return;
    }
    // the permutation is one, and the leaves of all levels cover every particle once
    std::vector<int> seen(n, 0);
    for (int32_t o : t.perm) {
        check(o >= 0 && o < n, tag + ": perm range");
        if (o >= 0 && o < n) ++seen[o];
    }
    for (int64_t i = 0; i < n; ++i) check(seen[i] == 1, tag + ": perm bijective");
    int64_t covered = 0;
    for (int l = 0; l <= t.depth; ++l) {
        const sfmm::TreeLevel& lv = t.lv[l];
        for (int32_t f = 0; f < lv.n_leaf(); ++f) {
            const int32_t a = lv.leaf2node[f];
            covered += lv.pend[a] - lv.pbegin[a];
        }
        sfmm::RefLists rl;
        sfmm::reference_lists(t, l, rl);
        check(rl.c_node.size() == static_cast<size_t>(lv.n_node()), tag + ": reference lists");
    }
    check(covered == n, tag + ": leaves cover the particles");
    for (size_t k = 0; k < t.near_begin.size(); ++k)
        check(t.near_begin[k] >= 0 && t.near_begin[k] <= t.near_end[k] && t.near_end[k] <= n, tag + ": near range");
    const sfmm::TreeStats st = sfmm::tree_stats(t);
    check(st.n_node >= 1 && st.n_leaf >= 1, tag + ": stats");
}

}  // namespace

int main()
{
    for (int threads : {1, 4}) {
        tree_case("uniform", 20000, 20, 64, threads, 2);
        tree_case("cluster", 20000, 20, 32, threads, 2);
        tree_case("uniform", 5000, 3, 8, threads, 0);  // max_level caps the depth
        tree_case("plane", 5000, 20, 16, threads, 1);
        tree_case("line", 3000, 20, 16, threads, 2);
        tree_case("same", 300, 20, 64, threads, 2);
        tree_case("uniform", 2, 20, 1, threads, 2);
        tree_case("uniform", 1, 20, 64, threads, 2);
        tree_case("line", 3000, 20, 16, threads, 2, true);
        tree_case("same", 300, 20, 64, threads, 2, true);  // coincident points: down to max_level
        tree_case("uniform", 1, 20, 64, threads, 2, true);
    }
    for (int p : {1, 4, 8, 12, 16}) {
        sfmm::Tables tb;
        sfmm::build_tables(p, tb);
        std::vector<double> rotA, coaxA;
        sfmm::build_m2l_fragments(tb, p % 2 == 0, rotA, coaxA);
        check(!rotA.empty() && !coaxA.empty(), "fragments p=" + std::to_string(p));
        if (p <= 12) {
            std::vector<uint16_t> ops;
            std::vector<float> az;
            std::vector<int32_t> grp;
            int e = 0;
            sfmm::build_dense_m2l(tb, ops, e, az, grp);
            check(!ops.empty() && !az.empty(), "dense operators p=" + std::to_string(p));
        }
    }
    std::printf("%s (%d failed checks)\n", failures ? "FAILED" : "ok", failures);
    return failures ? 1 : 0;
}
