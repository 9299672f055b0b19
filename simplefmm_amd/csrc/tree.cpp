// Host-side adaptive octree for the FMM plan.  See tree.hpp for the semantics
// kept from the reference partitioner and the construction used here.
#include "tree.hpp"

#include "morton.hpp"

#include <algorithm>
#include <cmath>
#include <functional>
#include <thread>

namespace sfmm {

namespace {

constexpr double kPi = 3.14159265358979323846;

// Runs f(begin, end, chunk) over [0, n) split into contiguous chunks on `threads` threads.
void parallel_chunks(int64_t n, int threads, const std::function<void(int64_t, int64_t, int)>& f)
{
    if (threads <= 1 || n < 4096) {
        f(0, n, 0);
        return;
    }
    std::vector<std::thread> pool;
    for (int c = 0; c < threads; ++c) {
        const int64_t a = n * c / threads, b = n * (c + 1) / threads;
        pool.emplace_back([&, a, b, c] { f(a, b, c); });
    }
    for (auto& th : pool) th.join();
}

// Stable LSD radix sort of (key, value) by key, `bits` low bits significant.
void radix_sort(std::vector<int64_t>& key, std::vector<int32_t>& val, int bits)
{
    const size_t n = key.size();
    std::vector<int64_t> k2(n);
    std::vector<int32_t> v2(n);
    constexpr int R = 11;
    for (int shift = 0; shift < bits; shift += R) {
        std::vector<size_t> cnt((1 << R) + 1, 0);
        for (size_t i = 0; i < n; ++i) ++cnt[((key[i] >> shift) & ((1 << R) - 1)) + 1];
        for (int b = 0; b < (1 << R); ++b) cnt[b + 1] += cnt[b];
        for (size_t i = 0; i < n; ++i) {
            const size_t d = cnt[(key[i] >> shift) & ((1 << R) - 1)]++;
            k2[d] = key[i];
            v2[d] = val[i];
        }
        key.swap(k2);
        val.swap(v2);
    }
}

}  // namespace

int32_t morton_axis(int64_t code, int axis) { return code_axis(code, axis); }

int64_t morton_code(int32_t i, int32_t j, int32_t k)
{
    return spread3(static_cast<uint32_t>(i)) | (spread3(static_cast<uint32_t>(j)) << 1) |
           (spread3(static_cast<uint32_t>(k)) << 2);
}

void BoxGeom::rotate(const double* v, double* r) const  // box.hpp:152-157
{
    rotate_xy(cphi, sphi, v[X], v[Y], r[0], r[1]);
    r[2] = v[Z];
}

void BoxGeom::centre(int level, int64_t code, double* out) const  // box.hpp:66-70,147-150
{
    for (int d = 0; d < 3; ++d) out[d] = corner[d] + width * ((morton_axis(code, d) + 0.5) / (1 << level));
}

static void init_axes(BoxGeom& b, int vertical_axis, double phideg)
{
    b.X = (vertical_axis + 1) % 3;
    b.Y = (vertical_axis + 2) % 3;
    b.Z = vertical_axis;
    b.phi = phideg * (kPi / 180);
    b.cphi = std::cos(b.phi);
    b.sphi = std::sin(b.phi);
}

BoxGeom make_box(int64_t n, const double* pos, int vertical_axis, double phideg)
{
    BoxGeom b;
    init_axes(b, vertical_axis, phideg);
    double lo[3] = {0, 0, 0}, hi[3] = {0, 0, 0};
    for (int64_t i = 0; i < n; ++i) {
        double r[3];
        b.rotate(pos + 3 * i, r);
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
    // box.hpp:101-106: the cube starts from the vertical extent with a 1e-3
    // margin on each side and doubles upward until it covers the widest extent.
    b.width = (1 + 2 * 1e-3) * dl[2];
    // With no vertical extent the reference's doubling never ends (or leaves a
    // zero-width box); the zero width is kept and rejected by build_tree.
    while (b.width > 0 && b.width < dmax) {
        c[2] += b.width / 2;
        b.width *= 2;
    }
    for (int d = 0; d < 3; ++d) b.corner[d] = c[d] - b.width / 2;
    return b;
}

BoxGeom make_box_explicit(const double* corner, double width, int vertical_axis, double phideg)
{
    BoxGeom b;
    init_axes(b, vertical_axis, phideg);
    b.width = width;
    for (int d = 0; d < 3; ++d) b.corner[d] = corner[d];
    return b;
}

int32_t TreeLevel::find(int64_t c) const
{
    auto it = std::lower_bound(code.begin(), code.end(), c);
    return (it != code.end() && *it == c) ? static_cast<int32_t>(it - code.begin()) : -1;
}

std::string build_tree(Tree& t, int64_t n, const double* pos, const BoxGeom& box, int max_level, int cap,
                       int threads)
{
    if (n <= 0) return "particle count must be positive";
    if (n > 0x7fffffffLL) return "particle count exceeds 2^31-1";
    if (max_level < 0 || max_level > 20) return "max_level must be in [0, 20] (20-bit Morton fields)";
    if (cap < 0) return "max_particles_per_node must be non-negative";
    if (!(box.width > 0) || !std::isfinite(box.width)) return "box width must be positive and finite";
    threads = std::max(1, threads);
    t = Tree{};
    t.n = n;
    t.max_level = max_level;
    t.cap = cap;
    t.box = box;
    const int L = max_level;

    // 1. cell codes at level L (box.hpp:57-63): clamp(int(u * 2^L)); the code of a
    //    coarser level l is this code shifted right by 3(L-l).
    std::vector<int64_t> key(n);
    t.perm.resize(n);
    std::vector<char> bad(threads, 0);
    parallel_chunks(n, threads, [&](int64_t a, int64_t b, int c) {
        for (int64_t i = a; i < b; ++i) {
            double r[3];
            box.rotate(pos + 3 * i, r);
            int32_t iv[3];
            for (int d = 0; d < 3; ++d)
                if (!cell_index(r[d], box.corner[d], box.width, L, iv[d])) bad[c] = 1;
            key[i] = morton_code(iv[0], iv[1], iv[2]);
            t.perm[i] = static_cast<int32_t>(i);
        }
    });
    for (char c : bad)
        if (c) return "a particle lies outside the box (or is not finite)";

    // 2. one stable sort by code = sort by (code, index)
    radix_sort(key, t.perm, 3 * L);

    // 3. levels top-down: a node subdivides while l < max_level and count > cap
    //    (partition.hpp:92-103); children are the non-empty octants in code order.
    t.lv.emplace_back();
    {
        TreeLevel& r = t.lv[0];
        r.code = {0};
        r.pbegin = {0};
        r.pend = {n};
        r.parent = {-1};
    }
    for (int l = 0;; ++l) {
        TreeLevel& cur = t.lv[l];
        const int32_t nn = cur.n_node();
        cur.child0.assign(nn, -1);
        cur.nchild.assign(nn, 0);
        TreeLevel next;
        if (l < L) {
            const int sh = 3 * (L - l - 1);
            for (int32_t a = 0; a < nn; ++a) {
                const int64_t b0 = cur.pbegin[a], b1 = cur.pend[a];
                if (b1 - b0 <= cap) continue;
                cur.child0[a] = next.n_node();
                int64_t s = b0;
                while (s < b1) {
                    const int64_t prefix = key[s] >> sh;
                    // end of this octant's run
                    const int64_t e = std::upper_bound(key.begin() + s, key.begin() + b1, ((prefix + 1) << sh) - 1) -
                                      key.begin();
                    next.code.push_back(prefix);
                    next.pbegin.push_back(s);
                    next.pend.push_back(e);
                    next.parent.push_back(a);
                    ++cur.nchild[a];
                    s = e;
                }
            }
        }
        if (next.code.empty()) break;
        t.lv.push_back(std::move(next));
    }
    t.depth = static_cast<int>(t.lv.size()) - 1;
    key.clear();
    key.shrink_to_fit();

    // 4. leaves; particles of a leaf ascending by index (partition.hpp:52,
    //    box.hpp:142: the per-level sort of (code, index) pairs)
    t.leaf_row.assign(t.depth + 2, 0);
    for (int l = 0; l <= t.depth; ++l) {
        TreeLevel& lv = t.lv[l];
        lv.node2leaf.assign(lv.n_node(), -1);
        for (int32_t a = 0; a < lv.n_node(); ++a) {
            if (lv.nchild[a] == 0) {
                lv.node2leaf[a] = lv.n_leaf();
                lv.leaf2node.push_back(a);
                t.leaf_level.push_back(l);
                t.leaf_node.push_back(a);
            }
        }
        t.leaf_row[l + 1] = t.leaf_row[l] + lv.n_leaf();
    }
    parallel_chunks(static_cast<int64_t>(t.leaf_node.size()), threads, [&](int64_t a, int64_t b, int) {
        for (int64_t g = a; g < b; ++g) {
            const TreeLevel& lv = t.lv[t.leaf_level[g]];
            const int32_t node = t.leaf_node[g];
            std::sort(t.perm.begin() + lv.pbegin[node], t.perm.begin() + lv.pend[node]);
        }
    });

    // 5. same-level neighbours and interaction lists, top-down.  The neighbours of
    //    X are the children of the neighbours of parent(X) within +-1 of X; the other
    //    children of the parent's neighbours form X's interaction list
    //    (tree.hpp:122-174, partition.hpp:283-297).
    t.lv[0].nb_row = {0, 1};
    t.lv[0].nb = {0};
    t.lv[0].il_row = {0};
    for (int l = 1; l <= t.depth; ++l) {
        TreeLevel& lv = t.lv[l];
        const TreeLevel& up = t.lv[l - 1];
        const int32_t nn = lv.n_node();
        std::vector<int32_t> ijk(3 * static_cast<size_t>(nn));
        for (int32_t a = 0; a < nn; ++a)
            for (int d = 0; d < 3; ++d) ijk[3 * a + d] = morton_axis(lv.code[a], d);
        const int nth = nn > 20000 ? threads : 1;
        std::vector<std::vector<int32_t>> nbv(nth), srcv(nth);
        std::vector<std::vector<int16_t>> offv(nth);
        std::vector<std::vector<int64_t>> nbc(nth), ilc(nth);
        std::vector<int64_t> cuts(nth + 1);
        for (int c = 0; c <= nth; ++c) cuts[c] = static_cast<int64_t>(nn) * c / nth;
        auto work = [&](int c) {
            for (int64_t a = cuts[c]; a < cuts[c + 1]; ++a) {
                const int32_t P = lv.parent[a];
                int64_t cnt_nb = 0, cnt_il = 0;
                for (int64_t k = up.nb_row[P]; k < up.nb_row[P + 1]; ++k) {
                    const int32_t Y = up.nb[k];
                    for (int32_t z = up.child0[Y]; z >= 0 && z < up.child0[Y] + up.nchild[Y]; ++z) {
                        const int di = ijk[3 * z] - ijk[3 * a], dj = ijk[3 * z + 1] - ijk[3 * a + 1],
                                  dk = ijk[3 * z + 2] - ijk[3 * a + 2];
                        if (std::abs(di) <= 1 && std::abs(dj) <= 1 && std::abs(dk) <= 1) {
                            nbv[c].push_back(z);
                            ++cnt_nb;
                        } else if (l >= 2) {
                            srcv[c].push_back(z);
                            offv[c].push_back(static_cast<int16_t>(offset_id(di, dj, dk)));
                            ++cnt_il;
                        }
                    }
                }
                nbc[c].push_back(cnt_nb);
                ilc[c].push_back(cnt_il);
            }
        };
        if (nth == 1) {
            work(0);
        } else {
            std::vector<std::thread> pool;
            for (int c = 0; c < nth; ++c) pool.emplace_back(work, c);
            for (auto& th : pool) th.join();
        }
        lv.nb_row.assign(nn + 1, 0);
        lv.il_row.assign(nn + 1, 0);
        int64_t a = 0;
        for (int c = 0; c < nth; ++c)
            for (size_t k = 0; k < nbc[c].size(); ++k, ++a) {
                lv.nb_row[a + 1] = lv.nb_row[a] + nbc[c][k];
                lv.il_row[a + 1] = lv.il_row[a] + ilc[c][k];
            }
        lv.nb.reserve(lv.nb_row[nn]);
        lv.il_src.reserve(lv.il_row[nn]);
        lv.il_off.reserve(lv.il_row[nn]);
        for (int c = 0; c < nth; ++c) {
            lv.nb.insert(lv.nb.end(), nbv[c].begin(), nbv[c].end());
            lv.il_src.insert(lv.il_src.end(), srcv[c].begin(), srcv[c].end());
            lv.il_off.insert(lv.il_off.end(), offv[c].begin(), offv[c].end());
        }
    }

    // 6. near-field sources of each leaf as particle ranges: the coarser leaves
    //    adjacent to an ancestor and the whole subtrees of the same-level
    //    neighbours (partition.hpp:324-366).  Subtrees are contiguous in the
    //    sorted order, so each is one range; adjacent ranges are merged.
    const int64_t nleaves = static_cast<int64_t>(t.leaf_node.size());
    std::vector<std::vector<int64_t>> rb(threads), re(threads), rc(threads);
    std::vector<int64_t> lcuts(threads + 1);
    for (int c = 0; c <= threads; ++c) lcuts[c] = nleaves * c / threads;
    auto near_work = [&](int c) {
        std::vector<std::pair<int64_t, int64_t>> rg;
        for (int64_t g = lcuts[c]; g < lcuts[c + 1]; ++g) {
            rg.clear();
            const int l = t.leaf_level[g];
            int32_t a = t.leaf_node[g];
            const TreeLevel& me = t.lv[l];
            for (int64_t k = me.nb_row[a]; k < me.nb_row[a + 1]; ++k) {
                const int32_t Y = me.nb[k];
                rg.push_back({me.pbegin[Y], me.pend[Y]});
            }
            int32_t anc = a;
            for (int la = l - 1; la >= 0; --la) {
                anc = t.lv[la + 1].parent[anc];
                const TreeLevel& A = t.lv[la];
                for (int64_t k = A.nb_row[anc]; k < A.nb_row[anc + 1]; ++k) {
                    const int32_t Y = A.nb[k];
                    if (A.nchild[Y] == 0) rg.push_back({A.pbegin[Y], A.pend[Y]});
                }
            }
            std::sort(rg.begin(), rg.end());
            int64_t cnt = 0;
            for (size_t k = 0; k < rg.size(); ++k) {
                if (cnt > 0 && re[c].back() == rg[k].first) {
                    re[c].back() = rg[k].second;
                } else {
                    rb[c].push_back(rg[k].first);
                    re[c].push_back(rg[k].second);
                    ++cnt;
                }
            }
            rc[c].push_back(cnt);
        }
    };
    {
        std::vector<std::thread> pool;
        for (int c = 0; c < threads; ++c) pool.emplace_back(near_work, c);
        for (auto& th : pool) th.join();
    }
    t.near_row.assign(nleaves + 1, 0);
    int64_t g = 0;
    for (int c = 0; c < threads; ++c)
        for (int64_t v : rc[c]) {
            t.near_row[g + 1] = t.near_row[g] + v;
            ++g;
        }
    for (int c = 0; c < threads; ++c) {
        t.near_begin.insert(t.near_begin.end(), rb[c].begin(), rb[c].end());
        t.near_end.insert(t.near_end.end(), re[c].begin(), re[c].end());
    }
    return {};
}

void reference_lists(const Tree& t, int l, RefLists& o)
{
    o = RefLists{};
    const TreeLevel& lv = t.lv[l];
    o.c_node.assign(lv.code.begin(), lv.code.end());
    o.idx_row.push_back(0);
    for (int32_t f = 0; f < lv.n_leaf(); ++f) {
        const int32_t a = lv.leaf2node[f];
        o.leaf2node.push_back(a);
        for (int64_t s = lv.pbegin[a]; s < lv.pend[a]; ++s) o.idx_val.push_back(t.perm[s]);
        o.idx_row.push_back(static_cast<int64_t>(o.idx_val.size()));
    }
    o.ilist_row.push_back(0);
    if (l >= 2) {
        for (int32_t a = 0; a < lv.n_node(); ++a) {
            for (int64_t k = lv.il_row[a]; k < lv.il_row[a + 1]; ++k) {
                const int off = lv.il_off[k];
                o.ilist_val.insert(o.ilist_val.end(),
                                   {lv.il_src[k], off % 7 - 3, (off / 7) % 7 - 3, off / 49 - 3});
            }
            o.ilist_row.push_back(static_cast<int64_t>(o.ilist_val.size() / 4));
        }
    }
    o.clist_row.push_back(0);
    if (l < t.depth) {
        const TreeLevel& dn = t.lv[l + 1];
        for (int32_t a = 0; a < lv.n_node(); ++a) {
            for (int32_t z = lv.child0[a]; z >= 0 && z < lv.child0[a] + lv.nchild[a]; ++z)
                o.clist_val.insert(o.clist_val.end(), {z, dn.code[z]});
            o.clist_row.push_back(static_cast<int64_t>(o.clist_val.size() / 2));
        }
    }
    if (l >= 1)
        for (int32_t a = 0; a < lv.n_node(); ++a) o.plist.insert(o.plist.end(), {lv.parent[a], lv.code[a]});
    // nlist (partition.hpp:343-366): rows sorted by (level, leaf)
    o.nlist_row.push_back(0);
    for (int32_t f = 0; f < lv.n_leaf(); ++f) {
        const int32_t a = lv.leaf2node[f];
        std::vector<int32_t> anc(l + 1);
        anc[l] = a;
        for (int la = l - 1; la >= 0; --la) anc[la] = t.lv[la + 1].parent[anc[la + 1]];
        for (int la = 0; la < l; ++la) {
            const TreeLevel& A = t.lv[la];
            for (int64_t k = A.nb_row[anc[la]]; k < A.nb_row[anc[la] + 1]; ++k) {
                const int32_t Y = A.nb[k];
                if (A.node2leaf[Y] >= 0) o.nlist_val.insert(o.nlist_val.end(), {la, A.node2leaf[Y]});
            }
        }
        for (int ld = l; ld <= t.depth; ++ld) {
            const TreeLevel& D = t.lv[ld];
            const int sh = 3 * (ld - l);
            for (int64_t k = lv.nb_row[a]; k < lv.nb_row[a + 1]; ++k) {
                const int64_t c = lv.code[lv.nb[k]];
                const auto lo = std::lower_bound(D.code.begin(), D.code.end(), c << sh) - D.code.begin();
                const auto hi = std::lower_bound(D.code.begin(), D.code.end(), (c + 1) << sh) - D.code.begin();
                for (auto z = lo; z < hi; ++z)
                    if (D.node2leaf[z] >= 0) o.nlist_val.insert(o.nlist_val.end(), {ld, D.node2leaf[z]});
            }
        }
        o.nlist_row.push_back(static_cast<int64_t>(o.nlist_val.size() / 2));
    }
}

TreeStats tree_stats(const Tree& t)
{
    TreeStats s;
    for (int l = 0; l <= t.depth; ++l) {
        const TreeLevel& lv = t.lv[l];
        s.n_node += lv.n_node();
        s.n_leaf += lv.n_leaf();
        if (l >= 2) s.m2l_pairs += lv.il_row.back();
        if (l >= 3) s.m2m_edges += lv.n_node();
    }
    for (size_t g = 0; g < t.leaf_node.size(); ++g) {
        const TreeLevel& lv = t.lv[t.leaf_level[g]];
        const int64_t cnt = lv.pend[t.leaf_node[g]] - lv.pbegin[t.leaf_node[g]];
        int64_t src = 0;
        for (int64_t k = t.near_row[g]; k < t.near_row[g + 1]; ++k) src += t.near_end[k] - t.near_begin[k];
        s.near_pairs += cnt * src - cnt;
        if (t.leaf_level[g] >= 2) s.far_leaf_particles += cnt;
    }
    return s;
}

}  // namespace sfmm
