// Driver for the reference FMM (icyspirit/SimpleFMM), used only to write golden
// fixtures and to time the reference CPU path.  It contains no reference code:
// the reference headers are included from the reference checkout at build time
// (oracle/Makefile passes -I<reference dir>), and the binary is written to
// oracle/_ref/ only.  Nothing in the product links or loads it.
//
// Call sequence mirrors main.cpp:38,46-47,54,61,66 (Box -> OctreePartitioner ->
// refine -> FMM3D<double,3,p,Vector,true> -> rinv / rinv<true>).
//
// usage: ref_harness IN OUT P CAP MAXLEVEL MODE [TREEOUT] [REPEAT] [SELFLIST]
//   MODE = potgrad : IN holds scalar charges q; writes phi and grad(phi) using
//                    the curl identity (two rinv<true> runs, Q=(q,0,0), (0,q,0))
//          vector  : IN holds 3-vector charges; writes rinv (3 potentials) and
//                    rinv<true> (curl field), as main.cpp does
//          potential: IN holds scalar charges; writes phi only, from an FMM3D
//                    built without gradient tables (support_gradient = false,
//                    fmm.hpp:134,151-153), which fits larger N
//          time    : like potgrad but prints wall times only (ctor, rinv, rinv<true>)
//   Every evaluation is written three times: full (rinv), far (rinv_nonear) and
//   near (N2N), unless MODE ends in "_full".
//   TREEOUT (or "-") receives the per-level octree lists.
//   SELFLIST (optional): 'SFMS', int64 n, int64 nnz, int64 row[n+1], int64 val[nnz];
//   FMM3D is then built with self_generator(i) = val[row[i]..row[i+1]) (fmm.hpp:204)
//   and evaluated with is_self(i, j) = j in that list (rinv(Q, U, is_self),
//   fmm.hpp:864-873); without it is_self(i, j) = (i == j) (fmm.hpp:875-881).
#include "box.hpp"
#include "fmm.hpp"
#include "matrix.hpp"

#include <atomic>
#include <chrono>
#include <cstdint>
#include <cstdio>
#include <cstring>
#include <memory>
#include <string>
#include <vector>

using V3 = Vector<double, 3>;

// MPICH 3.3.2 (the MPI in this image) rejects MPI_Win_sync outside a
// passive-target epoch and aborts.  The reference calls it (distributed.hpp:216)
// only to order plain stores into an MPI-3 shared-memory window before a
// barrier; in MPI's unified memory model that is a processor memory barrier.
// The harness provides exactly that through the MPI profiling interface, so the
// reference runs unmodified.
extern "C" int MPI_Win_sync(MPI_Win)
{
    std::atomic_thread_fence(std::memory_order_seq_cst);
    return MPI_SUCCESS;
}

namespace {

struct Input {
    int64_t n = 0;
    int32_t qdim = 0;
    std::vector<V3> pos;
    std::vector<double> q;  // qdim*n
};

bool read_input(const char* path, Input& in)
{
    FILE* f = std::fopen(path, "rb");
    if (!f) return false;
    char magic[4];
    int32_t version = 0;
    bool ok = std::fread(magic, 1, 4, f) == 4 && std::memcmp(magic, "SFMI", 4) == 0 &&
              std::fread(&version, 4, 1, f) == 1 && std::fread(&in.n, 8, 1, f) == 1 &&
              std::fread(&in.qdim, 4, 1, f) == 1;
    if (ok) {
        std::vector<double> buf(3 * in.n);
        ok = std::fread(buf.data(), 8, buf.size(), f) == buf.size();
        in.pos.resize(in.n);
        for (int64_t i = 0; i < in.n; ++i) in.pos[i] = V3{{buf[3 * i], buf[3 * i + 1], buf[3 * i + 2]}};
        in.q.resize(in.qdim * in.n);
        ok = ok && std::fread(in.q.data(), 8, in.q.size(), f) == in.q.size();
    }
    std::fclose(f);
    return ok;
}

struct SelfList {
    std::vector<int64_t> row, val;
    bool given() const { return !row.empty(); }
    std::vector<int> of(int i) const { return std::vector<int>(val.begin() + row[i], val.begin() + row[i + 1]); }
    bool has(int i, int j) const
    {
        for (int64_t k = row[i]; k < row[i + 1]; ++k)
            if (val[k] == j) return true;
        return false;
    }
};

bool read_self_list(const char* path, int64_t n, SelfList& s)
{
    FILE* f = std::fopen(path, "rb");
    if (!f) return false;
    char magic[4];
    int64_t nn = 0, nnz = 0;
    bool ok = std::fread(magic, 1, 4, f) == 4 && std::memcmp(magic, "SFMS", 4) == 0 &&
              std::fread(&nn, 8, 1, f) == 1 && std::fread(&nnz, 8, 1, f) == 1 && nn == n && nnz >= 0;
    if (ok) {
        s.row.resize(n + 1);
        s.val.resize(nnz);
        ok = std::fread(s.row.data(), 8, n + 1, f) == static_cast<size_t>(n + 1) &&
             std::fread(s.val.data(), 8, nnz, f) == static_cast<size_t>(nnz) && s.row[0] == 0 && s.row[n] == nnz;
    }
    std::fclose(f);
    return ok;
}

void put_block(FILE* f, const char* name, const double* data, int64_t count)
{
    char tag[16] = {0};
    std::strncpy(tag, name, 15);
    std::fwrite(tag, 1, 16, f);
    std::fwrite(&count, 8, 1, f);
    std::fwrite(data, 8, count, f);
}

void put_iblock(FILE* f, const char* name, const std::vector<int64_t>& v)
{
    char tag[16] = {0};
    std::strncpy(tag, name, 15);
    std::fwrite(tag, 1, 16, f);
    int64_t count = static_cast<int64_t>(v.size());
    std::fwrite(&count, 8, 1, f);
    std::fwrite(v.data(), 8, count, f);
}

double now()
{
    return std::chrono::duration<double>(std::chrono::steady_clock::now().time_since_epoch()).count();
}

template <class Part>
void dump_tree(const char* path, const Part& part)
{
    FILE* f = std::fopen(path, "wb");
    if (!f) return;
    std::fwrite("SFMT", 1, 4, f);
    int64_t depth = part.level();
    std::fwrite(&depth, 8, 1, f);
    for (int l = 0; l <= part.level(); ++l) {
        const auto& lv = part.octreeLevel(l);
        std::vector<int64_t> head{l, lv.n_node(), lv.n_leaf()};
        put_iblock(f, "head", head);
        put_iblock(f, "c_node", std::vector<int64_t>(lv.c_node().begin(), lv.c_node().end()));
        std::vector<int64_t> leaf2node, irow{0}, ival;
        for (int i = 0; i < lv.n_leaf(); ++i) {
            leaf2node.push_back(lv.from_leaf(i));
            for (int k = 0; k < lv.indices().nnz(i); ++k) ival.push_back(std::get<0>(lv.indices().value(i, k)));
            irow.push_back(static_cast<int64_t>(ival.size()));
        }
        put_iblock(f, "leaf2node", leaf2node);
        put_iblock(f, "idx_row", irow);
        put_iblock(f, "idx_val", ival);
        std::vector<int64_t> row{0}, val;
        if (l >= 2) {
            for (int i = 0; i < lv.n_node(); ++i) {
                for (int k = 0; k < lv.ilist().nnz(i); ++k) {
                    const auto& e = lv.ilist().value(i, k);
                    const auto& d = std::get<1>(e);
                    val.push_back(std::get<0>(e));
                    val.push_back(d.i);
                    val.push_back(d.j);
                    val.push_back(d.k);
                }
                row.push_back(static_cast<int64_t>(val.size() / 4));
            }
        }
        put_iblock(f, "ilist_row", row);
        put_iblock(f, "ilist_val", val);
        row.assign(1, 0);
        val.clear();
        if (l < part.level()) {
            for (int i = 0; i < lv.n_node(); ++i) {
                for (int k = 0; k < lv.clist().nnz(i); ++k) {
                    const auto& e = lv.clist().value(i, k);
                    val.push_back(std::get<0>(e));
                    val.push_back(std::get<1>(e));
                }
                row.push_back(static_cast<int64_t>(val.size() / 2));
            }
        }
        put_iblock(f, "clist_row", row);
        put_iblock(f, "clist_val", val);
        val.clear();
        for (const auto& e : lv.plist()) {
            val.push_back(e.first);
            val.push_back(e.second);
        }
        put_iblock(f, "plist", val);
        row.assign(1, 0);
        val.clear();
        for (int i = 0; i < lv.n_leaf(); ++i) {
            for (int k = 0; k < lv.nlist().nnz(i); ++k) {
                const auto& e = lv.nlist().value(i, k);
                val.push_back(std::get<0>(e));
                val.push_back(std::get<1>(e));
            }
            row.push_back(static_cast<int64_t>(val.size() / 2));
        }
        put_iblock(f, "nlist_row", row);
        put_iblock(f, "nlist_val", val);
    }
    std::fclose(f);
}

template <int P, bool G>
int run(const Input& in, const char* out_path, int cap, int max_level, const std::string& mode,
        const char* tree_path, int repeat, const SelfList& self)
{
    const int64_t n = in.n;
    const bool full_only = mode.size() > 5 && mode.substr(mode.size() - 5) == "_full";
    const std::string base = full_only ? mode.substr(0, mode.size() - 5) : mode;

    double t0 = now();
    Box<double, 3, Vector> box(in.pos, 2);
    OctreePartitioner<double, Vector> part(in.pos, box);
    part.refine(max_level, cap);
    double t1 = now();
    if (tree_path && std::strcmp(tree_path, "-") != 0) dump_tree(tree_path, part);
    MPI_Barrier(MPI_COMM_WORLD);
    double t2 = now();
    using Fmm = FMM3D<double, 3, P, Vector, G>;
    std::unique_ptr<Fmm> fmm_ptr =
        self.given() ? std::make_unique<Fmm>(part, MPI_COMM_WORLD, [&self](int i) { return self.of(i); })
                     : std::make_unique<Fmm>(part, MPI_COMM_WORLD);
    const Fmm& fmm = *fmm_ptr;
    MPI_Barrier(MPI_COMM_WORLD);
    double t3 = now();

    int rank = 0;
    MPI_Comm_rank(MPI_COMM_WORLD, &rank);
    const auto is_self = [&self](int i, int j) { return self.given() ? self.has(i, j) : i == j; };

    auto allreduce = [&](std::vector<V3>& U) {
        MPI_Allreduce(MPI_IN_PLACE, U.data(), 3 * static_cast<int>(n), MPI_DOUBLE, MPI_SUM, MPI_COMM_WORLD);
    };
    // which: 0 full, 1 far, 2 near
    auto eval = [&](const std::vector<V3>& Q, bool grad, int which) {
        std::vector<V3> U(n);
        for (auto& u : U) u = V3{{0.0, 0.0, 0.0}};
        if constexpr (G) {
            if (which == 0) {
                if (grad) fmm.template rinv<true>(Q.data(), U.data(), is_self);
                else fmm.template rinv<false>(Q.data(), U.data(), is_self);
            } else if (which == 1) {
                if (grad) fmm.template rinv_nonear<true>(Q.data(), U.data());
                else fmm.template rinv_nonear<false>(Q.data(), U.data());
            } else {
                if (grad) fmm.template N2N<true>(Q.data(), U.data(), is_self);
                else fmm.template N2N<false>(Q.data(), U.data(), is_self);
            }
        } else {
            (void)grad;
            if (which == 0) fmm.template rinv<false>(Q.data(), U.data(), is_self);
            else if (which == 1) fmm.template rinv_nonear<false>(Q.data(), U.data());
            else fmm.template N2N<false>(Q.data(), U.data(), is_self);
        }
        allreduce(U);
        return U;
    };

    if (base == "time" && G) {
        std::vector<V3> Q1(n), Q2(n);
        for (int64_t i = 0; i < n; ++i) {
            Q1[i] = V3{{in.q[i], 0.0, 0.0}};
            Q2[i] = V3{{0.0, in.q[i], 0.0}};
        }
        for (int r = 0; r < repeat; ++r) {
            MPI_Barrier(MPI_COMM_WORLD);
            double a = now();
            auto U = eval(Q1, false, 0);
            MPI_Barrier(MPI_COMM_WORLD);
            double b = now();
            auto F1 = eval(Q1, true, 0);
            MPI_Barrier(MPI_COMM_WORLD);
            double c = now();
            auto F2 = eval(Q2, true, 0);
            MPI_Barrier(MPI_COMM_WORLD);
            double d = now();
            if (rank == 0) {
                std::printf("{\"n\": %lld, \"p\": %d, \"tree_s\": %.6f, \"ctor_s\": %.6f, \"rinv_s\": %.6f, "
                            "\"rinv_curl_s\": %.6f, \"rinv_curl2_s\": %.6f, \"depth\": %d}\n",
                            static_cast<long long>(n), P, t1 - t0, t3 - t2, b - a, c - b, d - c, part.level());
                std::fflush(stdout);
            }
        }
        return 0;
    }

    if (rank != 0) {
        // every rank must take part in the collective evaluations
    }
    FILE* f = rank == 0 ? std::fopen(out_path, "wb") : nullptr;
    if (rank == 0 && !f) return 3;
    if (f) {
        std::fwrite("SFMO", 1, 4, f);
        int64_t hdr[4] = {n, P, part.level(), 0};
        std::fwrite(hdr, 8, 4, f);
    }
    const int nwhich = full_only ? 1 : 3;
    const char* suffix[3] = {"full", "far", "near"};
    if (base == "potential") {
        std::vector<V3> Q1(n);
        for (int64_t i = 0; i < n; ++i) Q1[i] = V3{{in.q[i], 0.0, 0.0}};
        for (int w = 0; w < nwhich; ++w) {
            auto U = eval(Q1, false, w);
            if (f) {
                std::vector<double> phi(n);
                for (int64_t i = 0; i < n; ++i) phi[i] = U[i][0];
                put_block(f, (std::string("phi_") + suffix[w]).c_str(), phi.data(), n);
            }
        }
    } else if (base == "potgrad" && G) {
        std::vector<V3> Q1(n), Q2(n);
        for (int64_t i = 0; i < n; ++i) {
            Q1[i] = V3{{in.q[i], 0.0, 0.0}};
            Q2[i] = V3{{0.0, in.q[i], 0.0}};
        }
        for (int w = 0; w < nwhich; ++w) {
            auto U = eval(Q1, false, w);
            auto F1 = eval(Q1, true, w);
            auto F2 = eval(Q2, true, w);
            if (f) {
                std::vector<double> phi(n), grad(3 * n), redundant(n);
                for (int64_t i = 0; i < n; ++i) {
                    phi[i] = U[i][0];
                    grad[3 * i + 0] = F2[i][2];
                    grad[3 * i + 1] = -F1[i][2];
                    grad[3 * i + 2] = F1[i][1];
                    // F2_x = -dphi/dz duplicates grad_z: kept as a consistency check
                    redundant[i] = -F2[i][0];
                }
                put_block(f, (std::string("phi_") + suffix[w]).c_str(), phi.data(), n);
                put_block(f, (std::string("grad_") + suffix[w]).c_str(), grad.data(), 3 * n);
                put_block(f, (std::string("gzdup_") + suffix[w]).c_str(), redundant.data(), n);
            }
        }
    } else if (base == "vector" && G) {
        std::vector<V3> Q(n);
        for (int64_t i = 0; i < n; ++i) Q[i] = V3{{in.q[3 * i], in.q[3 * i + 1], in.q[3 * i + 2]}};
        for (int w = 0; w < nwhich; ++w) {
            auto U = eval(Q, false, w);
            auto F = eval(Q, true, w);
            if (f) {
                put_block(f, (std::string("pot_") + suffix[w]).c_str(), &U[0][0], 3 * n);
                put_block(f, (std::string("curl_") + suffix[w]).c_str(), &F[0][0], 3 * n);
            }
        }
    } else {
        if (f) std::fclose(f);
        return 4;
    }
    if (f) std::fclose(f);
    return 0;
}

}  // namespace

int main(int argc, char** argv)
{
    MPI_Init(&argc, &argv);
    int rc = 0;
    if (argc < 7) {
        std::fprintf(stderr, "usage: %s IN OUT P CAP MAXLEVEL MODE [TREEOUT] [REPEAT]\n", argv[0]);
        rc = 2;
    } else {
        Input in;
        if (!read_input(argv[1], in)) {
            std::fprintf(stderr, "cannot read %s\n", argv[1]);
            rc = 2;
        } else {
            const int p = std::atoi(argv[3]);
            const int cap = std::atoi(argv[4]);
            const int max_level = std::atoi(argv[5]);
            const std::string mode = argv[6];
            const char* tree = argc > 7 ? argv[7] : nullptr;
            const int repeat = argc > 8 ? std::atoi(argv[8]) : 1;
            SelfList self;
            if (argc > 9 && !read_self_list(argv[9], in.n, self)) {
                std::fprintf(stderr, "cannot read self list %s\n", argv[9]);
                rc = 2;
            }
            // the potential-only modes build FMM3D without gradient tables
            const bool g = mode.rfind("potential", 0) != 0;
#define RUN(PP) rc = g ? run<PP, true>(in, argv[2], cap, max_level, mode, tree, repeat, self) \
                       : run<PP, false>(in, argv[2], cap, max_level, mode, tree, repeat, self)
            if (rc == 0) {
                switch (p) {
                case 4: RUN(4); break;
                case 6: RUN(6); break;
                case 8: RUN(8); break;
                case 10: RUN(10); break;
                case 12: RUN(12); break;
                case 16: RUN(16); break;
                default: std::fprintf(stderr, "unsupported p=%d\n", p); rc = 2;
                }
            }
#undef RUN
        }
    }
    MPI_Finalize();
    return rc;
}
