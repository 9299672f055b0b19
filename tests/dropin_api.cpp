// Calls the C++ drop-in (simplefmm_amd/cpp/fmm.hpp) the way code written for the
// reference's FMM3D does (fmm.hpp:204,423-442,773-881): a self_generator
// constructor, comm(), slist(), and rinv / rinv<true> / rinv_nonear / N2N with
// the library's own predicate and with a plain lambda.  Built and run by
// tests/test_gpu_api.py; the outputs are compared with the oracle there.
//
// usage: dropin_api IN SELFLIST OUT
//   IN: 'SFMI' input with 3-vector charges; SELFLIST: 'SFMS' lists;
//   OUT: 'SFMD', int64 n, then blocks of 3n doubles in this order:
//        pot (self_list()), pot (lambda), curl (self_list()), curl (lambda),
//        near curl (N2N, lambda), far pot (rinv_nonear);
//        then int64 nnz, int64 row[n+1], int64 val[nnz] of slist().
#include <cstdint>
#include <cstdio>
#include <cstring>
#include <fstream>
#include <iostream>
#include <vector>

#include "fmm.hpp"

using V3 = Vector<double, 3>;

namespace {

bool read_all(const char* path, std::vector<char>& bytes)
{
    std::ifstream f(path, std::ios::binary);
    bytes.assign(std::istreambuf_iterator<char>(f), std::istreambuf_iterator<char>());
    return static_cast<bool>(f) || f.eof();
}

}  // namespace

int main(int argc, char** argv)
{
    if (argc != 4) {
        std::cerr << "usage: dropin_api IN SELFLIST OUT" << std::endl;
        return 2;
    }
    std::vector<char> in, sl;
    if (!read_all(argv[1], in) || in.size() < 20 || std::memcmp(in.data(), "SFMI", 4) != 0) return 2;
    int64_t n;
    int32_t qdim;
    std::memcpy(&n, in.data() + 8, 8);
    std::memcpy(&qdim, in.data() + 16, 4);
    if (qdim != 3 || in.size() != 20 + static_cast<size_t>(n) * 6 * sizeof(double)) return 2;
    std::vector<V3> pos(n), Q(n);
    std::memcpy(pos.data(), in.data() + 20, n * sizeof(V3));
    std::memcpy(Q.data(), in.data() + 20 + n * sizeof(V3), n * sizeof(V3));

    if (!read_all(argv[2], sl) || sl.size() < 20 || std::memcmp(sl.data(), "SFMS", 4) != 0) return 2;
    int64_t sn, nnz;
    std::memcpy(&sn, sl.data() + 4, 8);
    std::memcpy(&nnz, sl.data() + 12, 8);
    if (sn != n) return 2;
    std::vector<int64_t> row(n + 1), val(nnz);
    std::memcpy(row.data(), sl.data() + 20, (n + 1) * 8);
    std::memcpy(val.data(), sl.data() + 20 + (n + 1) * 8, nnz * 8);

    Box<double, 3, Vector> box(pos, 2);
    OctreePartitioner<double, Vector> part(pos, box);
    part.refine(20, 24);
    const std::function<std::vector<int>(int)> generator = [&](int i) {
        return std::vector<int>(val.begin() + row[i], val.begin() + row[i + 1]);
    };
    const int comm = 0;
    FMM3D<double, 3, 6, Vector, true> fmm(part, comm, generator);
    const auto c = fmm.comm();
    (void)c;
    const auto member = [&](int i, int j) {
        for (int64_t k = row[i]; k < row[i + 1]; ++k)
            if (val[k] == j) return true;
        return false;
    };

    std::vector<std::vector<V3>> out(6, std::vector<V3>(n, V3{{0.0, 0.0, 0.0}}));
    fmm.rinv(Q.data(), out[0].data(), fmm.self_list());
    fmm.rinv(Q.data(), out[1].data(), member);
    fmm.rinv<true>(Q.data(), out[2].data(), fmm.self_list());
    fmm.rinv<true>(Q.data(), out[3].data(), member);
    fmm.N2N<true>(Q.data(), out[4].data(), member);
    fmm.rinv_nonear(Q.data(), out[5].data());

    const auto& slist = fmm.slist();
    FILE* f = std::fopen(argv[3], "wb");
    if (!f) return 3;
    std::fwrite("SFMD", 1, 4, f);
    std::fwrite(&n, 8, 1, f);
    for (const auto& o : out) std::fwrite(o.data(), sizeof(V3), n, f);
    const int64_t snnz = slist.nnz();
    std::fwrite(&snnz, 8, 1, f);
    std::vector<int64_t> srow(n + 1, 0), sval;
    for (int i = 0; i < slist.nrow(); ++i) {
        for (int k = 0; k < slist.nnz(i); ++k) sval.push_back(std::get<0>(slist.value(i, k)));
        srow[i + 1] = static_cast<int64_t>(sval.size());
    }
    std::fwrite(srow.data(), 8, n + 1, f);
    std::fwrite(sval.data(), 8, sval.size(), f);
    std::fclose(f);
    std::cout << "dropin_api: n = " << n << ", slist nnz = " << snnz << std::endl;
    return 0;
}
