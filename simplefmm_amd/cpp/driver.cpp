// Command-line counterpart of the reference driver (main.cpp:14-115): random
// particles in the unit cube with random 3-vector charges (or the particles and
// charges of an input file), Box + partitioner + FMM3D<double,3,p,Vector,true>,
// rinv and rinv<true>, then the relative L2 error against an FP64 direct sum.
// The O(N^2) check of main.cpp:70-96 runs on the GPU over a sample of targets
// (sfmm_direct) instead of all N.
//
// usage: sfmm_driver [N] [max_particles_per_node] [n_targets] [seed] [--in FILE] [--out FILE] [--device-tree]
//   --device-tree: build the octree on the GPU (the same tree as the host build)
//   --in  FILE: 'SFMI' | int32 1 | int64 n | int32 qdim (1 or 3) | double pos[3n] | double q[qdim n]
//   --out FILE: 'SFMR' | int32 1 | int64 n | int32 nout | uint64 FNV-1a hash of the input file |
//               double out[nout n]; qdim 3: potential (3) then curl field (3) per particle,
//               qdim 1: potential and gradient (4) per particle.
#include <chrono>
#include <cstdint>
#include <cstdio>
#include <cstdlib>
#include <cstring>
#include <fstream>
#include <iostream>
#include <string>
#include <vector>

#include "fmm.hpp"

constexpr int dim = 3;
constexpr int expansion_order = 8;
using Vector_t = Vector<double, dim>;

namespace {

// portable seeded generator (splitmix64), in place of std::rand
struct SplitMix {
    uint64_t s;
    double next()
    {
        uint64_t z = (s += 0x9E3779B97F4A7C15ULL);
        z = (z ^ (z >> 30)) * 0xBF58476D1CE4E5B9ULL;
        z = (z ^ (z >> 27)) * 0x94D049BB133111EBULL;
        return static_cast<double>((z ^ (z >> 31)) >> 11) * (1.0 / 9007199254740992.0);
    }
};

uint64_t fnv1a(const std::vector<char>& bytes)
{
    uint64_t h = 1469598103934665603ULL;
    for (char c : bytes) {
        h ^= static_cast<unsigned char>(c);
        h *= 1099511628211ULL;
    }
    return h;
}

// Reads an input file; returns false (with a message) when it is malformed.
bool read_input(const std::string& path, std::vector<Vector_t>& pos, std::vector<double>& q, int& qdim,
                uint64_t& hash)
{
    std::ifstream f(path, std::ios::binary);
    std::vector<char> bytes((std::istreambuf_iterator<char>(f)), std::istreambuf_iterator<char>());
    if (bytes.size() < 20 || std::memcmp(bytes.data(), "SFMI", 4) != 0) {
        std::cerr << "not an SFMI file: " << path << std::endl;
        return false;
    }
    int32_t ver;
    int64_t n;
    std::memcpy(&ver, bytes.data() + 4, 4);
    std::memcpy(&n, bytes.data() + 8, 8);
    std::memcpy(&qdim, bytes.data() + 16, 4);
    if (ver != 1 || n <= 0 || (qdim != 1 && qdim != 3) ||
        bytes.size() != 20 + static_cast<size_t>(n) * (3 + qdim) * sizeof(double)) {
        std::cerr << "malformed SFMI file: " << path << std::endl;
        return false;
    }
    pos.resize(n);
    q.resize(static_cast<size_t>(n) * qdim);
    std::memcpy(pos.data(), bytes.data() + 20, static_cast<size_t>(n) * 3 * sizeof(double));
    std::memcpy(q.data(), bytes.data() + 20 + static_cast<size_t>(n) * 3 * sizeof(double), q.size() * sizeof(double));
    hash = fnv1a(bytes);
    return true;
}

bool write_output(const std::string& path, int64_t n, int32_t nout, uint64_t hash, const std::vector<double>& out)
{
    std::ofstream f(path, std::ios::binary);
    const int32_t ver = 1;
    f.write("SFMR", 4);
    f.write(reinterpret_cast<const char*>(&ver), 4);
    f.write(reinterpret_cast<const char*>(&n), 8);
    f.write(reinterpret_cast<const char*>(&nout), 4);
    f.write(reinterpret_cast<const char*>(&hash), 8);
    f.write(reinterpret_cast<const char*>(out.data()), static_cast<std::streamsize>(out.size() * sizeof(double)));
    return static_cast<bool>(f);
}

double seconds()
{
    return std::chrono::duration<double>(std::chrono::steady_clock::now().time_since_epoch()).count();
}

}  // namespace

int main(int argc, char** argv)
{
    std::vector<std::string> args;
    std::string in_path, out_path;
    bool device_tree = false;
    for (int i = 1; i < argc; ++i) {
        const std::string a = argv[i];
        if ((a == "--in" || a == "--out") && i + 1 < argc) (a == "--in" ? in_path : out_path) = argv[++i];
        else if (a == "--device-tree") device_tree = true;
        else args.push_back(a);
    }
    size_t N = args.size() > 0 ? std::strtoull(args[0].c_str(), nullptr, 10) : 200000;
    const int max_particles_per_node = args.size() > 1 ? std::atoi(args[1].c_str()) : 128;
    const int64_t ntarget = args.size() > 2 ? std::atoll(args[2].c_str()) : 2000;
    SplitMix rng{args.size() > 3 ? std::strtoull(args[3].c_str(), nullptr, 10) : 1};

    std::vector<Vector_t> positions, vector_charges;
    std::vector<double> scalar_charges;
    int qdim = 3;
    uint64_t in_hash = 0;
    if (!in_path.empty()) {
        std::vector<double> q;
        if (!read_input(in_path, positions, q, qdim, in_hash)) return 2;
        N = positions.size();
        if (qdim == 3) {
            vector_charges.resize(N);
            std::memcpy(vector_charges.data(), q.data(), q.size() * sizeof(double));
        } else {
            scalar_charges = q;
        }
    } else {
        positions.reserve(N);
        vector_charges.reserve(N);
        for (size_t i = 0; i < N; ++i) {
            positions.push_back({{rng.next(), rng.next(), rng.next()}});
            vector_charges.push_back({{rng.next(), rng.next(), rng.next()}});
        }
    }

    const int vertical_axis = 2;
    Box<double, dim, Vector> box(positions, vertical_axis);
    std::cout << box << std::endl;

    OctreePartitioner<double, Vector> partitioner(positions, box);
    partitioner.refine(20, max_particles_per_node, device_tree);
    std::cout << partitioner << std::endl;

    int comm = 0;  // stands for the communicator argument; ignored
    FMM3D<double, dim, expansion_order, Vector, true> fmm(partitioner, comm);

    if (qdim == 1) {  // scalar charges: potential and gradient
        std::vector<double> U(4 * N, 0.0);
        const double t0 = seconds();
        fmm.rinv_potgrad(scalar_charges.data(), U.data());
        std::cout << "Elapsed time for FMM potential+gradient: " << seconds() - t0 << " [s]" << std::endl;
        if (!out_path.empty() && !write_output(out_path, static_cast<int64_t>(N), 4, in_hash, U)) return 3;
        return 0;
    }

    std::vector<Vector_t> potentials(N), fields(N);
    double t0 = seconds();
    fmm.rinv(vector_charges.data(), potentials.data());
    double t1 = seconds();
    fmm.rinv<true>(vector_charges.data(), fields.data());
    double t2 = seconds();
    std::cout << "Elapsed time for FMM potential calculation: " << t1 - t0 << " [s]" << std::endl;
    std::cout << "Elapsed time for FMM field calculation: " << t2 - t1 << " [s]" << std::endl;

    // sampled FP64 direct sums on the device
    std::vector<int64_t> targets;
    for (int64_t k = 0; k < ntarget && k < static_cast<int64_t>(N); ++k)
        targets.push_back(static_cast<int64_t>((k * 2654435761ULL) % N));
    sfmm_tree_desc td{};
    td.n = static_cast<int64_t>(N);
    td.positions = reinterpret_cast<const double*>(positions.data());
    td.max_level = 20;
    td.max_per_leaf = max_particles_per_node;
    td.vertical_axis = vertical_axis;
    sfmm_tree tree = nullptr;
    sfmm_plan plan = nullptr;
    sfmm_plan_desc pd{};
    pd.p = 1;
    pd.max_ncharge = 3;
    sfmm_compat::check(sfmm_tree_create(&td, &tree), "direct-sum tree");
    sfmm_compat::check(sfmm_plan_create(tree, &pd, &plan), "direct-sum plan");
    std::vector<double> dpot(3 * targets.size()), dfield(3 * targets.size());
    const double* q = reinterpret_cast<const double*>(vector_charges.data());
    sfmm_compat::check(sfmm_direct(plan, SFMM_POTENTIAL, 3, q, static_cast<int64_t>(targets.size()), targets.data(),
                                   dpot.data()),
                       "direct potential");
    sfmm_compat::check(sfmm_direct(plan, SFMM_CURL, 3, q, static_cast<int64_t>(targets.size()), targets.data(),
                                   dfield.data()),
                       "direct field");
    sfmm_plan_destroy(plan);
    sfmm_tree_destroy(tree);

    // relative L2 as main.cpp:99-111, over the sampled targets
    double np = 0, dp = 0, nf = 0, df = 0;
    for (size_t k = 0; k < targets.size(); ++k)
        for (int c = 0; c < 3; ++c) {
            const double a = potentials[targets[k]][c], b = dpot[3 * k + c];
            const double e = fields[targets[k]][c], f = dfield[3 * k + c];
            np += (a - b) * (a - b);
            dp += b * b;
            nf += (e - f) * (e - f);
            df += f * f;
        }
    std::cout << "Relative l2 error of potential = " << std::sqrt(np / dp) << std::endl;
    std::cout << "Relative l2 error of field = " << std::sqrt(nf / df) << std::endl;
    if (!out_path.empty()) {
        std::vector<double> out(6 * N);
        for (size_t i = 0; i < N; ++i)
            for (int c = 0; c < 3; ++c) {
                out[6 * i + c] = potentials[i][c];
                out[6 * i + 3 + c] = fields[i][c];
            }
        if (!write_output(out_path, static_cast<int64_t>(N), 6, in_hash, out)) return 3;
    }
    return 0;
}
