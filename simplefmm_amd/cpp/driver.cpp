// Command-line counterpart of the reference driver (main.cpp:14-115): random
// particles in the unit cube with random 3-vector charges, Box + partitioner +
// FMM3D<double,3,p,Vector,true>, rinv and rinv<true>, then the relative L2
// error against an FP64 direct sum.  The O(N^2) check of main.cpp:70-96 runs
// on the GPU over a sample of targets (sfmm_direct) instead of all N.
//
// usage: sfmm_driver [N] [max_particles_per_node] [n_targets] [seed]
#include <chrono>
#include <cstdint>
#include <cstdio>
#include <cstdlib>
#include <iostream>
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

double seconds()
{
    return std::chrono::duration<double>(std::chrono::steady_clock::now().time_since_epoch()).count();
}

}  // namespace

int main(int argc, char** argv)
{
    const size_t N = argc > 1 ? std::strtoull(argv[1], nullptr, 10) : 200000;
    const int max_particles_per_node = argc > 2 ? std::atoi(argv[2]) : 128;
    const int64_t ntarget = argc > 3 ? std::atoll(argv[3]) : 2000;
    SplitMix rng{argc > 4 ? std::strtoull(argv[4], nullptr, 10) : 1};

    std::vector<Vector_t> positions, vector_charges;
    positions.reserve(N);
    vector_charges.reserve(N);
    for (size_t i = 0; i < N; ++i) {
        positions.push_back({{rng.next(), rng.next(), rng.next()}});
        vector_charges.push_back({{rng.next(), rng.next(), rng.next()}});
    }

    const int vertical_axis = 2;
    Box<double, dim, Vector> box(positions, vertical_axis);
    std::cout << box << std::endl;

    OctreePartitioner<double, Vector> partitioner(positions, box);
    partitioner.refine(20, max_particles_per_node);
    std::cout << partitioner << std::endl;

    int comm = 0;  // stands for the communicator argument; ignored
    FMM3D<double, dim, expansion_order, Vector, true> fmm(partitioner, comm);

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
    return 0;
}
