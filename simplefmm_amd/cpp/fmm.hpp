// C++17 drop-in for the public interface of icyspirit/SimpleFMM as its main.cpp
// drives it (box.hpp:73-178, partition.hpp:380-500, fmm.hpp:134-882,
// matrix.hpp:164-165): the class templates keep their names, template
// parameters, constructor arguments and accumulate-into-U semantics, and run on
// the MI355X through the sfmm C-ABI (include/sfmm.h).
//
// Differences a caller can see:
//   * the communicator argument is stored and returned by comm() but not used
//     (one process, one GPU): its type is MPI_Comm when <mpi.h> is included
//     first, an empty tag otherwise;
//   * T must be double; p must be in 1..16; gradient (curl) needs N == 3, as in
//     the reference (fmm.hpp:550);
//   * rinv(Q, U, is_self) and N2N(Q, U, is_self) with the predicate returned by
//     self_list() (j in self_generator(i)) exclude the listed pairs on the GPU.
//     Any other callable is evaluated on the host, as the reference evaluates it,
//     for every near-field pair (i, j != i): the GPU sums the near field without
//     those exclusions and the pairs with is_self(i, j) true are then subtracted
//     on the host in FP64.  This is exact but costs one host call per near pair.
//     The pair (i, i) is always excluded (the reference would divide by zero);
//   * far-field pairs of the generator (slist(), the reference's _slist) are
//     subtracted in every near-field sum, as in the reference (fmm.hpp:798-807);
//   * the per-level stage methods N2M, M2M, M2L, L2L, L2N (fmm.hpp:524-771) are
//     not exposed: the expansions live on the device in a level-scaled real form,
//     not in the reference's LevelData; rinv_nonear (far field), N2N (near field)
//     and the per-stage device times (sfmm_last_stage_ms) cover their uses;
//   * a failing call prints the library's message and aborts, where the
//     reference would assert.
#pragma once

#include <array>
#include <cmath>
#include <cstdio>
#include <cstdlib>
#include <functional>
#include <tuple>
#include <ostream>
#include <type_traits>
#include <utility>
#include <vector>

#include "../../include/sfmm.h"

#ifdef MPI_VERSION
using sfmm_comm_t = MPI_Comm;
#else
struct sfmm_comm_t {};  // no MPI: an opaque tag
#endif

namespace sfmm_compat {

[[noreturn]] inline void die(const char* what)
{
    std::fprintf(stderr, "sfmm: %s: %s\n", what, sfmm_last_error());
    std::abort();
}

inline void check(int rc, const char* what)
{
    if (rc != SFMM_OK) die(what);
}

}  // namespace sfmm_compat

// Vector<T, N>: contiguous std::array with the arithmetic main.cpp uses.
template <typename T, size_t N>
struct Vector : std::array<T, N> {
    Vector operator+(const Vector& o) const { Vector r; for (size_t i = 0; i < N; ++i) r[i] = (*this)[i] + o[i]; return r; }
    Vector operator-(const Vector& o) const { Vector r; for (size_t i = 0; i < N; ++i) r[i] = (*this)[i] - o[i]; return r; }
    Vector operator*(T s) const { Vector r; for (size_t i = 0; i < N; ++i) r[i] = (*this)[i] * s; return r; }
    Vector operator/(T s) const { Vector r; for (size_t i = 0; i < N; ++i) r[i] = (*this)[i] / s; return r; }
    friend Vector operator*(T s, const Vector& v) { return v * s; }
    Vector& operator+=(const Vector& o) { for (size_t i = 0; i < N; ++i) (*this)[i] += o[i]; return *this; }
    Vector& operator-=(const Vector& o) { for (size_t i = 0; i < N; ++i) (*this)[i] -= o[i]; return *this; }
    T dot(const Vector& o) const { T s = 0; for (size_t i = 0; i < N; ++i) s += (*this)[i] * o[i]; return s; }
    T norm() const { return std::sqrt(dot(*this)); }
    Vector cross(const Vector& o) const
    {
        static_assert(N == 3, "cross product of 3-vectors");
        const auto& a = *this;
        return Vector{{a[1] * o[2] - a[2] * o[1], a[2] * o[0] - a[0] * o[2], a[0] * o[1] - a[1] * o[0]}};
    }
    friend std::ostream& operator<<(std::ostream& os, const Vector& v)
    {
        for (size_t i = 0; i < N; ++i) os << v[i] << ' ';
        return os;
    }
};

template <typename T, int dim, template <typename, size_t> typename Container>
class Box {
    static_assert(dim == 3 && std::is_same_v<T, double>, "the evaluator handles 3-D FP64 boxes");

public:
    using Coord_t = Container<T, dim>;

    Box(const std::vector<Coord_t>& positions, int vertical_axis = dim - 1, T phideg = 0) :
        _vertical_axis(vertical_axis), _phideg(phideg), _given(false)
    {
        double out[4];
        sfmm_compat::check(sfmm_box_from_positions(static_cast<int64_t>(positions.size()),
                                                   reinterpret_cast<const double*>(positions.data()), vertical_axis,
                                                   phideg, out),
                           "Box");
        _width = out[0];
        _corner = Coord_t{{out[1], out[2], out[3]}};
    }

    Box(const Coord_t& corner, T width, int vertical_axis = dim - 1, T phideg = 0) :
        _vertical_axis(vertical_axis), _phideg(phideg), _given(true), _width(width), _corner(corner)
    {
    }

    T width() const noexcept { return _width; }
    const Coord_t& corner() const noexcept { return _corner; }
    int vertical_axis() const noexcept { return _vertical_axis; }
    T phideg() const noexcept { return _phideg; }
    bool given() const noexcept { return _given; }

    friend std::ostream& operator<<(std::ostream& os, const Box& b)
    {
        return os << "Box width = " << b._width << ", corner = " << b._corner;
    }

private:
    int _vertical_axis;
    T _phideg;
    bool _given;
    T _width;
    Coord_t _corner;
};

template <typename T, template <typename, size_t> typename Container>
class OctreePartitioner {
public:
    static constexpr int dim = 3;
    using Coord_t = Container<T, dim>;
    using Box_t = Box<T, dim, Container>;

    OctreePartitioner(const std::vector<Coord_t>& positions, const Box_t& box) : _positions(positions), _box(box) {}
    OctreePartitioner(const OctreePartitioner&) = delete;
    OctreePartitioner& operator=(const OctreePartitioner&) = delete;
    ~OctreePartitioner()
    {
        if (_tree) sfmm_tree_destroy(_tree);
    }

    // partition.hpp:405-446; on_device builds the same tree on the GPU
    void refine(int max_level, int max_particles_per_node, bool on_device = false)
    {
        sfmm_tree_desc d{};
        d.on_device = on_device ? 1 : 0;
        d.n = static_cast<int64_t>(_positions.size());
        d.positions = reinterpret_cast<const double*>(_positions.data());
        d.max_level = max_level;
        d.max_per_leaf = max_particles_per_node;
        d.vertical_axis = _box.vertical_axis();
        d.phideg = _box.phideg();
        d.box_given = _box.given() ? 1 : 0;
        for (int k = 0; k < 3; ++k) d.corner[k] = _box.corner()[k];
        d.width = _box.width();
        if (_tree) sfmm_tree_destroy(_tree);
        _tree = nullptr;
        sfmm_compat::check(sfmm_tree_create(&d, &_tree), "OctreePartitioner::refine");
    }

    const std::vector<Coord_t>& positions() const noexcept { return _positions; }
    const Box_t& box() const noexcept { return _box; }
    int level() const
    {
        sfmm_tree_info info{};
        sfmm_compat::check(sfmm_tree_get_info(_tree, &info), "OctreePartitioner::level");
        return info.depth;
    }
    sfmm_tree handle() const noexcept { return _tree; }

    friend std::ostream& operator<<(std::ostream& os, const OctreePartitioner& self)
    {
        for (int l = 0; l <= self.level(); ++l) {
            int64_t c[6];
            sfmm_compat::check(sfmm_tree_level_counts(self._tree, l, c), "OctreePartitioner::operator<<");
            os << "level = " << l << ", n_node = " << c[0] << ", n_leaf = " << c[1] << ", n_particle = " << c[2]
               << std::endl;
        }
        return os;
    }

private:
    const std::vector<Coord_t>& _positions;
    const Box_t& _box;
    sfmm_tree _tree = nullptr;
};

// Rows of (int) values, the read interface of the reference's CSRP<> (csr.hpp:58-103).
class CSRList {
public:
    CSRList() = default;
    CSRList(std::vector<int64_t> row, std::vector<int64_t> val) : _row(std::move(row)), _val(std::move(val)) {}
    int nrow() const noexcept { return _row.empty() ? 0 : static_cast<int>(_row.size() - 1); }
    int nnz(int row) const noexcept { return static_cast<int>(_row[row + 1] - _row[row]); }
    int nnz() const noexcept { return _row.empty() ? 0 : static_cast<int>(_row.back()); }
    std::tuple<int> value(int row, int k) const noexcept { return {static_cast<int>(_val[_row[row] + k])}; }

private:
    std::vector<int64_t> _row, _val;
};

template <typename T, size_t N, int p, template <typename, size_t> typename Container, bool support_gradient = false>
class FMM3D {
    static_assert(std::is_same_v<T, double>, "the MI355X evaluator computes in FP64");
    static_assert(p >= 1 && p <= 16, "expansion order must be in 1..16");
    static_assert(sizeof(Container<T, N>) == N * sizeof(T), "charges must be contiguous N-vectors");

public:
    using Partitioner_t = OctreePartitioner<T, Container>;

    using comm_type = sfmm_comm_t;

    // is_self(i, j) = "j is in self_generator(i)", evaluated on the GPU by rinv / N2N
    class SelfList {
    public:
        bool operator()(int i, int j) const noexcept
        {
            if (_row->empty()) return i == j;
            for (int64_t k = (*_row)[i]; k < (*_row)[i + 1]; ++k)
                if ((*_val)[k] == j) return true;
            return false;
        }

    private:
        friend class FMM3D;
        SelfList(const std::vector<int64_t>* row, const std::vector<int64_t>* val) : _row(row), _val(val) {}
        const std::vector<int64_t>* _row;
        const std::vector<int64_t>* _val;
    };

    template <typename Comm>
    FMM3D(const Partitioner_t& partitioner, Comm comm, const std::function<std::vector<int>(int)>& self_generator) :
        _partitioner(partitioner)
    {
        keep_comm(comm);
        const int64_t n = static_cast<int64_t>(partitioner.positions().size());
        _gen_row.assign(1, 0);
        for (int64_t i = 0; i < n; ++i) {
            for (int j : self_generator(static_cast<int>(i))) _gen_val.push_back(j);
            _gen_row.push_back(static_cast<int64_t>(_gen_val.size()));
        }
        create(_gen_row.data(), _gen_val.data());
    }

    template <typename Comm>
    FMM3D(const Partitioner_t& partitioner, Comm comm) : _partitioner(partitioner)
    {
        keep_comm(comm);
        create(nullptr, nullptr);
    }

    FMM3D(const FMM3D&) = delete;
    FMM3D& operator=(const FMM3D&) = delete;
    ~FMM3D()
    {
        if (_plan) sfmm_plan_destroy(_plan);
    }

    const Partitioner_t& partitioner() const noexcept { return _partitioner; }
    comm_type comm() const noexcept { return _comm; }

    // _slist (fmm.hpp:398-408,439-442): the generator's entries outside each particle's near field
    const CSRList& slist() const
    {
        if (!_slist_ready) {
            const int64_t n = static_cast<int64_t>(_partitioner.positions().size());
            int64_t nnz = 0;
            sfmm_compat::check(sfmm_plan_slist(_plan, nullptr, nullptr, &nnz), "FMM3D::slist");
            std::vector<int64_t> row(n + 1), val(nnz);
            sfmm_compat::check(sfmm_plan_slist(_plan, row.data(), val.data(), &nnz), "FMM3D::slist");
            _slist = CSRList(std::move(row), std::move(val));
            _slist_ready = true;
        }
        return _slist;
    }

    // The predicate rinv / N2N evaluate on the GPU: membership in self_generator(i).
    SelfList self_list() const noexcept { return SelfList(&_gen_row, &_gen_val); }

    template <bool gradient = false>
    void rinv(const Container<T, N>* Q, Container<T, N>* U) const noexcept
    {
        eval<gradient>(Q, U, SFMM_BOTH);
    }

    template <bool gradient = false, typename IsSelf>
    void rinv(const Container<T, N>* Q, Container<T, N>* U, const IsSelf& is_self) const noexcept
    {
        if constexpr (std::is_same_v<IsSelf, SelfList>) {
            eval<gradient>(Q, U, SFMM_BOTH | SFMM_SELF_LIST);
        } else {
            eval<gradient>(Q, U, SFMM_BOTH);
            subtract_excluded<gradient>(Q, U, is_self);
        }
    }

    template <bool gradient = false>
    void rinv_nonear(const Container<T, N>* Q, Container<T, N>* U) const noexcept
    {
        eval<gradient>(Q, U, SFMM_FAR);
    }

    template <bool gradient = false, typename IsSelf>
    void N2N(const Container<T, N>* Q, Container<T, N>* U, const IsSelf& is_self) const noexcept
    {
        if constexpr (std::is_same_v<IsSelf, SelfList>) {
            eval<gradient>(Q, U, SFMM_NEAR | SFMM_SELF_LIST);
        } else {
            eval<gradient>(Q, U, SFMM_NEAR);
            subtract_excluded<gradient>(Q, U, is_self);
        }
    }

    // Scalar charges q[i]: out[4i..4i+3] += phi, dphi/dx, dphi/dy, dphi/dz.
    void rinv_potgrad(const T* q, T* out, int which = SFMM_BOTH) const noexcept
    {
        sfmm_compat::check(sfmm_evaluate(_plan, SFMM_POTGRAD, 1, q, out, which), "FMM3D::rinv_potgrad");
    }

private:
    template <typename Comm>
    void keep_comm(const Comm& comm)
    {
        if constexpr (std::is_convertible_v<Comm, comm_type>) _comm = comm;
    }

    // The near-field pairs (i, j != i) a general is_self excludes, subtracted on the
    // host in FP64 after the GPU summed them (N2N, fmm.hpp:786-796).
    template <bool gradient, typename IsSelf>
    void subtract_excluded(const Container<T, N>* Q, Container<T, N>* U, const IsSelf& is_self) const
    {
        const auto& x = _partitioner.positions();
        const int64_t n = static_cast<int64_t>(x.size());
        std::vector<int64_t> src;
        for (int64_t i = 0; i < n; ++i) {
            int64_t cnt = 0;
            sfmm_compat::check(sfmm_tree_near_sources(_partitioner.handle(), i, nullptr, 0, &cnt), "FMM3D::N2N");
            src.resize(static_cast<size_t>(cnt));
            sfmm_compat::check(sfmm_tree_near_sources(_partitioner.handle(), i, src.data(), cnt, &cnt), "FMM3D::N2N");
            for (int64_t j : src) {
                if (j == i || !is_self(static_cast<int>(i), static_cast<int>(j))) continue;
                T R[3];
                for (int d = 0; d < 3; ++d) R[d] = x[i][d] - x[j][d];
                const T r = std::sqrt(R[0] * R[0] + R[1] * R[1] + R[2] * R[2]);
                if constexpr (gradient) {
                    const T r3 = r * r * r;
                    const auto& q = Q[j];
                    U[i][0] -= (q[1] * R[2] - q[2] * R[1]) / r3;
                    U[i][1] -= (q[2] * R[0] - q[0] * R[2]) / r3;
                    U[i][2] -= (q[0] * R[1] - q[1] * R[0]) / r3;
                } else {
                    for (size_t c = 0; c < N; ++c) U[i][c] -= Q[j][c] / r;
                }
            }
        }
    }

    void create(const int64_t* self_row, const int64_t* self_val)
    {
        sfmm_plan_desc d{};
        d.p = p;
        d.precision = SFMM_FP64;
        d.max_ncharge = N < 4 ? static_cast<int>(N) : 4;
        d.self_row = self_row;
        d.self_val = self_val;
        sfmm_compat::check(sfmm_plan_create(_partitioner.handle(), &d, &_plan), "FMM3D");
        std::printf("FMM with p = %d, n_particle = %zu\n", p, _partitioner.positions().size());
    }

    template <bool gradient>
    void eval(const Container<T, N>* Q, Container<T, N>* U, int which) const noexcept
    {
        static_assert(support_gradient || !gradient, "gradient needs support_gradient");
        const int64_t n = static_cast<int64_t>(_partitioner.positions().size());
        const T* q = reinterpret_cast<const T*>(Q);
        T* u = reinterpret_cast<T*>(U);
        if constexpr (gradient) {
            static_assert(N == 3, "the curl field needs 3-vector charges");
            sfmm_compat::check(sfmm_evaluate(_plan, SFMM_CURL, 3, q, u, which), "FMM3D::rinv<true>");
        } else if constexpr (N <= 4) {
            sfmm_compat::check(sfmm_evaluate(_plan, SFMM_POTENTIAL, static_cast<int>(N), q, u, which), "FMM3D::rinv");
        } else {
            // the potential is linear per component: evaluate N components in groups of <= 4
            for (size_t c0 = 0; c0 < N; c0 += 4) {
                const int nc = static_cast<int>(N - c0 < 4 ? N - c0 : 4);
                std::vector<T> qs(static_cast<size_t>(n) * nc), us(static_cast<size_t>(n) * nc, T(0));
                for (int64_t i = 0; i < n; ++i)
                    for (int c = 0; c < nc; ++c) qs[i * nc + c] = q[i * N + c0 + c];
                sfmm_compat::check(sfmm_evaluate(_plan, SFMM_POTENTIAL, nc, qs.data(), us.data(), which),
                                   "FMM3D::rinv");
                for (int64_t i = 0; i < n; ++i)
                    for (int c = 0; c < nc; ++c) u[i * N + c0 + c] += us[i * nc + c];
            }
        }
    }

    const Partitioner_t& _partitioner;
    sfmm_plan _plan = nullptr;
    comm_type _comm{};
    std::vector<int64_t> _gen_row, _gen_val;  // self_generator lists (empty: i -> {i})
    mutable CSRList _slist;
    mutable bool _slist_ready = false;
};
