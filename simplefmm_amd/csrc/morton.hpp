// Cell codes shared by the host tree build (g++) and the device tree build (hipcc),
// so both compute bit-identical codes.
//
// A point's cell at level L is clamp(int(u * 2^L)) per axis with u = (r - corner) / width
// and r the point rotated about the box's vertical axis (box.hpp:57-63,152-157); the
// three 21-bit indices are interleaved as x bit 0, y bit 1, z bit 2 (zorder.hpp:26-58).
#pragma once

#include <cstdint>

#if !defined(__HIPCC__) && !defined(__host__)
#define __host__
#define __device__
#endif

namespace sfmm {

__host__ __device__ inline int64_t spread3(uint32_t v)  // bit b -> bit 3b, 21 bits
{
    uint64_t x = v & 0x1fffff;
    x = (x | x << 32) & 0x1f00000000ffffULL;
    x = (x | x << 16) & 0x1f0000ff0000ffULL;
    x = (x | x << 8) & 0x100f00f00f00f00fULL;
    x = (x | x << 4) & 0x10c30c30c30c30c3ULL;
    x = (x | x << 2) & 0x1249249249249249ULL;
    return static_cast<int64_t>(x);
}

__host__ __device__ inline uint32_t compact3(uint64_t x)  // inverse of spread3
{
    x &= 0x1249249249249249ULL;
    x = (x ^ (x >> 2)) & 0x10c30c30c30c30c3ULL;
    x = (x ^ (x >> 4)) & 0x100f00f00f00f00fULL;
    x = (x ^ (x >> 8)) & 0x1f0000ff0000ffULL;
    x = (x ^ (x >> 16)) & 0x1f00000000ffffULL;
    x = (x ^ (x >> 32)) & 0x1fffffULL;
    return static_cast<uint32_t>(x);
}

__host__ __device__ inline int32_t code_axis(int64_t code, int axis)
{
    return static_cast<int32_t>(compact3(static_cast<uint64_t>(code) >> axis));
}

// interaction offset d = source - target, |d_i| <= 3, as one id in [0, 343)
__host__ __device__ inline int offset_id(int di, int dj, int dk) { return ((dk + 3) * 7 + (dj + 3)) * 7 + (di + 3); }

// Rotation about the vertical axis (box.hpp:152-157): (x, y) -> (c x - s y, s x + c y).
// Products and sums are rounded separately on both compilers (no fused multiply-add).
__host__ __device__ inline void rotate_xy(double c, double s, double x, double y, double& rx, double& ry)
{
#if defined(__clang__)
#pragma clang fp contract(off)
#endif
    rx = c * x - s * y;
    ry = s * x + c * y;
}

// Cell index of rotated coordinate r at level L (box.hpp:57-63).  Returns false when
// u = (r - corner) / width is outside [0, 1] or not finite (box.hpp:60 asserts it).
__host__ __device__ inline bool cell_index(double r, double corner, double width, int L, int32_t& out)
{
#if defined(__clang__)
#pragma clang fp contract(off)
#endif
    const double u = (r - corner) / width;
    if (!(u >= 0 && u <= 1)) {
        out = 0;
        return false;
    }
    const int32_t imax = (int32_t(1) << L) - 1;
    const int32_t v = static_cast<int32_t>(u * (1 << L));
    out = v < 0 ? 0 : (v > imax ? imax : v);
    return true;
}

}  // namespace sfmm
