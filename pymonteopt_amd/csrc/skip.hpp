// Per-voxel radius of the largest uniform cube around each voxel, which lets the walk cross
// the faces inside a same-label region in one step.
//
// Let B be the voxels that lie on the volume's outer face or have one of their 26
// neighbours under another label. If a voxel u is at Chebyshev distance d >= k from B, the
// cube of half-width k voxels around u holds u's label only: every voxel of the cube of
// half-width j < k is outside B, so its 26 neighbours, which cover the cube of half-width
// j + 1, share its label. The radius is r(u) = min(cap, d(u, B)), 0 on B itself.
//
// d(u, B) = min over x' of max(|x - x'|, min over y' of max(|y - y'|, min over z' of
// max(|z - z'|, b(x', y', z')))), with b = 0 on B and cap + 1 elsewhere, so the distance is
// three passes of a windowed min-max along z, y and x. Capped at cap, each pass looks at
// most cap voxels each way, and stops once the offset reaches the running minimum.
#pragma once

#include <hip/hip_runtime.h>
#include <cstdint>

namespace pmo {

// b: 0 on B, cap + 1 elsewhere
__global__ void __launch_bounds__(256) skip_seed_kernel(const uint8_t* __restrict__ lab, uint32_t nx, uint32_t ny,
                                                        uint32_t nz, uint32_t cap, uint8_t* __restrict__ out)
{
    const uint32_t n = nx * ny * nz;
    const uint32_t syz = ny * nz;
    for (uint32_t i = blockIdx.x * blockDim.x + threadIdx.x; i < n; i += gridDim.x * blockDim.x) {
        const uint32_t z = i % nz;
        const uint32_t y = (i / nz) % ny;
        const uint32_t x = i / syz;
        uint32_t v = cap + 1u;
        if (x == 0u || y == 0u || z == 0u || x + 1u == nx || y + 1u == ny || z + 1u == nz) {
            v = 0u;
        } else {
            const uint8_t c = lab[i];
            for (int dx = -1; dx <= 1 && v != 0u; ++dx)
                for (int dy = -1; dy <= 1; ++dy) {
                    const uint8_t* row = lab + (int64_t)i + (int64_t)dx * syz + (int64_t)dy * nz;
                    if (row[-1] != c || row[0] != c || row[1] != c) { v = 0u; break; }
                }
        }
        out[i] = (uint8_t)v;
    }
}

// out[i] = min over |d| <= cap (inside the axis) of max(|d|, in[i + d * stride]); the last
// pass clamps to cap
__global__ void __launch_bounds__(256) skip_pass_kernel(const uint8_t* __restrict__ in, uint8_t* __restrict__ out,
                                                        uint32_t n, uint32_t len, uint32_t stride, uint32_t cap,
                                                        uint32_t clamp)
{
    for (uint32_t i = blockIdx.x * blockDim.x + threadIdx.x; i < n; i += gridDim.x * blockDim.x) {
        const uint32_t c = (i / stride) % len;
        uint32_t m = in[i];
        // four offsets per round, their eight loads independent of each other; a round
        // starts only while an offset can still lower m (min-max with offset d is >= d)
        for (uint32_t d0 = 1u; d0 <= cap && d0 < m; d0 += 4u) {
#pragma unroll
            for (uint32_t j = 0u; j < 4u; ++j) {
                const uint32_t d = d0 + j;
                const uint32_t lo = (d <= cap && c >= d) ? in[i - d * stride] : 255u;
                const uint32_t hi = (d <= cap && c + d < len) ? in[i + d * stride] : 255u;
                const uint32_t a = lo < hi ? lo : hi;
                const uint32_t v = a > d ? a : d;
                m = v < m ? v : m;
            }
        }
        out[i] = (uint8_t)(clamp && m > cap ? cap : m);
    }
}

}  // namespace pmo
