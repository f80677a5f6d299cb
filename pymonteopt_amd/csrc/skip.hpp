// Per-voxel radius of the largest uniform cube around each voxel, which lets the walk cross
// the faces inside a same-label region in one step.
//
// Let B be the voxels that lie on the volume's outer face, have one of their 26 neighbours
// under another label, or hold label 255 (the marker of a mixed neighbourhood below; only a
// volume of 255 tissue labels has it, as its end label, and a packet never moves inside an
// end voxel). If a voxel u is at Chebyshev distance d >= k from B, the cube of half-width k
// voxels around u holds u's label only: every voxel of the cube of half-width j < k is
// outside B, so its 26 neighbours, which cover the cube of half-width j + 1, share its label.
// The radius is r(u) = min(cap, d(u, B)), 0 on B itself.
//
// B, separably: w_z = the label where the three voxels z-1, z, z+1 of a column share it,
// else 255; w_zy the same over y-1, y, y+1 of w_z; w_zyx over x of w_zy. w_zyx is a label
// (not 255) exactly where the 3x3x3 cube around the voxel holds that one label.
//
// d(u, B) = min over x' of max(|x - x'|, min over y' of max(|y - y'|, min over z' of
// max(|z - z'|, b(x', y', z')))), with b = 0 on B and cap + 1 elsewhere: three passes of a
// windowed min-max along z, y and x. Capped at cap, each pass looks at most cap voxels each
// way, and stops once the offset reaches the running minimum.
//
// The z passes read neighbouring bytes (one thread per voxel, cached lines). The y and x
// passes stride by a row or a plane, so there each thread sweeps one line along the axis,
// threads of a block on consecutive z (coalesced), keeping the window in registers (the
// seed) or in its own column of an LDS ring (the min-max): each byte is read once.
#pragma once

#include <hip/hip_runtime.h>
#include <cstdint>

namespace pmo {

constexpr uint32_t kSkipMixed = 255u;
constexpr uint32_t kSkipSweepBlock = 256u;

// w_z: one thread per voxel
__global__ void __launch_bounds__(256) skip_seed_z_kernel(const uint8_t* __restrict__ lab, uint32_t n, uint32_t nz,
                                                          uint8_t* __restrict__ out)
{
    for (uint32_t i = blockIdx.x * blockDim.x + threadIdx.x; i < n; i += gridDim.x * blockDim.x) {
        const uint32_t z = i % nz;
        uint32_t w = kSkipMixed;
        if (z > 0u && z + 1u < nz) {
            const uint32_t c = lab[i];
            if (lab[i - 1] == c && lab[i + 1] == c) w = c;
        }
        out[i] = (uint8_t)w;
    }
}

// The volume seen as [n_outer][len][S]; the thread of (outer = outer0 + blockIdx.y, inner)
// sweeps the line base + k*S, k = 0 .. len-1, base = outer*len*S + inner.
//
// w along the axis: the value where positions k-1, k, k+1 share it, else 255 (255 at both
// ends). final: write b instead (cap + 1 where w is a label, else 0).
__global__ void __launch_bounds__(256) skip_seed_sweep_kernel(const uint8_t* __restrict__ in, uint8_t* __restrict__ out,
                                                              uint32_t outer0, uint32_t len, uint32_t S, uint32_t cap,
                                                              uint32_t final_pass)
{
    const uint32_t inner = blockIdx.x * blockDim.x + threadIdx.x;
    if (inner >= S) return;
    const size_t base = ((size_t)outer0 + blockIdx.y) * len * S + inner;
    uint32_t prev = kSkipMixed, cur = in[base];
    for (uint32_t k = 0; k < len; ++k) {
        const uint32_t next = k + 1u < len ? (uint32_t)in[base + (size_t)(k + 1u) * S] : kSkipMixed;
        uint32_t w = (k > 0u && k + 1u < len && prev == cur && next == cur) ? cur : kSkipMixed;
        if (final_pass) w = w != kSkipMixed ? cap + 1u : 0u;
        out[base + (size_t)k * S] = (uint8_t)w;
        prev = cur;
        cur = next;
    }
}

// min-max along z (stride 1): out[i] = min over |d| <= cap (inside the axis) of
// max(|d|, in[i + d]); the last pass clamps to cap
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

// min-max along a strided axis by line sweeps (layout as skip_seed_sweep_kernel). The ring
// in LDS holds, per thread, the 2*cap + 1 values around the output position: slot j of
// thread t at ring[j * blockDim.x + t], written and read by that thread only.
__global__ void __launch_bounds__(256) skip_minmax_sweep_kernel(const uint8_t* __restrict__ in, uint8_t* __restrict__ out,
                                                                uint32_t outer0, uint32_t len, uint32_t S, uint32_t cap,
                                                                uint32_t clamp)
{
    extern __shared__ uint8_t ring[];
    const uint32_t tid = threadIdx.x, bd = blockDim.x;
    const uint32_t inner = blockIdx.x * bd + tid;
    if (inner >= S) return;
    const size_t base = ((size_t)outer0 + blockIdx.y) * len * S + inner;
    const uint32_t W = 2u * cap + 1u;
    for (uint32_t j = 0; j < W; ++j) ring[j * bd + tid] = (uint8_t)255u;
    // value t goes to slot t mod W; output k = t - cap once value t = k + cap is in
    uint32_t ts = 0;                                   // t mod W
    for (uint32_t t = 0; t < len + cap; ++t) {
        ring[ts * bd + tid] = t < len ? in[base + (size_t)t * S] : (uint8_t)255u;
        if (t >= cap) {
            const uint32_t k = t - cap;
            const uint32_t ks = ts >= cap ? ts - cap : ts + W - cap;      // k mod W
            uint32_t m = ring[ks * bd + tid];
            for (uint32_t d = 1u; d <= cap && d < m; ++d) {
                const uint32_t ls = ks >= d ? ks - d : ks + W - d;
                const uint32_t hs = ks + d < W ? ks + d : ks + d - W;
                const uint32_t lo = k >= d ? (uint32_t)ring[ls * bd + tid] : 255u;
                const uint32_t hi = ring[hs * bd + tid];                   // 255 past the end
                const uint32_t a = lo < hi ? lo : hi;
                const uint32_t v = a > d ? a : d;
                m = v < m ? v : m;
            }
            out[base + (size_t)k * S] = (uint8_t)(clamp && m > cap ? cap : m);
        }
        ts = ts + 1u == W ? 0u : ts + 1u;
    }
}

}  // namespace pmo
