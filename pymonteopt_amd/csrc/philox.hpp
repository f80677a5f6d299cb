// Counter-based Philox4x32-10 (Salmon, Moraes, Dror, Shaw, SC'11) for the photon walk.
//
// Draw assignment (shared with the CPU oracle, so both see the same numbers):
//   counter = (block, tag, packet_id_lo, packet_id_hi), key = (seed, stream)
//   tag 0, block i : the four uniforms of hop/interaction i of the packet
//                    (.x step length, .y roulette, .z HG cos(theta), .w azimuth)
//   tag 1, block j : Fresnel draws 4j..4j+3 (one per refractive-index change)
//   tag 2, block 0 : launch position of the flat-disc / Gaussian beams
// A packet's whole history is a pure function of (seed, stream, packet id), which is what
// the adjoint's path replay and the batched/multi-call sharding rely on.
#pragma once

#include <hip/hip_runtime.h>
#include <stdint.h>

namespace pmo {

constexpr uint32_t kPhiloxM0 = 0xD2511F53u;
constexpr uint32_t kPhiloxM1 = 0xCD9E8D57u;
constexpr uint32_t kPhiloxW0 = 0x9E3779B9u;
constexpr uint32_t kPhiloxW1 = 0xBB67AE85u;

enum : uint32_t { kTagWalk = 0u, kTagFresnel = 1u, kTagSource = 2u };

__host__ __device__ __forceinline__ uint4 philox4x32_10(uint4 c, uint32_t k0, uint32_t k1)
{
#pragma unroll
    for (int r = 0; r < 10; ++r) {
        const uint64_t m0 = (uint64_t)kPhiloxM0 * c.x;
        const uint64_t m1 = (uint64_t)kPhiloxM1 * c.z;
        c = make_uint4((uint32_t)(m1 >> 32) ^ c.y ^ k0, (uint32_t)m1,
                       (uint32_t)(m0 >> 32) ^ c.w ^ k1, (uint32_t)m0);
        k0 += kPhiloxW0;
        k1 += kPhiloxW1;
    }
    return c;
}

// The same block for a key held in scalar registers (uniform across the wave): the round
// keys are stepped in place, one scalar add each, so the compiler cannot precompute all
// twenty of them into registers that, in a kernel short of scalar registers, it would spill
// to vector-register lanes and read back with a VALU v_readlane per round.
__device__ __forceinline__ uint4 philox4x32_10_sk(uint4 c, uint32_t k0, uint32_t k1)
{
#pragma unroll
    for (int r = 0; r < 10; ++r) {
        const uint64_t m0 = (uint64_t)kPhiloxM0 * c.x;
        const uint64_t m1 = (uint64_t)kPhiloxM1 * c.z;
        c = make_uint4((uint32_t)(m1 >> 32) ^ c.y ^ k0, (uint32_t)m1,
                       (uint32_t)(m0 >> 32) ^ c.w ^ k1, (uint32_t)m0);
        k0 += kPhiloxW0;
        k1 += kPhiloxW1;
        asm volatile("" : "+s"(k0), "+s"(k1));
    }
    return c;
}

// A packet's blocks of one tag, counter (block, tag, plo, phi): round 1's x and y words and
// round 2's product M0 * x depend only on the tag, plo and the key, not on the block index.
// philox_packet computes them once per packet; philox_packet_block then gives block `blk`
// for 16 of the 18 64-bit multiply-adds (the quarter-rate instructions of a block), equal
// to philox4x32_10(make_uint4(blk, tag, plo, phi), k0, k1) bit for bit.
struct PhiloxPacket {
    uint32_t x1, y1;   // round 1: hi(M1 plo) ^ tag ^ k0, lo(M1 plo)
    uint32_t h2, w2;   // round 2: hi(M0 x1), lo(M0 x1)
};

__host__ __device__ __forceinline__ PhiloxPacket philox_packet(uint32_t tag, uint32_t plo, uint32_t k0)
{
    const uint64_t m1 = (uint64_t)kPhiloxM1 * plo;
    PhiloxPacket q;
    q.x1 = (uint32_t)(m1 >> 32) ^ tag ^ k0;
    q.y1 = (uint32_t)m1;
    const uint64_t m0 = (uint64_t)kPhiloxM0 * q.x1;
    q.h2 = (uint32_t)(m0 >> 32);
    q.w2 = (uint32_t)m0;
    return q;
}

__host__ __device__ __forceinline__ uint4 philox_packet_block(const PhiloxPacket& q, uint32_t blk, uint32_t phi,
                                                              uint32_t k0, uint32_t k1)
{
    // round 1: only the product of the block index is new
    const uint64_t a = (uint64_t)kPhiloxM0 * blk;
    const uint32_t z1 = (uint32_t)(a >> 32) ^ phi ^ k1;
    const uint32_t w1 = (uint32_t)a;
    k0 += kPhiloxW0;
    k1 += kPhiloxW1;
    // round 2: M0 x1 comes from the packet
    const uint64_t b = (uint64_t)kPhiloxM1 * z1;
    uint4 c = make_uint4((uint32_t)(b >> 32) ^ q.y1 ^ k0, (uint32_t)b, q.h2 ^ w1 ^ k1, q.w2);
    k0 += kPhiloxW0;
    k1 += kPhiloxW1;
#pragma unroll
    for (int r = 2; r < 10; ++r) {
        const uint64_t m0 = (uint64_t)kPhiloxM0 * c.x;
        const uint64_t m1 = (uint64_t)kPhiloxM1 * c.z;
        c = make_uint4((uint32_t)(m1 >> 32) ^ c.y ^ k0, (uint32_t)m1,
                       (uint32_t)(m0 >> 32) ^ c.w ^ k1, (uint32_t)m0);
        k0 += kPhiloxW0;
        k1 += kPhiloxW1;
    }
    return c;
}

// philox_packet_block for keys held in scalar registers, round keys stepped in place as in
// philox4x32_10_sk
__device__ __forceinline__ uint4 philox_packet_block_sk(const PhiloxPacket& q, uint32_t blk, uint32_t phi,
                                                        uint32_t k0, uint32_t k1)
{
    const uint64_t a = (uint64_t)kPhiloxM0 * blk;
    const uint32_t z1 = (uint32_t)(a >> 32) ^ phi ^ k1;
    const uint32_t w1 = (uint32_t)a;
    k0 += kPhiloxW0;
    k1 += kPhiloxW1;
    asm volatile("" : "+s"(k0), "+s"(k1));
    const uint64_t b = (uint64_t)kPhiloxM1 * z1;
    uint4 c = make_uint4((uint32_t)(b >> 32) ^ q.y1 ^ k0, (uint32_t)b, q.h2 ^ w1 ^ k1, q.w2);
    k0 += kPhiloxW0;
    k1 += kPhiloxW1;
    asm volatile("" : "+s"(k0), "+s"(k1));
#pragma unroll
    for (int r = 2; r < 10; ++r) {
        const uint64_t m0 = (uint64_t)kPhiloxM0 * c.x;
        const uint64_t m1 = (uint64_t)kPhiloxM1 * c.z;
        c = make_uint4((uint32_t)(m1 >> 32) ^ c.y ^ k0, (uint32_t)m1,
                       (uint32_t)(m0 >> 32) ^ c.w ^ k1, (uint32_t)m0);
        k0 += kPhiloxW0;
        k1 += kPhiloxW1;
        asm volatile("" : "+s"(k0), "+s"(k1));
    }
    return c;
}

// uniform on (0, 1]: top 24 bits plus one, times 2^-24; exact in float and double, never 0
// so -log(u) is finite (the reference's curand_uniform range, _cukernel.py:75).
// Written as k 2^-24 + 2^-24 with k = x >> 8 < 2^24: both steps are exact, so the value is
// (k + 1) 2^-24 whether or not they are fused, in one convert and one FMA.
template <typename R>
__host__ __device__ __forceinline__ R u01(uint32_t x)
{
    const R s = (R)(1.0 / 16777216.0);
    return (R)(x >> 8) * s + s;
}

}  // namespace pmo
