// Host check of the per-packet Philox split (pmo::philox_packet + philox_packet_block) against
// the plain Philox4x32-10 of the same counter and key, on pseudo-random counters and keys and
// on the edge words 0 and 2^32-1. Prints the number of mismatching blocks; exit status 1 if any.
#include <cstdio>
#include <cstdint>

#include "philox.hpp"

static uint64_t s = 0x9E3779B97F4A7C15ull;
static uint32_t next32()
{
    s ^= s << 13; s ^= s >> 7; s ^= s << 17;
    return (uint32_t)(s >> 16);
}

int main()
{
    const uint32_t edge[4] = {0u, 1u, 0x7FFFFFFFu, 0xFFFFFFFFu};
    long bad = 0, n = 0;
    for (int i = 0; i < 2000000; ++i) {
        uint32_t w[6];
        for (int j = 0; j < 6; ++j) w[j] = (i < 4096) ? edge[(i >> (2 * j)) & 3] : next32();
        const uint32_t blk = w[0], tag = w[1], plo = w[2], phi = w[3], k0 = w[4], k1 = w[5];
        const pmo::PhiloxPacket q = pmo::philox_packet(tag, plo, k0);
        const uint4 a = pmo::philox_packet_block(q, blk, phi, k0, k1);
        const uint4 b = pmo::philox4x32_10(make_uint4(blk, tag, plo, phi), k0, k1);
        bad += (a.x != b.x || a.y != b.y || a.z != b.z || a.w != b.w);
        ++n;
    }
    printf("%ld of %ld blocks differ\n", bad, n);
    return bad != 0;
}
