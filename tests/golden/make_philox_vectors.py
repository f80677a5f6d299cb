"""Writes tests/golden/philox_vectors.json: Philox4x32-10 outputs computed by rocRAND's
philox4x32_10_engine (header-only source under /opt/rocm/include/rocrand), compiled here
for the host. An implementation written independently of the walk's and the oracle's.

rocRAND's engine(seed, subsequence, offset) starts at counter
(offset/4 lo, offset/4 hi, subsequence lo, subsequence hi) with key (seed lo, seed hi), so
any (counter, key) pair with counter.y < 2^30 maps onto it (offset = 4 * (y:x) must fit in
64 bits). The three Random123 known-answer vectors (kat_vectors of Random123 1.x for
philox4x32_10) are recorded separately as published values.
"""
import json
import os
import random
import subprocess
import tempfile

HERE = os.path.dirname(os.path.abspath(__file__))

SRC = r"""
#include <rocrand/rocrand_philox4x32_10.h>
#include <cstdio>
#include <cstdlib>
int main(int argc, char** argv) {
    unsigned c[4], k[2];
    while (scanf("%x %x %x %x %x %x", &c[0], &c[1], &c[2], &c[3], &k[0], &k[1]) == 6) {
        unsigned long long seed = (unsigned long long)k[0] | ((unsigned long long)k[1] << 32);
        unsigned long long sub = (unsigned long long)c[2] | ((unsigned long long)c[3] << 32);
        unsigned long long off = 4ull * ((unsigned long long)c[0] | ((unsigned long long)c[1] << 32));
        rocrand_device::philox4x32_10_engine e(seed, sub, off);
        uint4 r = e.next4();
        printf("%08x %08x %08x %08x\n", r.x, r.y, r.z, r.w);
    }
    return 0;
}
"""


def main():
    rng = random.Random(20241203)
    cases = [((0, 0, 0, 0), (0, 0))]
    # the walk's counter shapes: (block, tag, packet lo, packet hi), key (seed, stream)
    for tag in (0, 1, 2):
        for _ in range(8):
            cases.append(((rng.randrange(0, 5000), tag, rng.randrange(0, 2 ** 32), rng.randrange(0, 4)),
                          (rng.randrange(0, 100000), rng.randrange(0, 600))))
    for _ in range(16):
        c = [rng.randrange(0, 2 ** 32) for _ in range(4)]
        c[1] = rng.randrange(0, 2 ** 30)
        cases.append((tuple(c), tuple(rng.randrange(0, 2 ** 32) for _ in range(2))))
    with tempfile.TemporaryDirectory() as td:
        src = os.path.join(td, "p.cpp")
        exe = os.path.join(td, "p")
        open(src, "w").write(SRC)
        subprocess.run(["/opt/rocm/bin/hipcc", "-O1", src, "-o", exe], check=True)
        inp = "".join("%x %x %x %x %x %x\n" % (*c, *k) for c, k in cases)
        out = subprocess.run([exe], input=inp, capture_output=True, text=True, check=True).stdout.split("\n")
    vec = []
    for (c, k), line in zip(cases, out):
        vec.append({"ctr": list(c), "key": list(k), "out": [int(x, 16) for x in line.split()]})
    kat = [{"ctr": [0, 0, 0, 0], "key": [0, 0], "out": [0x6627e8d5, 0xe169c58d, 0xbc57ac4c, 0x9b00dbd8]},
           {"ctr": [0xFFFFFFFF] * 4, "key": [0xFFFFFFFF] * 2,
            "out": [0x408f276d, 0x41c83b0e, 0xa20bc7c6, 0x6d5451fd]},
           {"ctr": [0x243f6a88, 0x85a308d3, 0x13198a2e, 0x03707344], "key": [0xa4093822, 0x299f31d0],
            "out": [0xd16cfe09, 0x94fdcceb, 0x5001e420, 0x24126ea1]}]
    json.dump({"generator": "rocrand_device::philox4x32_10_engine (host build)", "vectors": vec,
               "random123_kat": kat},
              open(os.path.join(HERE, "philox_vectors.json"), "w"), indent=1)
    print("wrote", len(vec), "vectors")


if __name__ == "__main__":
    main()
