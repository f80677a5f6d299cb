/*
 * oracle.c -- TEST INFRASTRUCTURE ONLY (see oracle.h). CPU restatement of pyMonteOpt's
 * voxel walk, pymopt/voxel/_kernel.py:23-172, built twice (double / float state).
 */
#define _GNU_SOURCE
#include "oracle.h"

#include <math.h>
#include <omp.h>
#include <stdlib.h>
#include <string.h>

/* Philox4x32-10 (Salmon, Moraes, Dror, Shaw, "Parallel random numbers: as easy as 1, 2, 3",
   SC'11): 10 rounds of (hi, lo) = M * c with the Weyl key schedule. */
void og_philox4x32_10(const uint32_t ctr[4], const uint32_t key[2], uint32_t out[4])
{
    uint32_t c0 = ctr[0], c1 = ctr[1], c2 = ctr[2], c3 = ctr[3];
    uint32_t k0 = key[0], k1 = key[1];
    for (int r = 0; r < 10; r++) {
        uint64_t m0 = (uint64_t)0xD2511F53u * c0;
        uint64_t m1 = (uint64_t)0xCD9E8D57u * c2;
        uint32_t n0 = (uint32_t)(m1 >> 32) ^ c1 ^ k0;
        uint32_t n1 = (uint32_t)m1;
        uint32_t n2 = (uint32_t)(m0 >> 32) ^ c3 ^ k1;
        uint32_t n3 = (uint32_t)m0;
        c0 = n0; c1 = n1; c2 = n2; c3 = n3;
        k0 += 0x9E3779B9u;
        k1 += 0xBB67AE85u;
    }
    out[0] = c0; out[1] = c1; out[2] = c2; out[3] = c3;
}

/* uniform on (0, 1]: the top 24 bits plus one, times 2^-24 (exact in float and double) */
double og_u01(uint32_t x) { return (double)((x >> 8) + 1u) * (1.0 / 16777216.0); }

#define REAL double
#define SFX f64
#include "walk_impl.h"
#undef REAL
#undef SFX

#define REAL float
#define SFX f32
#include "walk_impl.h"
#undef REAL
#undef SFX

int og_trace_f64(const og_volume* vol, const og_source* src, const og_run* run, uint64_t pid,
                 og_event* ev, int32_t cap)
{
    if (!vol || !src || !run || !ev || cap <= 0) return -1;
    return packet_f64(vol, src, run, NULL, NULL, pid, ev, cap);
}

double og_fresnel(double ni, double nt, double abs_vk, double* cos_t)
{
    double at = 0;
    double r = fresnel_lit_f64(ni, nt, abs_vk, &at, 0);
    if (cos_t) *cos_t = cos(at);
    return r;
}

double og_hg_cos(double g, double u) { return hg_f64(g, u); }

void og_spin(double v[3], double cos_th, double phi) { spin_f64(v, cos_th, phi); }

double og_boundary(const double p[3], const double v[3], double half, int32_t* k)
{
    int kk = 0;
    double d = boundary_f64(p, v, half, &kk);
    if (k) *k = kk;
    return d;
}
