// OBD receiver on gfx950: the packets a walk sends back out of the sample, traced through
// the receiver optics for every focus offset Z of a scan (pymopt/optics/_classes.py:381-567).
//
// The receiver is:
//   lens 1: plano-convex, plano face toward the sample (Lens1, :295-303);
//   slit 1: annular, front and back face (Slit, :50-85);
//   the LD mount, optional (LdFixingPart, :336-373);
//   slit 2;
//   lens 2: plano-convex, convex face toward the sample (Lens2, :305-313);
//   the photodiode disc (Photodiode, :316-334).
//
// One lane takes one exit packet and loops over the focus steps. The planes and sphere
// centres of each step are computed on the host in FP64 and staged in LDS. The weight
// reaching the photodiode at step s accumulates in an FP64 LDS bin, flushed once per block.
//
// Fresnel at the four lens faces follows :152-188. The reference keeps a ray when a fresh
// uniform exceeds R; here the uniform of face f at step s for packet j is word f of
// Philox4x32-10 with counter (s, 3, j_lo, j_hi). With `expect` the weight is multiplied by
// the transmission 1 - R instead, which is the cut's mean.
#pragma once

#include "walk.hpp"

namespace pmo {

enum : uint32_t { kTagObd = 3u };

struct ObdStep {           // layout of one row of pmo_obd_trace's `steps` (include/pmo.h)
    float z_lens1, c_lens1, z_slit1, z_lfp, z_slit2, z_lens2, c_lens2, z_pd, x_shift, pad0, pad1, pad2;
};

struct ObdParams {
    const float* p;        // [3][stride] x, y, z of the exit states
    const float* v;        // [3][stride]
    const float* w;        // [n]
    unsigned long long n, stride;
    const ObdStep* steps;
    int32_t n_steps;
    float r1, n1, outer1_sq, r2, n2, outer2_sq;
    float d_in, d_out, slit_th, pd_r_sq;
    int32_t ld_fix;
    float lfp_a1, lfp_b1, lfp_a2, lfp_b2, lfp_b3, lfp_th;
    int32_t inversion;
    float side;
    int32_t expect;
    uint32_t seed, stream;
    double* out;           // [n_steps]
};

// Vector refraction across a face with unit normal nn along the propagation
// (_classes.py:152-188): v' = a v + (cos_t - a cos_i) nn, a = ni/nt. Returns false on total
// internal reflection. R is the unpolarised Fresnel reflectance in cosine form, equal to
// the reference's 0.5 [sin^2(ai-at)/sin^2(ai+at) + tan^2(ai-at)/tan^2(ai+at)].
__device__ __forceinline__ bool obd_refract(float& vx, float& vy, float& vz, float nx, float ny, float nz, float ni,
                                            float nt, float& R)
{
    const float a = ni / nt;
    const float ci = nx * vx + ny * vy + nz * vz;
    const float s2 = a * a * (1.f - ci * ci);
    if (ni > nt && s2 >= 1.f) return false;
    const float ct = __builtin_copysignf(__builtin_sqrtf(__builtin_fmaxf(1.f - s2, 0.f)), ci);
    const float aci = __builtin_fabsf(ci), act = __builtin_fabsf(ct);
    const float rs = (ni * aci - nt * act) / (ni * aci + nt * act);
    const float rp = (ni * act - nt * aci) / (ni * act + nt * aci);
    R = 0.5f * (rs * rs + rp * rp);
    const float g = ct - a * ci;
    vx = a * vx + g * nx;
    vy = a * vy + g * ny;
    vz = a * vz + g * nz;
    const float inv = __builtin_amdgcn_rsqf(vx * vx + vy * vy + vz * vz);
    vx *= inv; vy *= inv; vz *= inv;
    return true;
}

// Fresnel cut of a face: the ray goes on if the uniform exceeds R (stochastic), or its
// weight is scaled by 1 - R (expect).
__device__ __forceinline__ bool obd_pass(bool expect, float u, float R, float& wt)
{
    if (expect) {
        wt *= 1.f - R;
        return true;
    }
    return u > R;
}

// Move along v to the plane z = zp (_classes.py:64-67).
__device__ __forceinline__ void obd_to_plane(float& x, float& y, float& z, float vx, float vy, float vz, float zp)
{
    const float a = (zp - z) / vz;
    x += a * vx;
    y += a * vy;
    z = zp;
}

// Entry point on lens 1's plano face z = z_lens1 from the sample plane z = 0, with
// ivz = 1 / vz precomputed per packet (the scan moves this plane only).
__device__ __forceinline__ void obd_lens1_entry(const ObdStep& S, float& x, float& y, float vx, float vy, float ivz)
{
    const float a = S.z_lens1 * ivz;
    x += a * vx;
    y += a * vy;
}

// Weight reaching the photodiode at focus step s (0 if the ray is lost) of exit packet j,
// whose direction (vx, vy, vz) is a unit vector. The Fresnel uniforms are drawn only for
// rays that pass lens 1's aperture, where most rays end; the draws depend on (s, j) alone,
// so the order changes nothing.
__device__ float obd_trace(const ObdParams& P, const ObdStep& S, float x, float y, float vx, float vy, float vz,
                           float ivz, float w, int32_t s, unsigned long long j)
{
    const bool ex = P.expect != 0;
    float z = S.z_lens1, R = 0.f, wt = w;
    // lens 1, plano face (air -> glass)
    obd_lens1_entry(S, x, y, vx, vy, ivz);
    if (!(x * x + y * y <= P.outer1_sq)) return 0.f;
    float u[4] = {0.f, 0.f, 0.f, 0.f};
    if (!ex) {
        const uint4 r = philox4x32_10(make_uint4((uint32_t)s, kTagObd, (uint32_t)j, (uint32_t)(j >> 32)), P.seed,
                                      P.stream);
        u[0] = u01<float>(r.x); u[1] = u01<float>(r.y); u[2] = u01<float>(r.z); u[3] = u01<float>(r.w);
    }
    if (!obd_refract(vx, vy, vz, 0.f, 0.f, -1.f, 1.f, P.n1, R) || !obd_pass(ex, u[0], R, wt)) return 0.f;
    // lens 1, convex face (glass -> air): far side of the sphere about (0, 0, c_lens1)
    {
        const float pz = z - S.c_lens1;
        const float b = x * vx + y * vy + pz * vz;
        const float c = x * x + y * y + pz * pz - P.r1 * P.r1;
        const float disc = b * b - c;
        if (!(disc >= 0.f)) return 0.f;
        const float t = -b + __builtin_sqrtf(disc);
        x += t * vx; y += t * vy; z += t * vz;
        const float rr = x * x + y * y;
        if (!(rr <= P.outer1_sq)) return 0.f;
        const float d = __builtin_sqrtf(P.r1 * P.r1 - rr);
        float nx = x / d, ny = y / d, nz = -1.f;
        const float inv = __builtin_amdgcn_rsqf(nx * nx + ny * ny + 1.f);
        nx *= inv; ny *= inv; nz *= inv;
        if (!obd_refract(vx, vy, vz, nx, ny, nz, P.n1, 1.f, R) || !obd_pass(ex, u[1], R, wt)) return 0.f;
    }
    // slit 1, LD mount, slit 2: front and back face of each
    const float din2 = P.d_in > 0.f ? P.d_in * P.d_in : -1.f, dout2 = P.d_out * P.d_out;
#pragma unroll
    for (int e = 0; e < 6; ++e) {
        const bool lfp = (e == 2 || e == 3);
        if (lfp && !P.ld_fix) continue;
        float zp;
        if (e == 0) zp = S.z_slit1;
        else if (e == 1) zp = S.z_slit1 - P.slit_th;
        else if (e == 2) zp = S.z_lfp;
        else if (e == 3) zp = S.z_lfp - P.lfp_th;
        else if (e == 4) zp = S.z_slit2;
        else zp = S.z_slit2 - P.slit_th;
        obd_to_plane(x, y, z, vx, vy, vz, zp);
        if (lfp) {
            if (!((y > P.lfp_a1 * x + P.lfp_b1) || (y < P.lfp_a2 * x + P.lfp_b2) || (y < P.lfp_b3))) return 0.f;
        } else {
            const float rr = x * x + y * y;
            if (!(rr < dout2 && rr > din2)) return 0.f;
        }
    }
    // lens 2, convex face (air -> glass): near side of the sphere about (0, 0, c_lens2)
    {
        const float pz = z - S.c_lens2;
        const float b = x * vx + y * vy + pz * vz;
        const float c = x * x + y * y + pz * pz - P.r2 * P.r2;
        const float disc = b * b - c;
        if (!(disc >= 0.f)) return 0.f;
        const float t = -b - __builtin_sqrtf(disc);
        x += t * vx; y += t * vy; z += t * vz;
        const float rr = x * x + y * y;
        if (!(rr <= P.outer2_sq)) return 0.f;
        const float d = __builtin_sqrtf(P.r2 * P.r2 - rr);
        float nx = -x / d, ny = -y / d, nz = -1.f;
        const float inv = __builtin_amdgcn_rsqf(nx * nx + ny * ny + 1.f);
        nx *= inv; ny *= inv; nz *= inv;
        if (!obd_refract(vx, vy, vz, nx, ny, nz, 1.f, P.n2, R) || !obd_pass(ex, u[2], R, wt)) return 0.f;
    }
    // lens 2, plano face (glass -> air)
    obd_to_plane(x, y, z, vx, vy, vz, S.z_lens2);
    if (!(x * x + y * y <= P.outer2_sq)) return 0.f;
    if (!obd_refract(vx, vy, vz, 0.f, 0.f, -1.f, P.n2, 1.f, R) || !obd_pass(ex, u[3], R, wt)) return 0.f;
    // photodiode
    obd_to_plane(x, y, z, vx, vy, vz, S.z_pd);
    return (x * x + y * y < P.pd_r_sq) ? wt : 0.f;
}

// A step's lens-1 aperture test: the first test of obd_trace, computed the same way.
__device__ __forceinline__ bool obd_enters(const ObdParams& P, const ObdStep& S, float x, float y, float vx, float vy,
                                           float ivz)
{
    obd_lens1_entry(S, x, y, vx, vy, ivz);
    return x * x + y * y <= P.outer1_sq;
}

constexpr int kObdUnroll = 4;    // focus steps tested per pass (independent, for latency hiding)
constexpr int kObdQueue = 512;   // per-wave queue of (lane, step) pairs: < 64 + 64 * kObdUnroll held

// One lane holds one exit packet for the whole scan. Each step, the wave tests all its
// packets against lens 1's aperture, where about 93% of traces end, and appends the
// (lane, step) pairs that enter to a per-wave LDS queue. Whenever the queue holds 64
// pairs, every lane traces one of them through the rest of the receiver, reading the
// packet's state from its owner lane. Full traces thus run with the wave full, instead of
// once per step for a handful of lanes. Each (packet, step) is traced by the same
// arithmetic as before; only the summation order of I(Z) changes.
__global__ void __launch_bounds__(256) obd_kernel(const ObdParams P)
{
    extern __shared__ __align__(16) unsigned char smem[];
    ObdStep* steps = reinterpret_cast<ObdStep*>(smem);
    double* hist = reinterpret_cast<double*>(smem + (size_t)P.n_steps * sizeof(ObdStep));
    uint16_t* queues = reinterpret_cast<uint16_t*>(smem + (size_t)P.n_steps * (sizeof(ObdStep) + sizeof(double)));
    for (int32_t i = threadIdx.x; i < P.n_steps; i += blockDim.x) {
        steps[i] = P.steps[i];
        hist[i] = 0.0;
    }
    __syncthreads();

    const uint32_t lane = threadIdx.x & 63u;
    const unsigned long long lt_mask = (lane == 0) ? 0ull : (~0ull >> (64u - lane));
    uint16_t* q = queues + (threadIdx.x >> 6) * kObdQueue;
    const unsigned long long stride = (unsigned long long)gridDim.x * blockDim.x;
    // the wave's first packet; the loop bound is wave-uniform, so ballots see every lane
    for (unsigned long long j0 = (unsigned long long)blockIdx.x * blockDim.x + (threadIdx.x & ~63u); j0 < P.n;
         j0 += stride) {
        const unsigned long long j = j0 + lane;
        bool valid = j < P.n;
        float x = 0.f, y = 0.f, vx = 0.f, vy = 0.f, vz = 1.f, w = 0.f;
        if (valid) {
            x = P.p[j]; y = P.p[P.stride + j];
            const float z = P.p[2 * P.stride + j];
            vx = P.v[j]; vy = P.v[P.stride + j]; vz = P.v[2 * P.stride + j];
            w = P.w[j];
            if (P.side != 0.f) {
                // side view (_classes.py:504-508): axes x and z exchanged, shifted by half the side
                x = z - 0.5f * P.side;
                const float tv = vx; vx = vz; vz = tv;
            }
            // the trace starts on the sample plane, p_z = 0 (:515)
            valid = P.inversion ? vz > 0.f : vz < 0.f;
            if (P.inversion) { y = -y; vz = -vz; vy = -vy; }
            const float inv = __builtin_amdgcn_rsqf(vx * vx + vy * vy + vz * vz);
            vx *= inv; vy *= inv; vz *= inv;
        }
        const float ivz = 1.f / vz;
        uint32_t head = 0, tail = 0;   // wave-uniform queue positions
        for (int32_t s0 = 0; s0 <= P.n_steps; s0 += kObdUnroll) {
            if (s0 < P.n_steps) {
                bool enters[kObdUnroll];
#pragma unroll
                for (int u = 0; u < kObdUnroll; ++u) {
                    const int32_t s = s0 + u;
                    enters[u] = false;
                    if (valid && s < P.n_steps) {
                        const ObdStep& S = steps[s];
                        enters[u] = obd_enters(P, S, x + S.x_shift, y, vx, vy, ivz);
                    }
                }
#pragma unroll
                for (int u = 0; u < kObdUnroll; ++u) {
                    const unsigned long long m = __ballot(enters[u]);
                    if (enters[u])
                        q[(tail + (uint32_t)__popcll(m & lt_mask)) & (kObdQueue - 1)] =
                            (uint16_t)(((s0 + u) << 6) | (int32_t)lane);
                    tail += (uint32_t)__popcll(m);
                }
            }
            __builtin_amdgcn_fence(__ATOMIC_RELEASE, "wavefront");
            __builtin_amdgcn_wave_barrier();
            __builtin_amdgcn_fence(__ATOMIC_ACQUIRE, "wavefront");
            // trace 64 queued pairs at a time; after the last step, the remainder
            const bool last = s0 + kObdUnroll > P.n_steps;
            while (tail - head >= 64u || (last && tail != head)) {
                const uint32_t take = tail - head < 64u ? tail - head : 64u;
                const bool mine = lane < take;
                const uint32_t e = mine ? (uint32_t)q[(head + lane) & (kObdQueue - 1)] : 0u;
                const int src = (int)(e & 63u);
                const int32_t es = (int32_t)(e >> 6);
                const float tx = __shfl(x, src), ty = __shfl(y, src);
                const float tvx = __shfl(vx, src), tvy = __shfl(vy, src), tvz = __shfl(vz, src);
                const float tivz = __shfl(ivz, src), tw = __shfl(w, src);
                if (mine) {
                    const ObdStep& S = steps[es];
                    const float d = obd_trace(P, S, tx + S.x_shift, ty, tvx, tvy, tvz, tivz, tw, es, j0 + (uint32_t)src);
                    if (d > 0.f) atomicAdd(&hist[es], (double)d);
                }
                head += take;
                __builtin_amdgcn_fence(__ATOMIC_RELEASE, "wavefront");
                __builtin_amdgcn_wave_barrier();
                __builtin_amdgcn_fence(__ATOMIC_ACQUIRE, "wavefront");
            }
            if (last) break;
        }
    }
    __syncthreads();
    for (int32_t i = threadIdx.x; i < P.n_steps; i += blockDim.x)
        if (hist[i] != 0.0) gmem_add(P.out + i, hist[i]);
}

}  // namespace pmo
