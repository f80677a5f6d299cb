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

// Weight reaching the photodiode at one focus step (0 if the ray is lost).
__device__ float obd_trace(const ObdParams& P, const ObdStep& S, float x, float y, float vx, float vy, float vz,
                           float w, const float u[4])
{
    const bool ex = P.expect != 0;
    float z = 0.f, R = 0.f, wt = w;
    {
        const float inv = __builtin_amdgcn_rsqf(vx * vx + vy * vy + vz * vz);
        vx *= inv; vy *= inv; vz *= inv;
    }
    // lens 1, plano face (air -> glass)
    obd_to_plane(x, y, z, vx, vy, vz, S.z_lens1);
    if (!(x * x + y * y <= P.outer1_sq)) return 0.f;
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

__global__ void __launch_bounds__(256) obd_kernel(const ObdParams P)
{
    extern __shared__ __align__(16) unsigned char smem[];
    ObdStep* steps = reinterpret_cast<ObdStep*>(smem);
    double* hist = reinterpret_cast<double*>(smem + (size_t)P.n_steps * sizeof(ObdStep));
    for (int32_t i = threadIdx.x; i < P.n_steps; i += blockDim.x) {
        steps[i] = P.steps[i];
        hist[i] = 0.0;
    }
    __syncthreads();

    const unsigned long long stride = (unsigned long long)gridDim.x * blockDim.x;
    for (unsigned long long j = (unsigned long long)blockIdx.x * blockDim.x + threadIdx.x; j < P.n; j += stride) {
        float x = P.p[j], y = P.p[P.stride + j], z = P.p[2 * P.stride + j];
        float vx = P.v[j], vy = P.v[P.stride + j], vz = P.v[2 * P.stride + j];
        const float w = P.w[j];
        if (P.side != 0.f) {
            // side view (_classes.py:504-508): axes x and z exchanged, shifted by half the side
            const float t = x; x = z - 0.5f * P.side; z = t + 0.5f * P.side;
            const float tv = vx; vx = vz; vz = tv;
        }
        (void)z;  // the trace starts on the sample plane, p_z = 0 (:515)
        if (!(P.inversion ? vz > 0.f : vz < 0.f)) continue;
        if (P.inversion) { y = -y; vz = -vz; vy = -vy; }
        for (int32_t s = 0; s < P.n_steps; ++s) {
            const ObdStep S = steps[s];
            float u[4] = {0.f, 0.f, 0.f, 0.f};
            if (!P.expect) {
                const uint4 r = philox4x32_10(make_uint4((uint32_t)s, kTagObd, (uint32_t)j, (uint32_t)(j >> 32)),
                                              P.seed, P.stream);
                u[0] = u01<float>(r.x); u[1] = u01<float>(r.y); u[2] = u01<float>(r.z); u[3] = u01<float>(r.w);
            }
            const float d = obd_trace(P, S, x + S.x_shift, y, vx, vy, vz, w, u);
            if (d > 0.f) atomicAdd(&hist[s], (double)d);
        }
    }
    __syncthreads();
    for (int32_t i = threadIdx.x; i < P.n_steps; i += blockDim.x)
        if (hist[i] != 0.0) gmem_add(P.out + i, hist[i]);
}

}  // namespace pmo
