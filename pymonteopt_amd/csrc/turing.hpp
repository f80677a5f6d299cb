// Turing-pattern trabecular bone model on gfx950 (pymopt/modeling_gpu/_classes.py:25-176).
//
// One step of the reaction-diffusion system (_calc_onestep, :112-115):
//   1. reaction, in place: u <- f(u, v) = u + dt (a1 u + a2 v + a3 u^3),
//                          v <- g(u, v) = v + dt (b1 u + b2 v)              (:51-52)
//   2. U = F[u], V = F[v]                                       (hipFFT R2C)
//   3. U <- Re U / lambda_u(k) / n^3, V likewise                (spectrum kernel)
//   4. u = F^-1[U], v = F^-1[V]                                 (hipFFT C2R)
//
// lambda(k) = 1 + 2 c (3 - cos a - cos b - cos g), with c = dt d / dx^2, is the symbol of
// the reference's implicit 7-point diffusion stencil (:83-104). It is evaluated from a
// cosine table, where the reference stores 1 / Re F[stencil]. Keeping only Re U is the
// reference's `cp.real(cp.fft.fftn(...))`: each step acts on the even part of the
// reaction term.
#pragma once

#include <hip/hip_runtime.h>
#include <hipfft/hipfft.h>
#include <stdint.h>

namespace pmo {

struct TuringCoef {
    float dt, a1, a2, a3, b1, b2;
};

__global__ void __launch_bounds__(256) turing_reaction_kernel(float* __restrict__ u, float* __restrict__ v,
                                                              unsigned long long n, TuringCoef c)
{
    const unsigned long long step = (unsigned long long)gridDim.x * blockDim.x;
    for (unsigned long long i = (unsigned long long)blockIdx.x * blockDim.x + threadIdx.x; i < n; i += step) {
        const float x = u[i], y = v[i];
        u[i] = x + c.dt * (c.a1 * x + c.a2 * y + c.a3 * x * x * x);
        v[i] = y + c.dt * (c.b1 * x + c.b2 * y);
    }
}

// U, V: [n][n][n/2+1] complex; cosine table of 2 pi i / n staged in LDS per block
__global__ void __launch_bounds__(256) turing_spectrum_kernel(float2* __restrict__ U, float2* __restrict__ V, int n,
                                                              float cu, float cv, float scale)
{
    extern __shared__ float cos_tab[];
    for (int i = threadIdx.x; i < n; i += blockDim.x) cos_tab[i] = (float)cospi(2.0 * (double)i / (double)n);
    __syncthreads();
    const int nh = n / 2 + 1;
    const unsigned long long total = (unsigned long long)n * n * nh;
    const unsigned long long step = (unsigned long long)gridDim.x * blockDim.x;
    for (unsigned long long e = (unsigned long long)blockIdx.x * blockDim.x + threadIdx.x; e < total; e += step) {
        const int k = (int)(e % nh);
        const unsigned long long t = e / nh;
        const int j = (int)(t % n), i = (int)(t / n);
        const float s = 3.f - cos_tab[i] - cos_tab[j] - cos_tab[k];
        const float lu = 1.f + 2.f * cu * s, lv = 1.f + 2.f * cv * s;
        U[e] = make_float2(U[e].x * scale / lu, 0.f);
        V[e] = make_float2(V[e].x * scale / lv, 0.f);
    }
}

// u <= thr -> 0 (TuringPattern._adjust_vbtv, :134-138); counts the zeroed cells
__global__ void __launch_bounds__(256) threshold_kernel(float* u, unsigned long long n, float thr,
                                                        unsigned long long* n_below)
{
    unsigned long long c = 0;
    const unsigned long long step = (unsigned long long)gridDim.x * blockDim.x;
    for (unsigned long long i = (unsigned long long)blockIdx.x * blockDim.x + threadIdx.x; i < n; i += step) {
        if (u[i] <= thr) {
            u[i] = 0.f;
            ++c;
        }
    }
    // wave sum, one atomic per wave
    for (int off = 32; off > 0; off >>= 1) c += __shfl_down(c, off);
    if ((threadIdx.x & 63u) == 0u && c) atomicAdd(n_below, c);
}

// strict local maxima along the last axis of rows of length len against `order` neighbours
// each side, out-of-range neighbours clipped to the row's edge (scipy argrelmax, mode
// 'clip'; TuringPattern._calc_microarchitecture, :140-160)
__global__ void __launch_bounds__(256) count_peaks_kernel(const float* u, unsigned long long rows, int len, int order,
                                                          unsigned long long* count)
{
    unsigned long long c = 0;
    const unsigned long long total = rows * (unsigned long long)len;
    const unsigned long long step = (unsigned long long)gridDim.x * blockDim.x;
    for (unsigned long long e = (unsigned long long)blockIdx.x * blockDim.x + threadIdx.x; e < total; e += step) {
        const int x = (int)(e % (unsigned long long)len);
        const float* row = u + (e - (unsigned long long)x);
        const float m = row[x];
        bool peak = true;
        for (int s = 1; s <= order && peak; ++s) {
            const int l = x - s < 0 ? 0 : x - s, r = x + s > len - 1 ? len - 1 : x + s;
            peak = m > row[l] && m > row[r];
        }
        c += peak ? 1ull : 0ull;
    }
    for (int off = 32; off > 0; off >>= 1) c += __shfl_down(c, off);
    if ((threadIdx.x & 63u) == 0u && c) atomicAdd(count, c);
}

// u > 0 -> 1, else 0 (TuringPattern._model_binarization, :162-165)
__global__ void __launch_bounds__(256) binarize_kernel(const float* u, unsigned long long n, int8_t* labels)
{
    const unsigned long long step = (unsigned long long)gridDim.x * blockDim.x;
    for (unsigned long long i = (unsigned long long)blockIdx.x * blockDim.x + threadIdx.x; i < n; i += step)
        labels[i] = u[i] > 0.f ? (int8_t)1 : (int8_t)0;
}

}  // namespace pmo
