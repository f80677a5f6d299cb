// Device side of the voxel photon walk for gfx950 (CDNA4, wave64).
//
// One packet per lane, persistent grid. Each loop iteration a lane performs one event of
// pymopt/voxel/_kernel.py:60-165: either a voxel-face crossing (with Fresnel test where
// the refractive index changes) or the end of a hop inside a voxel (drop, roulette,
// Henyey-Greenstein spin, next step length). A lane whose packet ended takes the next
// packet id from its wave's batch; a wave claims batches of ids from one global counter.
//
// Optics (per label) live in LDS; labels are read as bytes through L2; absorbed weight
// is coalesced per lane while consecutive deposits hit the same voxel and flushed with
// FP64 atomics; exit tallies go to per-block FP64 LDS histograms flushed once at the end.
#pragma once

#include <hip/hip_runtime.h>
#include <stdint.h>

#include "philox.hpp"

namespace pmo {

// ------------------------------------------------------------------ precision traits
template <typename R> struct Mth;

template <> struct Mth<float> {
    static __device__ __forceinline__ float rcp(float x) { return __builtin_amdgcn_rcpf(x); }
    static __device__ __forceinline__ float sqrt(float x) { return __builtin_amdgcn_sqrtf(x); }
    static __device__ __forceinline__ float rsqrt(float x) { return __builtin_amdgcn_rsqf(x); }
    static __device__ __forceinline__ float div(float a, float b) { return a * __builtin_amdgcn_rcpf(b); }
    // -ln u for u in [2^-24, 1] (never denormal): the hardware log2 times -ln 2
    static __device__ __forceinline__ float neglog(float u) { return __builtin_amdgcn_logf(u) * -0.693147180559945309f; }
    // sin and cos of 2 pi u for u in [0, 1]: the hardware sine and cosine take revolutions
    static __device__ __forceinline__ void sincos2pi(float u, float* s, float* c)
    {
        *s = __builtin_amdgcn_sinf(u);
        *c = __builtin_amdgcn_cosf(u);
    }
    static __device__ __forceinline__ float fabs(float x) { return __builtin_fabsf(x); }
    static __device__ __forceinline__ float copysign(float a, float b) { return __builtin_copysignf(a, b); }
    static __device__ __forceinline__ float fmax(float a, float b) { return __builtin_fmaxf(a, b); }
    static __device__ __forceinline__ float fmin(float a, float b) { return __builtin_fminf(a, b); }
    static __device__ __forceinline__ float floor(float x) { return __builtin_floorf(x); }
    static __device__ __forceinline__ float rint(float x) { return __builtin_rintf(x); }
    // x clamped to [-1, 1] (x is never NaN here): one median-of-three
    static __device__ __forceinline__ float clamp1(float x) { return __builtin_amdgcn_fmed3f(x, -1.0f, 1.0f); }
};

template <> struct Mth<double> {
    static __device__ __forceinline__ double rcp(double x) { return 1.0 / x; }
    static __device__ __forceinline__ double sqrt(double x) { return ::sqrt(x); }
    static __device__ __forceinline__ double rsqrt(double x) { return 1.0 / ::sqrt(x); }
    static __device__ __forceinline__ double div(double a, double b) { return a / b; }
    static __device__ __forceinline__ double neglog(double u) { return -::log(u); }
    static __device__ __forceinline__ void sincos2pi(double u, double* s, double* c)
    {
        ::sincos(6.283185307179586477 * u, s, c);
    }
    static __device__ __forceinline__ double fabs(double x) { return __builtin_fabs(x); }
    static __device__ __forceinline__ double copysign(double a, double b) { return __builtin_copysign(a, b); }
    static __device__ __forceinline__ double fmax(double a, double b) { return __builtin_fmax(a, b); }
    static __device__ __forceinline__ double fmin(double a, double b) { return __builtin_fmin(a, b); }
    static __device__ __forceinline__ double floor(double x) { return __builtin_floor(x); }
    static __device__ __forceinline__ double rint(double x) { return __builtin_rint(x); }
    static __device__ __forceinline__ double clamp1(double x) { return x < -1.0 ? -1.0 : (x > 1.0 ? 1.0 : x); }
};

// per-label constants in LDS, derived once per block from (mu_a, mu_s, g, n)
template <typename R> struct LabelConst {
    R mt;      // mu_a + mu_s   (reference mt, _kernel.py:56)
    R ra;      // mu_a / mt     (drop fraction, _kernel.py:130)
    R n;       // refractive index
    R g;
    R hg1;     // 1 + g^2
    R hg2;     // 1 - g^2
    R hg3;     // 1 - g
    R hg4;     // 2 g
    R hginv;   // 1 / (2 g)
    R ms;      // mu_s
    R inv_ma;  // 1 / mu_a, 0 where mu_a = 0 (r-z fluence weight)
    R pad;
};

// ------------------------------------------------------------------ kernel parameters
struct WalkParams {
    const uint8_t* labels;
    const uint16_t* label_radius;  // label | radius << 8 per voxel (SKIP kernels), or null
    const float4* optics;
    uint32_t nx, ny, nz, nvox;
    uint32_t sx, sy;           // strides of x and y in the C-order volume
    uint32_t n_labels;         // end label
    int32_t cx, cy;            // encoder centre voxel
    float pitch;
    // source
    int32_t src_kind, ix, iy, iz;
    float radius, dirx, diry, dirz, w0;
    // per-packet launch states (src_kind 3): [3][st_stride] add, p, v and [st_stride] w
    const int32_t* st_add;
    const float* st_p;
    const float* st_v;
    const float* st_w;
    unsigned long long st_stride;
    // run
    unsigned long long n_packets, first_packet;
    uint32_t seed, stream;
    float wth, roulette_m, inv_m;
    uint32_t flags;
    uint32_t batch;            // packet ids claimed per wave per counter atomic
    uint32_t max_events;       // per-packet event guard
    uint32_t quorum;           // lanes ready for an interaction before a wave runs that step
    uint32_t refill;           // idle lanes before a wave starts new packets (all idle: always)
    // tallies
    double* scalars;
    unsigned long long* events;
    double* absorbed;
    int32_t nr;
    float dr;
    double* rd_r;
    double* tt_r;
    double* rd_top_r;
    int32_t na;
    float da;
    double* rd_a;
    double* tt_a;
    int32_t nl;
    float dl, line_y;
    double* line_z;
    int32_t n_rings;
    const float* rings;
    double* ring_w;
    float* exit_state;
    int32_t* exit_add;         // [3][n_packets] end voxel address, or null
    float* exit_p;             // [3][n_packets] position in that voxel's frame, or null
    int32_t rz_nr, rz_nz;      // r-z internal fluence bins [i dr, (i+1) dr) x [k dz, (k+1) dz)
    float rz_dr, rz_dz;
    double* fluence_rz;
    unsigned long long* work;
    const uint32_t* key_dev;   // (seed, stream) in device memory, or null
    // LDS-privatised box of the absorbed tally around the launch voxel
    int32_t win_x0, win_y0, win_z0;
    uint32_t win_nx, win_ny, win_nz, win_cells;
    // absorbed weight outside the box, gathered per 2x2x2-voxel cube (one 64 B line of FP64)
    // in a per-lane LDS line buffer and added to `lines` (cube-major, 8 doubles per cube)
    // one line per request; null: one global atomic per voxel deposit
    double* lines;
    uint32_t nly, nlz;         // cubes along y and z
    uint32_t lbuf_stride;      // block size: element e of lane t is at lbuf[e * stride + (t ^ e)]
    // dynamic LDS layout (bytes)
    uint32_t lds_tab, lds_scal, lds_ev, lds_hist, lds_lbuf, lds_stage, lds_win;
    uint32_t h_rd, h_tt, h_top, h_rda, h_tta, h_line, h_ring, h_rz, h_total;  // offsets in doubles
};

// block-level accumulator slots in LDS (doubles)
enum { kBS_RD = 0, kBS_TT, kBS_OTHER, kBS_ABS, kBS_KILL, kBS_GAIN, kBS_W0, kBS_FACE, kBS_COUNT = 13 };
enum { kBE_PACKETS = 0, kBE_INTERACT, kBE_CROSS, kBE_FRESNEL, kBE_ROULETTE, kBE_KILLED, kBE_ESCAPED,
       kBE_GLOBAL_DEPOSITS, kBE_COUNT = 8 };

__device__ __forceinline__ void lds_add(double* p, double v) { atomicAdd(p, v); }
__device__ __forceinline__ void lds_add(unsigned long long* p, unsigned long long v) { atomicAdd(p, v); }

__device__ __forceinline__ void gmem_add(double* p, double v)
{
    __hip_atomic_fetch_add(p, v, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
}

// a pending absorbed-weight deposit lives in one of two places: slot >= 0 is a voxel index
// of the global tally, slot < 0 is -(cell + 1) of the block's LDS window
constexpr int32_t kNoSlot = 0x7fffffff;

__device__ __forceinline__ void flush_deposit(const WalkParams& P, double* win, int32_t slot, double v, uint32_t& n_glob)
{
    if (slot == kNoSlot) return;
    if (slot < 0) {
        lds_add(win + (-slot - 1), v);
    } else {
        gmem_add(P.absorbed + slot, v);
        ++n_glob;
    }
}

__device__ __forceinline__ int bin_of(float x, float width, int nbins)
{
    // bins (i*width, (i+1)*width]; x <= 0 in none (metrics/_scoring.py:22 uses alpha[i] < r)
    if (!(x > 0.f)) return -1;
    const int b = (int)__builtin_ceilf(x / width) - 1;
    return (b >= 0 && b < nbins) ? b : -1;
}

// ------------------------------------------------------------------ shared physics
// These are the per-event formulas of pymopt/voxel/_kernel.py, shared by every walk
// kernel so that a packet with the same draws takes the same path in each of them.

template <typename R>
__device__ __forceinline__ LabelConst<R> label_const(const float4 o)
{
    LabelConst<R> c;
    const R ma = (R)o.x, ms = (R)o.y, g = (R)o.z;
    c.mt = ma + ms;
    c.ra = (c.mt > (R)0) ? ma / c.mt : (R)0;
    c.n = (R)o.w;
    c.g = g;
    c.hg1 = (R)1 + g * g;
    c.hg2 = (R)1 - g * g;
    c.hg3 = (R)1 - g;
    c.hg4 = (R)2 * g;
    c.hginv = (g != (R)0) ? (R)1 / ((R)2 * g) : (R)0;
    c.ms = ms;
    c.inv_ma = (ma > (R)0) ? (R)1 / ma : (R)0;
    c.pad = (R)0;
    return c;
}

// Fresnel reflectance for |v_k| = ci going from index ni into nt (:84-93), in cosine form;
// ct receives cos of the refraction angle. R = 0 at exactly normal incidence like the
// reference (its ai == 0 branch) unless normal_limit asks for ((ni-nt)/(ni+nt))^2.
template <typename R>
__device__ __forceinline__ R fresnel_reflectance(R ni, R nt, R ci, bool normal_limit, R& ct)
{
    using M = Mth<R>;
    ct = (R)1;
    if (ci == (R)1) {
        const R q = (ni - nt) / (ni + nt);
        return normal_limit ? q * q : (R)0;
    }
    const R si = M::sqrt(M::fmax((R)1 - ci * ci, (R)0));
    const R sa = ni / nt * si;
    if (ni > nt && sa >= (R)1) return (R)1;        // total internal reflection (:85-86)
    ct = M::sqrt(M::fmax((R)1 - sa * sa, (R)0));
    const R sam = si * ct - ci * sa;               // sin(ai - at)
    const R sap = si * ct + ci * sa;               // sin(ai + at)
    const R cam = ci * ct + si * sa;               // cos(ai - at)
    const R cap = ci * ct - si * sa;               // cos(ai + at)
    const R q1 = M::div(sam, sap);
    const R q2 = M::div(cap, cam);
    return (R)0.5 * q1 * q1 * ((R)1 + q2 * q2);
}

// refraction into the next voxel along axis k (:97-106): tangential components scale by
// ni/nt, the normal one becomes sign(v_k) cos(at), then renormalise
template <typename R>
__device__ __forceinline__ void refract(int k, R vk, R rn, R ct, R& vx, R& vy, R& vz)
{
    using M = Mth<R>;
    R tx = vx * rn, ty = vy * rn, tz = vz * rn;
    const R nk = M::copysign(ct, vk);
    if (k == 0) tx = nk; else if (k == 1) ty = nk; else tz = nk;
    const R inv = M::rsqrt(tx * tx + ty * ty + tz * tz);
    vx = tx * inv; vy = ty * inv; vz = tz * inv;
}

// Henyey-Greenstein cos(theta) (:12-21), clamped to [-1, 1]
template <typename R>
__device__ __forceinline__ R hg_cos(const LabelConst<R>& L, R u)
{
    using M = Mth<R>;
    if (L.g != (R)0) {
        const R t = L.hg2 * M::rcp(L.hg3 + L.hg4 * u);
        const R c = (L.hg1 - t * t) * L.hginv;
        return M::clamp1(c);
    }
    return (R)2 * u - (R)1;
}

// new direction after scattering by theta (cos c) and azimuth 2 pi u_phi (:144-165)
template <typename R>
__device__ __forceinline__ void spin(R c, R u_phi, R& vx, R& vy, R& vz)
{
    using M = Mth<R>;
    const R sn = M::sqrt(M::fmax((R)1 - c * c, (R)0));
    R sf, cf;
    M::sincos2pi(u_phi, &sf, &cf);
    R nx_, ny_, nz_;
    if (M::fabs(vz) > (R)0.99999) {
        nx_ = sn * cf;
        ny_ = sn * sf;
        nz_ = vz > (R)0 ? c : -c;              // sign(v_z) * cos(theta), :153
    } else {
        const R den = M::sqrt((R)1 - vz * vz);
        const R iden = M::rcp(den);
        nx_ = sn * (vx * vz * cf - vy * sf) * iden + vx * c;
        ny_ = sn * (vy * vz * cf + vx * sf) * iden + vy * c;
        nz_ = -sn * cf * den + vz * c;
    }
    const R inv = M::rsqrt(nx_ * nx_ + ny_ * ny_ + nz_ * nz_);
    vx = nx_ * inv; vy = ny_ * inv; vz = nz_ * inv;
}

// lateral launch position of packet (plo, phi) for the flat-disc (kind 1) and Gaussian
// (kind 2) beams, as a voxel offset from the beam centre plus an in-voxel position
template <typename R>
__device__ __forceinline__ void launch_xy(int kind, float radius, R pitch, R half, uint32_t plo, uint32_t phi,
                                          uint32_t seed, uint32_t stream, int nx, int ny, int32_t& ax,
                                          int32_t& ay, R& px, R& py)
{
    using M = Mth<R>;
    const uint4 r = philox4x32_10(make_uint4(0u, kTagSource, plo, phi), seed, stream);
    const R u0 = u01<R>(r.x), u1 = u01<R>(r.y);
    R rad;
    if (kind == 1) rad = (R)radius * M::sqrt(u0);
    else rad = (R)(radius * 0.5f) * M::sqrt((R)2 * M::neglog(u0));
    R sa, ca;
    M::sincos2pi(u1, &sa, &ca);
    const R x = rad * ca, y = rad * sa;
    const int dx = (int)__builtin_floor((double)(x / pitch + (R)0.5));
    const int dy = (int)__builtin_floor((double)(y / pitch + (R)0.5));
    ax += dx; ay += dy;
    // in-voxel position, clamped: rounding of x / pitch can put it a hair outside the voxel
    px = x - pitch * (R)dx;
    py = y - pitch * (R)dy;
    px = px < -half ? -half : (px > half ? half : px);
    py = py < -half ? -half : (py > half ? half : py);
    // keep the launch voxel inside the padded volume's interior
    if (ax < 1) { ax = 1; px = -half; }
    if (ax > nx - 2) { ax = nx - 2; px = half; }
    if (ay < 1) { ay = 1; py = -half; }
    if (ay > ny - 2) { ay = ny - 2; py = half; }
}

// A position given relative to the centre of a voxel inside a uniform cube of half-width
// rad voxels, moved into the frame of the cube's voxel that holds it: returns the voxel
// offset, in [-rad, rad], and leaves p in [-half, half].
template <typename R>
__device__ __forceinline__ int cube_relocate(R& p, R pitch, R ipitch, R half, int rad)
{
    using M = Mth<R>;
    int o = (int)M::rint(p * ipitch);
    o = o < -rad ? -rad : (o > rad ? rad : o);
    p = M::fmin(M::fmax(p - (R)o * pitch, -half), half);
    return o;
}

__device__ __forceinline__ int32_t clamp_index(int32_t a, uint32_t n)
{
    return a < 0 ? 0 : (a >= (int32_t)n ? (int32_t)n - 1 : a);
}

// ------------------------------------------------------------------ line-buffered deposits
// A lane's line buffer holds the 8 FP64 sums of the 2x2x2 cube its packet last deposited in
// outside the LDS box; element e of lane t lives at lbuf[line_slot(e, t)] (element-major,
// lanes XOR-swizzled so that the eight lanes reading one buffer hit eight banks).
// (the stride is the block size, passed as a kernel argument: blockDim is read from the
// dispatch packet with a vector load, which would wait on every load in flight)
__device__ __forceinline__ uint32_t line_slot(uint32_t e, uint32_t t, uint32_t stride) { return e * stride + (t ^ e); }

// Hands the line buffers of the wave's lanes that `want` over to global memory, eight
// buffers per round: the eight lanes 8k..8k+7 read the eight elements of the k-th buffer and
// add the non-zero ones to its cube's line with one atomic instruction, so each buffer costs
// one 64 B request; the readers clear what they read. Wave-uniform: every lane of the wave
// must call it together. The stage (16 words per wave) passes buffer owners and lines.
__device__ __forceinline__ void flush_lines(double* lines, uint32_t stride, double* lbuf, int32_t* stage, bool want,
                                            int32_t& cur_line, unsigned long long lt_mask, uint32_t& n_req)
{
    const uint32_t lane = threadIdx.x & 63u;
    int32_t* st = stage + (threadIdx.x >> 6) * 16u;
    unsigned long long fm = __ballot(want);
    while (fm != 0ull) {
        const uint32_t nf = (uint32_t)__popcll(fm) < 8u ? (uint32_t)__popcll(fm) : 8u;
        const uint32_t rank = (uint32_t)__popcll(fm & lt_mask);
        const bool mine = want && rank < 8u;
        if (mine) { st[rank] = (int32_t)threadIdx.x; st[8u + rank] = cur_line; }
        __builtin_amdgcn_fence(__ATOMIC_RELEASE, "wavefront");
        __builtin_amdgcn_wave_barrier();
        __builtin_amdgcn_fence(__ATOMIC_ACQUIRE, "wavefront");
        const uint32_t slot = lane >> 3, j = lane & 7u;
        if (slot < nf) {
            const uint32_t owner = (uint32_t)st[slot];
            const uint32_t line = (uint32_t)st[8u + slot];
            double* q = lbuf + line_slot(j, owner, stride);
            const double v = *q;
            *q = 0.0;
            if (v != 0.0) gmem_add(lines + (size_t)line * 8u + j, v);
        }
        __builtin_amdgcn_fence(__ATOMIC_RELEASE, "wavefront");
        __builtin_amdgcn_wave_barrier();
        __builtin_amdgcn_fence(__ATOMIC_ACQUIRE, "wavefront");
        if (mine) { cur_line = -1; want = false; }
        if (lane == 0) n_req += nf;
        fm = __ballot(want);
    }
}

// ------------------------------------------------------------------ the walk kernel
// RZ: also tally the r-z fluence (a separate instantiation, so the plain walk keeps its
// register budget). SKIP: P.label_radius holds each voxel's label and uniform-cube radius
// (label | radius << 8, read with one load), and a packet in a
// voxel of radius r > 0 treats the cube of half-width r around it as one voxel: one step
// reaches its face or the hop's end. The faces inside the cube join voxels of one label,
// where the reference moves on without a draw (_kernel.py:82), so every packet makes the
// same draws (and events, unless rounding flips a decision at a face); only the arithmetic
// of the merged steps is rounded once instead of per face.
// LINES: the absorbed map outside the LDS box goes through per-lane line buffers (P.lines);
// a separate kernel with its own register budget (below), as it holds more per lane.
template <typename R, bool RZ, bool SKIP, bool LINES>
__device__ __forceinline__ void walk_body(const WalkParams& P)
{
    using M = Mth<R>;
    extern __shared__ __align__(16) unsigned char smem[];
    LabelConst<R>* tab = reinterpret_cast<LabelConst<R>*>(smem + P.lds_tab);
    double* bscal = reinterpret_cast<double*>(smem + P.lds_scal);
    unsigned long long* bev = reinterpret_cast<unsigned long long*>(smem + P.lds_ev);
    double* hist = reinterpret_cast<double*>(smem + P.lds_hist);
    double* win = reinterpret_cast<double*>(smem + P.lds_win);
    double* lbuf = reinterpret_cast<double*>(smem + P.lds_lbuf);
    int32_t* stage = reinterpret_cast<int32_t*>(smem + P.lds_stage);
    const bool use_lines = LINES && P.lines != nullptr;

    const uint32_t end = P.n_labels;
    for (uint32_t i = threadIdx.x; i <= end; i += blockDim.x) tab[i] = label_const<R>(P.optics[i]);
    for (uint32_t i = threadIdx.x; i < kBS_COUNT; i += blockDim.x) bscal[i] = 0.0;
    for (uint32_t i = threadIdx.x; i < kBE_COUNT; i += blockDim.x) bev[i] = 0ull;
    for (uint32_t i = threadIdx.x; i < P.h_total; i += blockDim.x) hist[i] = 0.0;
    for (uint32_t i = threadIdx.x; i < P.win_cells; i += blockDim.x) win[i] = 0.0;
    if (use_lines)
        for (uint32_t i = threadIdx.x; i < 8u * blockDim.x; i += blockDim.x) lbuf[i] = 0.0;
    __syncthreads();

    const uint32_t lane = threadIdx.x & 63u;
    const unsigned long long lt_mask = (lane == 0) ? 0ull : (~0ull >> (64u - lane));
    const R half = (R)P.pitch * (R)0.5;
    const R pitch = (R)P.pitch;
    const R ipitch = (R)1 / pitch;
    const bool normal_limit = (P.flags & 2u) != 0u;
    const uint32_t seed = P.key_dev ? __builtin_amdgcn_readfirstlane(P.key_dev[0]) : P.seed;
    const uint32_t skey = P.key_dev ? __builtin_amdgcn_readfirstlane(P.key_dev[1]) : P.stream;

    // lane state
    bool alive = false;
    R px = 0, py = 0, pz = 0, vx = 0, vy = 0, vz = 0, ivx = 0, ivy = 0, ivz = 0;
    R w = 0, st = 0, mt = 0, ni = 0, u_roul = 0, u_th = 0, u_phi = 0;
    R fb0 = 0, fb1 = 0, fb2 = 0;
    int32_t ax = 0, ay = 0, az = 0;    // voxel address (reference add[0..2])
    int32_t rad = 0;                   // uniform-cube radius around the voxel (SKIP), a lower bound
    uint32_t lab = 0, inter = 0, fres = 0, nev = 0;
    unsigned long long pid = 0;
    // lane accumulators (kept across packets)
    double acc_abs = 0.0;
    int32_t dep_slot = kNoSlot;      // where the pending deposit goes (see flush_deposit)
    double dep_acc = 0.0;
    uint32_t n_cross = 0, n_glob = 0;
    uint32_t n_pk = 0;               // packets this wave launched (wave-uniform)
    // line buffer: the cube it holds (-1 none) and a deposit into another cube waiting for
    // the buffer's hand-over at the top of the next iteration (p_line -1 none)
    int32_t cur_line = -1, p_line = -1;
    uint32_t p_e = 0;
    double p_dep = 0.0;

    // wave-uniform work batch
    unsigned long long b_next = 0, b_end = 0;
    bool exhausted = false;

    for (;;) {
        // ---------------------------------------------------- line buffers (all lanes here)
        if (use_lines) {
            flush_lines(P.lines, P.lbuf_stride, lbuf, stage, p_line >= 0 && cur_line >= 0, cur_line, lt_mask, n_glob);
            if (p_line >= 0) {
                cur_line = p_line;
                lds_add(lbuf + line_slot(p_e, threadIdx.x, P.lbuf_stride), p_dep);
                p_line = -1;
            }
        }
        // ---------------------------------------------------------- refill idle lanes
        const unsigned long long need = __ballot(!alive);
        if (need != 0ull && !exhausted && ((uint32_t)__popcll(need) >= P.refill || need == ~0ull)) {
            if (b_next == b_end) {
                unsigned long long claim = 0;
                if (lane == 0) claim = atomicAdd(P.work, (unsigned long long)P.batch);
                claim = __shfl(claim, 0);
                // wave-uniform: kept in scalar registers
                claim = ((unsigned long long)__builtin_amdgcn_readfirstlane((uint32_t)(claim >> 32)) << 32) |
                        (unsigned long long)__builtin_amdgcn_readfirstlane((uint32_t)claim);
                if (claim >= P.n_packets) {
                    exhausted = true;
                } else {
                    b_next = claim;
                    b_end = claim + P.batch < P.n_packets ? claim + P.batch : P.n_packets;
                }
            }
            if (!exhausted) {
                const uint32_t avail = (uint32_t)(b_end - b_next);
                const uint32_t rank = (uint32_t)__popcll(need & lt_mask);
                if (!alive && rank < avail) {
                    // ---------------- launch a packet (voxel/_classes.py:107-213)
                    pid = P.first_packet + b_next + rank;
                    const uint32_t plo = (uint32_t)pid, phi = (uint32_t)(pid >> 32);
                    ax = P.ix; ay = P.iy; az = P.iz;
                    px = 0; py = 0; pz = -half;
                    vx = (R)P.dirx; vy = (R)P.diry; vz = (R)P.dirz;
                    w = (R)P.w0;
                    if (P.src_kind == 3) {
                        // the caller's launch state (the reference kernels' add, p, v, w)
                        const unsigned long long q = pid - P.first_packet, sd = P.st_stride;
                        ax = P.st_add[q]; ay = P.st_add[sd + q]; az = P.st_add[2 * sd + q];
                        px = M::fmin(M::fmax((R)P.st_p[q], -half), half);
                        py = M::fmin(M::fmax((R)P.st_p[sd + q], -half), half);
                        pz = M::fmin(M::fmax((R)P.st_p[2 * sd + q], -half), half);
                        vx = (R)P.st_v[q]; vy = (R)P.st_v[sd + q]; vz = (R)P.st_v[2 * sd + q];
                        w = (R)P.st_w[q];
                    } else if (P.src_kind != 0) {
                        launch_xy<R>(P.src_kind, P.radius, pitch, half, plo, phi, seed, skey, (int)P.nx,
                                     (int)P.ny, ax, ay, px, py);
                    }
                    ivx = M::rcp(M::fabs(vx)); ivy = M::rcp(M::fabs(vy)); ivz = M::rcp(M::fabs(vz));
                    uint32_t l0 = end;   // an address outside the volume leaves at once
                    rad = 0;
                    if ((uint32_t)ax < P.nx && (uint32_t)ay < P.ny && (uint32_t)az < P.nz) {
                        const uint32_t li = ((uint32_t)ax * P.ny + (uint32_t)ay) * P.nz + (uint32_t)az;
                        if constexpr (SKIP) {
                            const uint32_t lr = P.label_radius[li];
                            l0 = lr & 255u;
                            rad = (int32_t)(lr >> 8);
                        } else {
                            l0 = P.labels[li];
                        }
                    }
                    lab = l0 < end ? l0 : end;
                    ni = tab[lab].n;
                    mt = tab[lab].mt;
                    inter = 0; fres = 0; nev = 0;
                    const uint4 r = philox4x32_10_sk(make_uint4(0u, kTagWalk, plo, phi), seed, skey);
                    st = M::neglog(u01<R>(r.x));
                    u_roul = u01<R>(r.y); u_th = u01<R>(r.z); u_phi = u01<R>(r.w);
                    alive = true;
                }
                const uint32_t launched = (uint32_t)__popcll(need) < avail ? (uint32_t)__popcll(need) : avail;
                b_next += launched;
                n_pk += launched;                 // wave-uniform

            }
        }
        if (__ballot(alive) == 0ull) break;

        // ------------------------------------------------------------ classify
        // Every live lane finds its next event: a face crossing if the remaining optical
        // depth reaches the exit face of its voxel, else the end of the hop inside it.
        // The wave then runs one of the two steps (a wave-uniform choice): the interaction
        // step once `quorum` lanes are ready for it or no lane has a crossing left, else
        // the crossing step. Lanes not served wait with their state unchanged, so each
        // packet's own sequence of events, and its draws, are the same as in any order.
        R db = (R)1000, smt = (R)0;
        int k = 0;
        bool want_int = false;
        const bool walking = alive && lab != end;
        if (walking) {
            // distance to the exit face, x < y < z on ties (_kernel.py:63-68)
            // only components with |v_i| > 0 take part (_kernel.py:65), as in the reference;
            // with SKIP, the exit face of the uniform cube around the voxel
            const R ext = SKIP ? (R)rad * pitch + half : half;
            const R d0 = vx != (R)0 ? (ext - (vx > (R)0 ? px : -px)) * ivx : (R)INFINITY;
            const R d1 = vy != (R)0 ? (ext - (vy > (R)0 ? py : -py)) * ivy : (R)INFINITY;
            const R d2 = vz != (R)0 ? (ext - (vz > (R)0 ? pz : -pz)) * ivz : (R)INFINITY;
            if (d0 < db) { db = d0; k = 0; }
            if (d1 < db) { db = d1; k = 1; }
            if (d2 < db) { db = d2; k = 2; }
            smt = db * mt;
            want_int = !(st >= smt);                               // (:69)
        }
        const unsigned long long m_int = __ballot(walking && want_int);
        const unsigned long long m_cross = __ballot(walking && !want_int);
        const bool run_int = m_cross == 0ull || (uint32_t)__popcll(m_int) >= P.quorum;
        if (!alive) continue;

        // ------------------------------------------------------------ one event
        bool ended = false;      // packet left through an end voxel
        bool killed = false;     // packet ended by roulette / event guard
        int xk = 0, xs = 1;      // axis and sign of the final crossing
        if (lab == end) {
            // launched inside an end voxel: leaves at once
            ended = true; xk = 2; xs = vz > (R)0 ? 1 : -1;
        } else if (want_int == run_int) {
            ++nev;
            if (!want_int && SKIP && rad > 0) {
                // ------------------------------------------------ to the uniform cube's face
                px += vx * db; py += vy * db; pz += vz * db;
                const R vk = k == 0 ? vx : (k == 1 ? vy : vz);
                st -= smt;
                const int s = vk > (R)0 ? 1 : -1;
                int ox = cube_relocate<R>(px, pitch, ipitch, half, rad);
                int oy = cube_relocate<R>(py, pitch, ipitch, half, rad);
                int oz = cube_relocate<R>(pz, pitch, ipitch, half, rad);
                const R face = M::copysign(half, vk);
                if (k == 0) { ox = s * rad; px = face; }
                else if (k == 1) { oy = s * rad; py = face; }
                else { oz = s * rad; pz = face; }
                ax = clamp_index(ax + ox, P.nx); ay = clamp_index(ay + oy, P.ny); az = clamp_index(az + oz, P.nz);
                n_cross += (uint32_t)(abs(ox) + abs(oy) + abs(oz));
                rad = 0;            // on the cube's face: the next step crosses it voxel by voxel
            } else if (!want_int) {
                // ------------------------------------------------ face crossing (:69-120)
                ++n_cross;
                px += vx * db; py += vy * db; pz += vz * db;
                const R vk = k == 0 ? vx : (k == 1 ? vy : vz);
                const R face = M::copysign(half, vk);
                if (k == 0) px = face; else if (k == 1) py = face; else pz = face;
                st -= smt;
                const int s = vk > (R)0 ? 1 : -1;
                const int32_t bx = ax + (k == 0 ? s : 0);
                const int32_t by = ay + (k == 1 ? s : 0);
                const int32_t bz = az + (k == 2 ? s : 0);
                uint32_t nxt = end;
                int32_t nrad = 0;
                if ((uint32_t)bx < P.nx && (uint32_t)by < P.ny && (uint32_t)bz < P.nz) {
                    const uint32_t nvi = (__umul24((uint32_t)bx, P.ny) + (uint32_t)by) * P.nz + (uint32_t)bz;
                    if constexpr (SKIP) {
                        // label and radius in one 2-byte load: one memory access per crossing
                        const uint32_t lr = P.label_radius[nvi];
                        nxt = lr & 255u;
                        nrad = (int32_t)(lr >> 8);
                    } else {
                        nxt = P.labels[nvi];
                    }
                    nxt = nxt < end ? nxt : end;
                }
                const R nt = tab[nxt].n;
                bool transmit = true;
                if (ni != nt) {
                    // Fresnel (:82-108) in cosine form; equal to the acos/asin/sin/tan form
                    R u;
                    if ((fres & 3u) == 0u) {
                        const uint4 r = philox4x32_10_sk(make_uint4(fres >> 2, kTagFresnel, (uint32_t)pid,
                                                                 (uint32_t)(pid >> 32)), seed, skey);
                        u = u01<R>(r.x); fb0 = u01<R>(r.y); fb1 = u01<R>(r.z); fb2 = u01<R>(r.w);
                    } else {
                        u = fb0; fb0 = fb1; fb1 = fb2;
                    }
                    ++fres;
                    R ct;
                    const R rf = fresnel_reflectance<R>(ni, nt, M::fabs(vk), normal_limit, ct);
                    if (rf < u) {
                        refract<R>(k, vk, ni / nt, ct, vx, vy, vz);
                        ivx = M::rcp(M::fabs(vx)); ivy = M::rcp(M::fabs(vy)); ivz = M::rcp(M::fabs(vz));
                    } else {
                        transmit = false;
                    }
                }
                if (transmit) {
                    ax = bx; ay = by; az = bz;
                    if (k == 0) px = -px; else if (k == 1) py = -py; else pz = -pz;
                    lab = nxt;
                    if (SKIP) rad = nrad;
                    if (lab == end) {
                        ended = true; xk = k; xs = s;
                    } else {
                        mt = tab[lab].mt;
                        ni = nt;
                    }
                } else {
                    if (k == 0) vx = -vx; else if (k == 1) vy = -vy; else vz = -vz;  // (:119-120)
                }
            } else {
                // ------------------------------------------------ hop ends in the voxel
                const R sm = M::div(st, mt);
                px += vx * sm; py += vy * sm; pz += vz * sm;       // (:122-124)
                if (SKIP && rad > 0) {
                    // the hop ended inside the uniform cube: into the frame of the voxel that
                    // holds its end; the cube around that voxel is at least rad - offset
                    const int ox = cube_relocate<R>(px, pitch, ipitch, half, rad);
                    const int oy = cube_relocate<R>(py, pitch, ipitch, half, rad);
                    const int oz = cube_relocate<R>(pz, pitch, ipitch, half, rad);
                    ax = clamp_index(ax + ox, P.nx); ay = clamp_index(ay + oy, P.ny); az = clamp_index(az + oz, P.nz);
                    const int mx = abs(ox), my = abs(oy), mz = abs(oz);
                    n_cross += (uint32_t)(mx + my + mz);
                    rad -= mx > my ? (mx > mz ? mx : mz) : (my > mz ? my : mz);
                }
                const LabelConst<R>& L = tab[lab];
                const R wn = w - w * L.ra;                         // drop (:130)
                const double dep = (double)w - (double)wn;        // exact in FP64
                w = wn;
                acc_abs += dep;
                if (RZ && L.inv_ma > (R)0) {
                    // r-z fluence: deposit / mu_a binned at the interaction point, in the
                    // encoded coordinates of voxel/_classes.py:261-269
                    const R x = (R)P.pitch * (R)(ax - P.cx) + px;
                    const R y = (R)P.pitch * (R)(ay - P.cy) + py;
                    const R z = (R)P.pitch * (R)(az - 1) + pz + half;
                    const int ir = (int)M::floor(M::sqrt(x * x + y * y) / (R)P.rz_dr);
                    const int iz = (int)M::floor(z / (R)P.rz_dz);
                    if ((uint32_t)ir < (uint32_t)P.rz_nr && (uint32_t)iz < (uint32_t)P.rz_nz)
                        lds_add(&hist[P.h_rz + (uint32_t)ir * (uint32_t)P.rz_nz + (uint32_t)iz], dep * (double)L.inv_ma);
                }
                if (P.absorbed) {
                    const uint32_t lx = (uint32_t)(ax - P.win_x0), ly = (uint32_t)(ay - P.win_y0),
                                   lz = (uint32_t)(az - P.win_z0);
                    const bool in_win = lx < P.win_nx && ly < P.win_ny && lz < P.win_nz;
                    if (use_lines && !in_win) {
                        // into the lane's line buffer if the cube is the one it holds, else
                        // wait for the hand-over of the buffer at the next iteration
                        const int32_t li = (int32_t)((__umul24((uint32_t)ax >> 1, P.nly) + ((uint32_t)ay >> 1)) * P.nlz +
                                                     ((uint32_t)az >> 1));
                        const uint32_t e = (((uint32_t)ax & 1u) << 2) | (((uint32_t)ay & 1u) << 1) | ((uint32_t)az & 1u);
                        if (li == cur_line) {
                            lds_add(lbuf + line_slot(e, threadIdx.x, P.lbuf_stride), dep);
                        } else {
                            p_line = li; p_e = e; p_dep = dep;
                        }
                    } else {
                        int32_t slot;
                        if (in_win)
                            slot = -(int32_t)((lx * P.win_ny + ly) * P.win_nz + lz) - 1;
                        else
                            slot = (int32_t)((__umul24((uint32_t)ax, P.ny) + (uint32_t)ay) * P.nz + (uint32_t)az);
                        if (slot == dep_slot) {
                            dep_acc += dep;
                        } else {
                            flush_deposit(P, win, dep_slot, dep_acc, n_glob);
                            dep_slot = slot;
                            dep_acc = dep;
                        }
                    }
                }
                if (w <= (R)P.wth) {                               // roulette (:132-141)
                    lds_add(&bev[kBE_ROULETTE], 1ull);
                    if ((R)P.inv_m < u_roul) {
                        killed = true;
                    } else {
                        const R wm = w * (R)P.roulette_m;
                        lds_add(&bscal[kBS_GAIN], (double)wm - (double)w);
                        w = wm;
                    }
                }
                if (!killed) {
                    spin<R>(hg_cos<R>(L, u_th), u_phi, vx, vy, vz);      // (:143-165)
                    ivx = M::rcp(M::fabs(vx)); ivy = M::rcp(M::fabs(vy)); ivz = M::rcp(M::fabs(vz));
                    ++inter;
                    const uint4 r = philox4x32_10_sk(make_uint4(inter, kTagWalk, (uint32_t)pid,
                                                                (uint32_t)(pid >> 32)), seed, skey);
                    st = M::neglog(u01<R>(r.x));
                    u_roul = u01<R>(r.y); u_th = u01<R>(r.z); u_phi = u01<R>(r.w);
                } else {
                    ++inter;
                }
            }
            if (!ended && !killed && nev >= P.max_events) killed = true;  // runaway guard
        }

        if (ended || killed) {
            // ------------------------------------------------------ packet end
            lds_add(&bev[kBE_INTERACT], (unsigned long long)inter);
            lds_add(&bev[kBE_FRESNEL], (unsigned long long)fres);
            const uint64_t slot = pid - P.first_packet;
            if (killed) {
                lds_add(&bev[kBE_KILLED], 1ull);
                lds_add(&bscal[kBS_KILL], (double)w);
                if (P.exit_state) {
                    // the reference zeroes add, p, v and w of a killed packet (_kernel.py:134-137);
                    // this is the encoding of that state
                    const unsigned long long n = P.n_packets;
                    float* e = P.exit_state;
                    e[slot] = -P.pitch * (float)P.cx;
                    e[n + slot] = -P.pitch * (float)P.cy;
                    e[2 * n + slot] = -0.5f * P.pitch;
                    e[3 * n + slot] = 0.f; e[4 * n + slot] = 0.f; e[5 * n + slot] = 0.f;
                    e[6 * n + slot] = 0.f;
                }
                if (P.exit_add) {
                    const unsigned long long n = P.n_packets;
                    P.exit_add[slot] = 0; P.exit_add[n + slot] = 0; P.exit_add[2 * n + slot] = 0;
                }
                if (P.exit_p) {
                    const unsigned long long n = P.n_packets;
                    P.exit_p[slot] = 0.f; P.exit_p[n + slot] = 0.f; P.exit_p[2 * n + slot] = 0.f;
                }
            } else {
                lds_add(&bev[kBE_ESCAPED], 1ull);
                const double wd = (double)w;
                // exit encoding, voxel/_classes.py:261-269: z is formed in double and rounded
                // to 6 decimals (np.round(..., 6)) before it becomes float, so exits through a
                // face lie exactly on it (the bottom face at the plate's thickness, the top at 0)
                const float x = P.pitch * (float)(ax - P.cx) + (float)px;
                const float y = P.pitch * (float)(ay - P.cy) + (float)py;
                const double z64 = (double)P.pitch * (double)(az - 1) + (double)pz + 0.5 * (double)P.pitch;
                const float z = (float)(__builtin_rint(z64 * 1e6) * 1e-6);
                const float fvz = (float)vz;
                lds_add(&bscal[fvz < 0.f ? kBS_RD : (fvz > 0.f ? kBS_TT : kBS_OTHER)], wd);
                lds_add(&bscal[kBS_FACE + 2 * xk + (xs > 0 ? 1 : 0)], wd);
                // top-face detectors select v_z < 0 and z < 1e-8 on the encoded exit, the
                // reference's selection (virtualOBD.py:165)
                const bool top = fvz < 0.f && z < 1e-8f;
                // squares and sum rounded one by one (no contraction), as numpy forms r from
                // the float32 exit dump, so host and device binning part only at bin edges
                const float r = __builtin_sqrtf(__fadd_rn(__fmul_rn(x, x), __fmul_rn(y, y)));
                if (P.nr > 0) {
                    const int b = bin_of(r, P.dr, P.nr);
                    if (b >= 0) {
                        if (fvz < 0.f && P.rd_r) lds_add(&hist[P.h_rd + b], wd);
                        if (fvz > 0.f && P.tt_r) lds_add(&hist[P.h_tt + b], wd);
                        if (top && P.rd_top_r) lds_add(&hist[P.h_top + b], wd);
                    }
                }
                if (P.na > 0) {
                    const float ang = acosf(__builtin_fabsf(fvz));
                    const int b = bin_of(ang, P.da, P.na);
                    if (b >= 0) {
                        if (fvz < 0.f && P.rd_a) lds_add(&hist[P.h_rda + b], wd);
                        if (fvz > 0.f && P.tt_a) lds_add(&hist[P.h_tta + b], wd);
                    }
                }
                if (P.nl > 0 && P.line_z && (float)vx < 0.f && __builtin_fabsf(y) < P.line_y) {
                    const int b = bin_of(z, P.dl, P.nl);
                    if (b >= 0) lds_add(&hist[P.h_line + b], wd);
                }
                if (top && P.n_rings > 0 && P.ring_w) {
                    for (int i = 0; i < P.n_rings; ++i)
                        if (r > P.rings[2 * i] && r <= P.rings[2 * i + 1]) lds_add(&hist[P.h_ring + i], wd);
                }
                if (P.exit_state) {
                    const unsigned long long n = P.n_packets;
                    float* e = P.exit_state;
                    e[slot] = x; e[n + slot] = y; e[2 * n + slot] = z;
                    e[3 * n + slot] = (float)vx; e[4 * n + slot] = (float)vy; e[5 * n + slot] = fvz;
                    e[6 * n + slot] = (float)w;
                }
                if (P.exit_add) {
                    // the reference kernel's exit state: the end voxel and the position in its
                    // frame (_kernel.py:110-116)
                    const unsigned long long n = P.n_packets;
                    P.exit_add[slot] = ax; P.exit_add[n + slot] = ay; P.exit_add[2 * n + slot] = az;
                }
                if (P.exit_p) {
                    const unsigned long long n = P.n_packets;
                    P.exit_p[slot] = (float)px; P.exit_p[n + slot] = (float)py; P.exit_p[2 * n + slot] = (float)pz;
                }
            }
            alive = false;
        }
    }

    // ---------------------------------------------------------------- block flush
    if (use_lines) {
        // the deposits still waiting, then every line buffer (all lanes are here)
        flush_lines(P.lines, P.lbuf_stride, lbuf, stage, p_line >= 0 && cur_line >= 0, cur_line, lt_mask, n_glob);
        if (p_line >= 0) {
            cur_line = p_line;
            lds_add(lbuf + line_slot(p_e, threadIdx.x, P.lbuf_stride), p_dep);
            p_line = -1;
        }
        flush_lines(P.lines, P.lbuf_stride, lbuf, stage, cur_line >= 0, cur_line, lt_mask, n_glob);
    }
    if (P.absorbed) flush_deposit(P, win, dep_slot, dep_acc, n_glob);
    lds_add(&bscal[kBS_ABS], acc_abs);
    if (lane == 0) {
        lds_add(&bscal[kBS_W0], (double)n_pk * (double)P.w0);
        lds_add(&bev[kBE_PACKETS], (unsigned long long)n_pk);
    }
    lds_add(&bev[kBE_CROSS], (unsigned long long)n_cross);
    lds_add(&bev[kBE_GLOBAL_DEPOSITS], (unsigned long long)n_glob);
    __syncthreads();
    for (uint32_t i = threadIdx.x; i < kBS_COUNT; i += blockDim.x)
        if (bscal[i] != 0.0) gmem_add(P.scalars + i, bscal[i]);
    if (P.events)
        for (uint32_t i = threadIdx.x; i < kBE_COUNT; i += blockDim.x)
            if (bev[i] != 0ull) atomicAdd(P.events + i, bev[i]);
    for (uint32_t i = threadIdx.x; i < P.h_total; i += blockDim.x) {
        const double v = hist[i];
        if (v == 0.0) continue;
        double* dst;
        if (i < P.h_tt) dst = P.rd_r + (i - P.h_rd);
        else if (i < P.h_top) dst = P.tt_r + (i - P.h_tt);
        else if (i < P.h_rda) dst = P.rd_top_r + (i - P.h_top);
        else if (i < P.h_tta) dst = P.rd_a + (i - P.h_rda);
        else if (i < P.h_line) dst = P.tt_a + (i - P.h_tta);
        else if (i < P.h_ring) dst = P.line_z + (i - P.h_line);
        else if (i < P.h_rz) dst = P.ring_w + (i - P.h_ring);
        else dst = P.fluence_rz + (i - P.h_rz);
        gmem_add(dst, v);
    }
    for (uint32_t i = threadIdx.x; i < P.win_cells; i += blockDim.x) {
        const double v = win[i];
        if (v == 0.0) continue;
        const uint32_t lz = i % P.win_nz, t = i / P.win_nz;
        const uint32_t ly = t % P.win_ny, lx = t / P.win_ny;
        const uint32_t g = (((uint32_t)P.win_x0 + lx) * P.ny + (uint32_t)P.win_y0 + ly) * P.nz + (uint32_t)P.win_z0 + lz;
        gmem_add(P.absorbed + g, v);
    }
}

template <typename R, bool RZ, bool SKIP>
__global__ void __launch_bounds__(1024) __attribute__((amdgpu_waves_per_eu(6))) walk_kernel(const WalkParams P)
{
    walk_body<R, RZ, SKIP, false>(P);
}

// the FP32 walk with line-buffered absorbed deposits: it runs at 4 waves per SIMD (one
// 1024-lane block per CU, or four 256-lane blocks), whose register budget holds the line
// buffer's bookkeeping without spills
template <bool SKIP>
__global__ void __launch_bounds__(1024) __attribute__((amdgpu_waves_per_eu(4))) walk_lines_kernel(const WalkParams P)
{
    walk_body<float, false, SKIP, true>(P);
}

// Fluence2D / Fluence3D binning of points (pymopt/fluence/intarnal_fluence.py). With
// lds_bins > 0 (all the bins) each block sums its points in an LDS copy of the bins and adds
// the non-zero ones to `out` at its end: points crowd a few bins near the beam, whose global
// FP64 atomics would otherwise serialise on their lines.
__global__ void __launch_bounds__(1024) fluence_bin_kernel(const float* p, unsigned long long stride, const float* w,
                                                           unsigned long long n, int dims, int nr, float dr, int nz,
                                                           float dz, double* out, uint32_t lds_bins)
{
    extern __shared__ __align__(16) unsigned char smem_h[];
    double* h = reinterpret_cast<double*>(smem_h);
    for (uint32_t i = threadIdx.x; i < lds_bins; i += blockDim.x) h[i] = 0.0;
    if (lds_bins) __syncthreads();
    const unsigned long long step = (unsigned long long)gridDim.x * blockDim.x;
    for (unsigned long long j = (unsigned long long)blockIdx.x * blockDim.x + threadIdx.x; j < n; j += step) {
        const float x = p[j], y = p[stride + j], z = p[2 * stride + j];
        const int iz = (int)__builtin_floorf(z / dz);
        if (!(iz >= 0 && iz < nz)) continue;
        unsigned long long b;
        if (dims == 2) {
            const int ir = (int)__builtin_floorf(__builtin_sqrtf(x * x + y * y) / dr);
            if (!(ir >= 0 && ir < nr)) continue;
            b = (unsigned long long)ir * nz + iz;
        } else {
            const int ix = (int)__builtin_floorf(x / dr) + nr, iy = (int)__builtin_floorf(y / dr) + nr;
            if (!(ix >= 0 && ix < 2 * nr && iy >= 0 && iy < 2 * nr)) continue;
            b = ((unsigned long long)ix * (2 * nr) + iy) * nz + iz;
        }
        if (lds_bins) lds_add(&h[b], (double)w[j]);
        else gmem_add(out + b, (double)w[j]);
    }
    if (lds_bins) {
        __syncthreads();
        for (uint32_t i = threadIdx.x; i < lds_bins; i += blockDim.x)
            if (h[i] != 0.0) gmem_add(out + i, h[i]);
    }
}

__global__ void zero_u64_kernel(unsigned long long* p, int n)
{
    const int i = threadIdx.x;
    if (i < n) p[i] = 0ull;
}

// clears n doubles (a fill kernel, so the walk can be graph-captured)
__global__ void __launch_bounds__(256) pack_label_radius_kernel(const uint8_t* labels, const uint8_t* radius,
                                                                 uint16_t* out, unsigned long long n)
{
    const unsigned long long stride = (unsigned long long)gridDim.x * blockDim.x;
    for (unsigned long long i = (unsigned long long)blockIdx.x * blockDim.x + threadIdx.x; i < n; i += stride)
        out[i] = (uint16_t)((uint32_t)labels[i] | ((uint32_t)radius[i] << 8));
}

__global__ void __launch_bounds__(256) zero_f64_kernel(double* p, unsigned long long n)
{
    const unsigned long long step = (unsigned long long)gridDim.x * blockDim.x;
    for (unsigned long long i = (unsigned long long)blockIdx.x * blockDim.x + threadIdx.x; i < n; i += step) p[i] = 0.0;
}

// map[x][y][z] += lines[cube of (x, y, z)][element]: the cube-major line sums into the C-order
// map (the absorbed map, or the replay's gradient); one thread per cube, its 8 elements to the
// up-to-8 voxels inside the volume
__global__ void __launch_bounds__(256) lines_to_absorbed_kernel(const double* lines, double* absorbed, uint32_t nx,
                                                                uint32_t ny, uint32_t nz, uint32_t nly, uint32_t nlz,
                                                                unsigned long long n_cubes)
{
    const unsigned long long step = (unsigned long long)gridDim.x * blockDim.x;
    for (unsigned long long c = (unsigned long long)blockIdx.x * blockDim.x + threadIdx.x; c < n_cubes; c += step) {
        const uint32_t cz = (uint32_t)(c % nlz);
        const unsigned long long t = c / nlz;
        const uint32_t cy = (uint32_t)(t % nly), cx = (uint32_t)(t / nly);
        const double* q = lines + c * 8u;
        for (uint32_t e = 0; e < 8u; ++e) {
            const double v = q[e];
            if (v == 0.0) continue;
            const uint32_t x = 2u * cx + (e >> 2), y = 2u * cy + ((e >> 1) & 1u), z = 2u * cz + (e & 1u);
            if (x < nx && y < ny && z < nz) absorbed[((size_t)x * ny + y) * nz + z] += v;
        }
    }
}

// probes of the FP64 atomic rate and the FP32 FMA issue rate (every index in bounds)
// XCD (accelerator complex die) the calling wave runs on: HW_REG_XCC_ID (hwreg 20), bits 3:0
__device__ __forceinline__ uint32_t xcc_id() { return __builtin_amdgcn_s_getreg((3 << 11) | 20) & 7u; }

// Atomic adds (no return) of T to pseudo-random addresses of buf[n], at device scope or
// (WG) workgroup scope; with per_xcd, buf holds 8 copies of n elements and each workgroup
// adds into its XCD's copy. On gfx950 both scopes compile to the same global_atomic_add
// (no scope bits), so they probe the same path.
template <class T, bool WG = false>
__global__ void probe_atomics_kernel(T* buf, uint32_t n, int iters, int per_xcd)
{
    uint32_t x = (blockIdx.x * blockDim.x + threadIdx.x) * 2654435761u ^ 2463534242u;
    T* base = per_xcd ? buf + (size_t)xcc_id() * n : buf;
    for (int i = 0; i < iters; ++i) {
        x ^= x << 13; x ^= x >> 17; x ^= x << 5;          // xorshift32
        if (WG) __hip_atomic_fetch_add(base + (x % n), (T)1, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_WORKGROUP);
        else __hip_atomic_fetch_add(base + (x % n), (T)1, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
    }
}

// The same adds with the lanes in groups of 2^lg (lg = 2..6): each group adds to the 2^lg
// consecutive doubles of its own pseudo-random, 2^lg-aligned block of buf[n] (n >= 2^lg),
// one double per lane. lg = 6: a wave on eight 64 B lines of one 512 B block; lg = 2, 3, 4:
// a wave on 16 random 32 B, 8 random 64 B or 4 random 128 B blocks. Every address is below
// (n >> lg) << lg <= n.
__global__ void probe_atomics_lines_kernel(double* buf, uint32_t n, int iters, uint32_t lg)
{
    uint32_t x = ((blockIdx.x * blockDim.x + threadIdx.x) >> lg) * 2654435761u ^ 2463534242u;
    const uint32_t lane = threadIdx.x & ((1u << lg) - 1u), nb = n >> lg;
    for (int i = 0; i < iters; ++i) {
        x ^= x << 13; x ^= x >> 17; x ^= x << 5;
        __hip_atomic_fetch_add(buf + ((size_t)(x % nb) << lg) + lane, 1.0, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
    }
}

__global__ void probe_valu_kernel(float* out, int iters)
{
    const uint32_t t = blockIdx.x * blockDim.x + threadIdx.x;
    float a[8], b = 1.0000001f, c = 1e-7f;
#pragma unroll
    for (int j = 0; j < 8; ++j) a[j] = (float)(t + j);
    for (int i = 0; i < iters; ++i) {
#pragma unroll
        for (int r = 0; r < 4; ++r) {
#pragma unroll
            for (int j = 0; j < 8; ++j) a[j] = __builtin_fmaf(a[j], b, c);
        }
    }
    float s = 0.f;
#pragma unroll
    for (int j = 0; j < 8; ++j) s += a[j];
    out[t] = s;
}

// Philox4x32-10 blocks back to back, as the walk draws them (counter = block, tag, lane id),
// folded into one word per lane so that none is optimised away
__global__ void probe_philox_kernel(uint32_t* out, int iters, uint32_t seed)
{
    const uint32_t t = blockIdx.x * blockDim.x + threadIdx.x;
    uint32_t acc = 0u;
    for (int i = 0; i < iters; ++i) {
        const uint4 r = philox4x32_10(make_uint4((uint32_t)i, kTagWalk, t, 0u), seed, 7u);
        acc ^= r.x + r.y + r.z + r.w;
    }
    out[t] = acc;
}

__global__ void philox_kernel(const uint4* ctr, uint32_t k0, uint32_t k1, uint4* out, int64_t n)
{
    const int64_t i = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
    if (i < n) out[i] = philox4x32_10(ctr[i], k0, k1);
}

}  // namespace pmo
