// Batched walk over many layered plates in one persistent launch (gfx950).
//
// Each instance is a plate (pymopt/voxel/_classes.py:352-424 PlateModel geometry): a box of
// nx x ny x nz padded voxels whose interior label depends on z only, given as a column of
// nz labels (slices 0 and nz-1 and the lateral shell are the end label). A homogeneous slab
// is a one-layer plate. Instances have their own optics, launch source and packet budget.
//
// Work is a schedule of packet batches (instance, first local id, count), ordered
// round-robin over instances so that every instance's tallies are updated throughout the
// launch instead of in one burst. A wave claims one batch per counter atomic and sums the
// tallies of that batch's instance in its own LDS block (the instance's scalars, then its
// Rd(r) and Tt(r) bins), which it adds to global memory when it claims a batch of another
// instance: a few dozen global atomics per batch instead of three per packet, which
// otherwise contend on each instance's few lines. A packet of an earlier instance, still
// walking when its wave moves on, adds its tallies to global memory directly. Packet
// (instance i, local id j) draws from Philox key (seed, stream + i) with packet id j, the
// same draws as a plain walk of instance i alone with stream + i, so each instance's
// result is that of its own plain walk.
#pragma once

#include "walk.hpp"

namespace pmo {

struct PlateInstance {     // device layout of pmo_plate_instance (include/pmo.h)
    const uint8_t* z_labels;
    const float4* optics;
    int32_t nx, ny, nz, n_labels;
    float pitch;
    int32_t src_kind, ix, iy, iz;
    float radius, dirx, diry, dirz, w0;
    unsigned long long n_packets;
    const float4* layers;       // (z_top, z_bottom, label, 0) per run of equal labels, top to bottom
    int32_t n_layers, reserved;
};

struct BatchParams {
    const PlateInstance* inst;
    int32_t n_inst;
    const uint4* schedule;      // (instance, first local id, count, 0)
    unsigned long long n_batches;
    uint32_t seed, stream;
    float wth, roulette_m, inv_m;
    uint32_t flags;
    uint32_t max_events;
    uint32_t quorum;            // lanes wanting an interaction before a wave runs that step, >= 1
    uint32_t refill;            // idle lanes before a wave starts new packets (all idle: always)
    double* scalars;            // [n_inst][kBS_COUNT]
    unsigned long long* events; // [kBE_COUNT]
    int32_t nr;
    float dr;
    double* rd_r;               // [n_inst][nr]
    double* tt_r;               // [n_inst][nr]
    unsigned long long* work;
    const uint32_t* key_dev;
    uint32_t wave_acc;          // doubles of a wave's LDS tally of its current instance (0: none)
};

// the fields of the current instance a lane needs on every event, kept in registers
struct PlateGeo {
    const uint8_t* z_labels;
    const float4* optics;
    int32_t nx, ny, nz;
    uint32_t end;
    const float4* layers;
    int32_t n_layers;
};

__device__ __forceinline__ uint32_t plate_label(const PlateGeo& G, int32_t bx, int32_t by, int32_t bz)
{
    if (bx < 1 || by < 1 || bx > G.nx - 2 || by > G.ny - 2 || (uint32_t)bz >= (uint32_t)G.nz) return G.end;
    const uint32_t l = G.z_labels[bz];
    return l < G.end ? l : G.end;
}

// label across the face on axis k: in a plate only z-crossings can change layers, so
// lateral crossings keep the label unless they enter the lateral shell
__device__ __forceinline__ uint32_t plate_next_label(const PlateGeo& G, uint32_t lab, int k, int32_t bx, int32_t by,
                                                     int32_t bz)
{
    if (k == 2) return plate_label(G, bx, by, bz);
    return (bx < 1 || by < 1 || bx > G.nx - 2 || by > G.ny - 2) ? G.end : lab;
}

// LAYER = false: the voxel walk of pymopt/voxel/_kernel.py on the plate's voxels, identical
// to pmo_walk on the same plate. LAYER = true: the same walk with the position kept in plate
// coordinates (mm from the top-face centre) and faces only where the label can change: the
// layer boundaries and the lateral walls. The voxel faces it skips change neither label nor
// index (_kernel.py:82,110-118 do nothing there but st -= db*mt), so the two agree in exact
// arithmetic and draw the same random numbers.
template <typename R, bool LAYER>
__global__ void __launch_bounds__(1024) __attribute__((amdgpu_waves_per_eu(5)))
plate_batch_kernel(const BatchParams P)
{
    using M = Mth<R>;
    __shared__ unsigned long long bev[kBE_COUNT];
    extern __shared__ __align__(16) unsigned char smem_b[];
    for (uint32_t i = threadIdx.x; i < kBE_COUNT; i += blockDim.x) bev[i] = 0ull;
    // the wave's tally of instance w_inst: [kBS_COUNT scalars | nr Rd(r) | nr Tt(r)]
    double* wacc = reinterpret_cast<double*>(smem_b) + (threadIdx.x >> 6) * P.wave_acc;
    for (uint32_t i = threadIdx.x & 63u; i < P.wave_acc; i += 64u) wacc[i] = 0.0;
    __syncthreads();

    const uint32_t lane = threadIdx.x & 63u;
    const unsigned long long lt_mask = (lane == 0) ? 0ull : (~0ull >> (64u - lane));
    const bool normal_limit = (P.flags & 2u) != 0u;
    const uint32_t seed = P.key_dev ? __builtin_amdgcn_readfirstlane(P.key_dev[0]) : P.seed;
    const uint32_t skey = P.key_dev ? __builtin_amdgcn_readfirstlane(P.key_dev[1]) : P.stream;

    bool alive = false;
    R px = 0, py = 0, pz = 0, vx = 0, vy = 0, vz = 0, ivx = 0, ivy = 0, ivz = 0;
    R imt = 0;                        // LAYER: 1 / mt of the current label
    R w = 0, st = 0, u_roul = 0, u_th = 0, u_phi = 0, fb0 = 0, fb1 = 0, fb2 = 0, half = 0;
    LabelConst<R> L;                  // optics of the current label
    PlateGeo G{nullptr, nullptr, 0, 0, 0, 0u, nullptr, 0};
    R zt = 0, zb = 0, wxl = 0, wxh = 0, wyl = 0, wyh = 0;   // LAYER: current layer, lateral walls
    int32_t ax = 0, ay = 0, az = 0;
    uint32_t lab = 0, inter = 0, fres = 0, nev = 0, pid = 0, key1 = 0;
    PhiloxPacket pq{0u, 0u, 0u, 0u};  // the packet's part of its interaction draws
    int32_t inst = -1;
    double acc_abs = 0.0, acc_w0 = 0.0;
    uint32_t n_cross = 0;

    // wave-uniform current batch
    uint32_t b_inst = 0, b_next = 0, b_end = 0;
    bool exhausted = false;
    int32_t w_inst = -1;              // wave-uniform: the instance the wave's LDS tally holds

    // a tally of the lane's instance: into the wave's LDS block when it holds that instance
    auto inst_add = [&](uint32_t j, double v) {
        if (inst == w_inst && P.wave_acc != 0u) lds_add(&wacc[j], v);
        else if (j < (uint32_t)kBS_COUNT) gmem_add(P.scalars + (size_t)inst * kBS_COUNT + j, v);
        else if (j < (uint32_t)kBS_COUNT + (uint32_t)P.nr) gmem_add(P.rd_r + (size_t)inst * P.nr + (j - kBS_COUNT), v);
        else gmem_add(P.tt_r + (size_t)inst * P.nr + (j - kBS_COUNT - (uint32_t)P.nr), v);
    };
    // every lane of the wave together: the LDS block to global memory, cleared
    auto flush_wave = [&]() {
        __builtin_amdgcn_fence(__ATOMIC_RELEASE, "wavefront");
        __builtin_amdgcn_wave_barrier();
        __builtin_amdgcn_fence(__ATOMIC_ACQUIRE, "wavefront");
        if (w_inst >= 0) {
            for (uint32_t j = lane; j < P.wave_acc; j += 64u) {
                const double v = wacc[j];
                if (v != 0.0) {
                    wacc[j] = 0.0;
                    double* dst = j < (uint32_t)kBS_COUNT ? P.scalars + (size_t)w_inst * kBS_COUNT + j
                                : (j < (uint32_t)kBS_COUNT + (uint32_t)P.nr
                                       ? P.rd_r + (size_t)w_inst * P.nr + (j - kBS_COUNT)
                                       : P.tt_r + (size_t)w_inst * P.nr + (j - kBS_COUNT - (uint32_t)P.nr));
                    gmem_add(dst, v);
                }
            }
        }
        __builtin_amdgcn_fence(__ATOMIC_RELEASE, "wavefront");
        __builtin_amdgcn_wave_barrier();
        __builtin_amdgcn_fence(__ATOMIC_ACQUIRE, "wavefront");
    };

    for (;;) {
        const unsigned long long need = __ballot(!alive);
        if (need != 0ull && !exhausted && ((uint32_t)__popcll(need) >= P.refill || need == ~0ull)) {
            if (b_next == b_end) {
                unsigned long long claim = 0;
                if (lane == 0) claim = atomicAdd(P.work, 1ull);
                claim = __shfl(claim, 0);
                if (claim >= P.n_batches) {
                    exhausted = true;
                } else {
                    const uint4 b = P.schedule[claim];
                    b_inst = b.x < (uint32_t)P.n_inst ? b.x : 0u;
                    b_next = b.y;
                    b_end = b.x < (uint32_t)P.n_inst ? b.y + b.z : b.y;
                    if (P.wave_acc != 0u && (int32_t)b_inst != w_inst) {
                        flush_wave();
                        w_inst = (int32_t)b_inst;
                    }
                }
            }
            if (!exhausted) {
                const uint32_t avail = b_end - b_next;
                const uint32_t rank = (uint32_t)__popcll(need & lt_mask);
                if (!alive && rank < avail) {
                    if ((int32_t)b_inst != inst) {
                        // leaving an instance: flush its per-lane sums
                        if (inst >= 0) {
                            gmem_add(P.scalars + (size_t)inst * kBS_COUNT + kBS_ABS, acc_abs);
                            gmem_add(P.scalars + (size_t)inst * kBS_COUNT + kBS_W0, acc_w0);
                        }
                        acc_abs = 0.0; acc_w0 = 0.0;
                        inst = (int32_t)b_inst;
                    }
                    const PlateInstance& I = P.inst[inst];
                    G.z_labels = I.z_labels; G.optics = I.optics;
                    G.nx = I.nx; G.ny = I.ny; G.nz = I.nz; G.end = (uint32_t)I.n_labels;
                    G.layers = I.layers; G.n_layers = I.n_layers;
                    pid = b_next + rank;
                    key1 = skey + (uint32_t)inst;
                    half = (R)I.pitch * (R)0.5;
                    ax = I.ix; ay = I.iy; az = I.iz;
                    px = 0; py = 0; pz = -half;
                    if (I.src_kind != 0)
                        launch_xy<R>(I.src_kind, I.radius, (R)I.pitch, half, pid, 0u, seed, key1, I.nx, I.ny, ax,
                                     ay, px, py);
                    vx = (R)I.dirx; vy = (R)I.diry; vz = (R)I.dirz;
                    if constexpr (!LAYER) {
                        ivx = M::rcp(M::fabs(vx)); ivy = M::rcp(M::fabs(vy)); ivz = M::rcp(M::fabs(vz));
                    }
                    w = (R)I.w0;
                    acc_w0 += (double)I.w0;
                    // absorbed weight telescoped over the packet: each drop is exactly
                    // (double)w - (double)w_new, so their sum is the launch weight, minus
                    // the final weight, plus the roulette gains (added at the end)
                    acc_abs += (double)w;
                    lab = plate_label(G, ax, ay, az);
                    if constexpr (LAYER) {
                        // plate coordinates of the launch point (the exit encoding of
                        // voxel/_classes.py:261-269) and the walls of the interior
                        const R pitch = (R)I.pitch;
                        const int32_t cx = (I.nx - 1) / 2, cy = (I.ny - 1) / 2;
                        px = pitch * (R)(ax - cx) + px;
                        py = pitch * (R)(ay - cy) + py;
                        pz = pitch * (R)(az - 1) + pz + half;
                        wxl = pitch * (R)(1 - cx) - half; wxh = pitch * (R)(I.nx - 2 - cx) + half;
                        wyl = pitch * (R)(1 - cy) - half; wyh = pitch * (R)(I.ny - 2 - cy) + half;
                        // the layer holding the launch depth; none (end label) without a table
                        az = -1;
                        if (lab != G.end && G.layers != nullptr) {
                            for (int32_t j = 0; j < G.n_layers; ++j) {
                                const float4 Ly = G.layers[j];
                                if (pz < (R)Ly.y) { az = j; zt = (R)Ly.x; zb = (R)Ly.y; lab = (uint32_t)Ly.z; break; }
                            }
                        }
                        if (az < 0 || lab > G.end) lab = G.end;
                    }
                    L = label_const<R>(G.optics[lab]);
                    if constexpr (LAYER) imt = M::rcp(L.mt);
                    inter = 0; fres = 0; nev = 0;
                    pq = philox_packet(kTagWalk, pid, seed);
                    const uint4 r = philox_packet_block(pq, 0u, 0u, seed, key1);
                    st = M::neglog(u01<R>(r.x));
                    u_roul = u01<R>(r.y); u_th = u01<R>(r.z); u_phi = u01<R>(r.w);
                    alive = true;
                }
                b_next += (uint32_t)__popcll(need) < avail ? (uint32_t)__popcll(need) : avail;
            }
        }
        if (__ballot(alive) == 0ull) break;

        // classify every walking lane as a face crossing or an interaction, then run one
        // kind per wave step: interactions once `quorum` lanes want one (or none crosses),
        // so the long interaction block does not run with most of the wave masked off
        const uint32_t end = G.end;
        R db = (R)1000, smt = (R)0;
        int k = 0;
        bool want_int = false;
        R ex = 0, ey = 0, ez = 0;             // LAYER: end point of the pending hop
        const bool walking = alive && lab != end;
        if (walking) {
            if constexpr (LAYER) {
                // the hop of length st / mt meets no wall or layer boundary exactly when it
                // ends inside the closed box of the current layer; its end point is then the
                // interaction point (the fused p + v st / mt of the voxel walk)
                const R s = st * imt;
                ex = px + vx * s; ey = py + vy * s; ez = pz + vz * s;
                want_int = ex >= wxl && ex <= wxh && ey >= wyl && ey <= wyh && ez >= zt && ez <= zb;
            } else {
                // only components with |v_i| > 0 take part (_kernel.py:65), as in the reference
                const R d0 = vx != (R)0 ? (half - (vx > (R)0 ? px : -px)) * ivx : (R)INFINITY;
                const R d1 = vy != (R)0 ? (half - (vy > (R)0 ? py : -py)) * ivy : (R)INFINITY;
                const R d2 = vz != (R)0 ? (half - (vz > (R)0 ? pz : -pz)) * ivz : (R)INFINITY;
                if (d0 < db) { db = d0; k = 0; }
                if (d1 < db) { db = d1; k = 1; }
                if (d2 < db) { db = d2; k = 2; }
                smt = db * L.mt;
                want_int = !(st >= smt);
            }
        }
        const unsigned long long m_int = __ballot(walking && want_int);
        const unsigned long long m_cross = __ballot(walking && !want_int);
        const bool run_int = m_cross == 0ull || (uint32_t)__popcll(m_int) >= P.quorum;
        if (!alive) continue;

        bool ended = false, killed = false;
        int xk = 2, xs = 1;
        if (lab == end) {
            ended = true; xs = vz > (R)0 ? 1 : -1;
        } else if (want_int == run_int) {
            ++nev;
            if (!want_int) {
                ++n_cross;
                if constexpr (LAYER) {
                    // distance to the wall or boundary the hop meets (_kernel.py:62-70)
                    const R d0 = vx != (R)0 ? (vx > (R)0 ? wxh - px : px - wxl) * M::rcp(M::fabs(vx)) : (R)INFINITY;
                    const R d1 = vy != (R)0 ? (vy > (R)0 ? wyh - py : py - wyl) * M::rcp(M::fabs(vy)) : (R)INFINITY;
                    const R d2 = vz != (R)0 ? (vz > (R)0 ? zb - pz : pz - zt) * M::rcp(M::fabs(vz)) : (R)INFINITY;
                    if (d0 < db) { db = d0; k = 0; }
                    if (d1 < db) { db = d1; k = 1; }
                    if (d2 < db) { db = d2; k = 2; }
                    smt = db * L.mt;
                }
                px += vx * db; py += vy * db; pz += vz * db;
                const R vk = k == 0 ? vx : (k == 1 ? vy : vz);
                const int s = vk > (R)0 ? 1 : -1;
                st -= smt;
                // where the end-point test and st >= db mt round to different answers the
                // remainder is a few ulps below zero: the next hop starts on the face
                if constexpr (LAYER) st = M::fmax(st, (R)0);
                uint32_t nxt = end;
                int32_t bx = ax, by = ay, bz = az;
                R nzt = zt, nzb = zb;
                if constexpr (LAYER) {
                    // onto the wall or layer boundary crossed; only z faces lead to a layer
                    if (k == 0) px = s > 0 ? wxh : wxl;
                    else if (k == 1) py = s > 0 ? wyh : wyl;
                    else pz = s > 0 ? zb : zt;
                    if (k == 2) {
                        bz = az + s;
                        if (bz >= 0 && bz < G.n_layers) {
                            const float4 Ly = G.layers[bz];
                            const uint32_t l = (uint32_t)Ly.z;
                            nxt = l < end ? l : end;
                            nzt = (R)Ly.x; nzb = (R)Ly.y;
                        }
                    }
                } else {
                    const R face = M::copysign(half, vk);
                    if (k == 0) px = face; else if (k == 1) py = face; else pz = face;
                    bx = ax + (k == 0 ? s : 0);
                    by = ay + (k == 1 ? s : 0);
                    bz = az + (k == 2 ? s : 0);
                    nxt = plate_next_label(G, lab, k, bx, by, bz);
                }
                bool transmit = true;
                LabelConst<R> Ln = L;
                if (nxt != lab) {
                    Ln = label_const<R>(G.optics[nxt]);
                    if (L.n != Ln.n) {
                        R u;
                        if ((fres & 3u) == 0u) {
                            const uint4 r = philox4x32_10(make_uint4(fres >> 2, kTagFresnel, pid, 0u), seed, key1);
                            u = u01<R>(r.x); fb0 = u01<R>(r.y); fb1 = u01<R>(r.z); fb2 = u01<R>(r.w);
                        } else {
                            u = fb0; fb0 = fb1; fb1 = fb2;
                        }
                        ++fres;
                        R ct;
                        const R rf = fresnel_reflectance<R>(L.n, Ln.n, M::fabs(vk), normal_limit, ct);
                        if (rf < u) {
                            refract<R>(k, vk, L.n / Ln.n, ct, vx, vy, vz);
                            if constexpr (!LAYER) {
                                ivx = M::rcp(M::fabs(vx)); ivy = M::rcp(M::fabs(vy)); ivz = M::rcp(M::fabs(vz));
                            }
                        } else {
                            transmit = false;
                        }
                    }
                }
                if (transmit) {
                    if constexpr (LAYER) {
                        az = bz; zt = nzt; zb = nzb;
                    } else {
                        ax = bx; ay = by; az = bz;
                        if (k == 0) px = -px; else if (k == 1) py = -py; else pz = -pz;
                    }
                    if (nxt == end) {
                        ended = true; xk = k; xs = s;
                    } else if (nxt != lab) {
                        L = Ln;
                        if constexpr (LAYER) imt = M::rcp(L.mt);
                    }
                    lab = nxt;
                } else {
                    if (k == 0) vx = -vx; else if (k == 1) vy = -vy; else vz = -vz;
                }
            } else {
                if constexpr (LAYER) {
                    px = ex; py = ey; pz = ez;
                } else {
                    const R sm = M::div(st, L.mt);
                    px += vx * sm; py += vy * sm; pz += vz * sm;
                }
                w = w - w * L.ra;
                if (w <= (R)P.wth) {
                    atomicAdd(&bev[kBE_ROULETTE], 1ull);
                    if ((R)P.inv_m < u_roul) {
                        killed = true;
                    } else {
                        const R wm = w * (R)P.roulette_m;
                        const double gain = (double)wm - (double)w;
                        inst_add(kBS_GAIN, gain);
                        acc_abs += gain;
                        w = wm;
                    }
                }
                if (!killed) {
                    spin<R>(hg_cos<R>(L, u_th), u_phi, vx, vy, vz);
                    if constexpr (!LAYER) {
                        ivx = M::rcp(M::fabs(vx)); ivy = M::rcp(M::fabs(vy)); ivz = M::rcp(M::fabs(vz));
                    }
                    ++inter;
                    // the instance key is per lane; when every lane in this step has the
                    // same one (a wave mostly walks one batch of one instance), its round
                    // keys step in scalar registers instead of one vector add per round
                    const uint32_t k1u = __builtin_amdgcn_readfirstlane(key1);
                    uint4 r;
                    if (__ballot(key1 != k1u) == 0ull) r = philox_packet_block_sk(pq, inter, 0u, seed, k1u);
                    else r = philox_packet_block(pq, inter, 0u, seed, key1);
                    st = M::neglog(u01<R>(r.x));
                    u_roul = u01<R>(r.y); u_th = u01<R>(r.z); u_phi = u01<R>(r.w);
                } else {
                    ++inter;
                }
            }
            if (!ended && !killed && nev >= P.max_events) killed = true;
        }

        if (ended || killed) {
            acc_abs -= (double)w;
            atomicAdd(&bev[kBE_INTERACT], (unsigned long long)inter);
            atomicAdd(&bev[kBE_FRESNEL], (unsigned long long)fres);
            atomicAdd(&bev[kBE_PACKETS], 1ull);
            if (killed) {
                atomicAdd(&bev[kBE_KILLED], 1ull);
                inst_add(kBS_KILL, (double)w);
            } else {
                atomicAdd(&bev[kBE_ESCAPED], 1ull);
                const double wd = (double)w;
                const float fvz = (float)vz;
                inst_add(fvz < 0.f ? kBS_RD : (fvz > 0.f ? kBS_TT : kBS_OTHER), wd);
                inst_add(kBS_FACE + 2 * xk + (xs > 0 ? 1 : 0), wd);
                if (P.nr > 0 && fvz != 0.f) {
                    float x, y;
                    if constexpr (LAYER) {
                        x = (float)px; y = (float)py;
                    } else {
                        const float pitch = P.inst[inst].pitch;
                        const int cx = (G.nx - 1) / 2, cy = (G.ny - 1) / 2;
                        x = pitch * (float)(ax - cx) + (float)px;
                        y = pitch * (float)(ay - cy) + (float)py;
                    }
                    const int b = bin_of(__builtin_sqrtf(x * x + y * y), P.dr, P.nr);
                    if (b >= 0 && (fvz < 0.f ? P.rd_r : P.tt_r) != nullptr)
                        inst_add((uint32_t)kBS_COUNT + (fvz < 0.f ? 0u : (uint32_t)P.nr) + (uint32_t)b, wd);
                }
            }
            alive = false;
        }
    }
    if (inst >= 0) {
        gmem_add(P.scalars + (size_t)inst * kBS_COUNT + kBS_ABS, acc_abs);
        gmem_add(P.scalars + (size_t)inst * kBS_COUNT + kBS_W0, acc_w0);
    }
    if (P.wave_acc != 0u) flush_wave();
    atomicAdd(&bev[kBE_CROSS], (unsigned long long)n_cross);
    __syncthreads();
    if (P.events)
        for (uint32_t i = threadIdx.x; i < kBE_COUNT; i += blockDim.x)
            if (bev[i] != 0ull) atomicAdd(P.events + i, bev[i]);
}

}  // namespace pmo
