// Differentiable walk: forward pass with perturbation-Monte-Carlo path statistics, and the
// adjoint by path replay (gfx950).
//
// One kernel, two modes selected at run time, so forward and replay run the very same
// instructions and a replayed packet retraces its forward path exactly:
//  - forward: the walk of pymopt/voxel/_kernel.py with Rd(r) and, for every packet leaving
//    with v_z < 0 into radial bin b, sum_j W L_j and W n_j per label j (L_j path length and
//    n_j interactions in label j). With implicit capture the weight is prod (mu_s/mu_t)
//    and the path density prod mu_t^n exp(-mu_t L), so
//        d E[W] / d mu_a,j = E[-W L_j],   d E[W] / d mu_s,j = E[W (n_j / mu_s,j - L_j)],
//    which gives the Jacobian of Rd(r) in one pass. Each packet's exit (bin, W) is recorded.
//  - replay: for a loss gradient g_b over Rd(r), packet p that left into bin b contributes
//    -c_b W_p l to d loss / d mu_a of every voxel it crosses for a length l, with
//    c_b = g_b / (area_b N). The path is regenerated from the packet's Philox counter
//    (no stored paths); the replayed exit is checked against the record.
// Optionally mu_a is a per-voxel field instead of the label's value, which makes the
// per-voxel gradient the derivative of this forward with respect to that field.
// The replay's deposits outside its LDS box may go through the walk's per-lane line buffers
// (one 2x2x2-voxel cube each, handed over eight lanes per atomic instruction) into a
// cube-major scratch that is added to the gradient map afterwards.
#pragma once

#include "walk.hpp"

namespace pmo {

constexpr int kMaxJacLabels = 4;

struct DiffParams {
    const uint8_t* labels;
    const uint8_t* skip;           // forward (SKIP): uniform-cube radius per voxel (pmo_skip_radius)
    const float4* optics;
    const float* mua_field;        // per-voxel mu_a (padded index) or null
    uint32_t nx, ny, nz, n_labels;
    int32_t cx, cy;
    float pitch;
    int32_t src_kind, ix, iy, iz;
    float radius, dirx, diry, dirz, w0;
    unsigned long long n_packets;
    uint32_t seed, stream;
    float wth, roulette_m, inv_m;
    uint32_t flags, batch, max_events;
    uint32_t quorum;               // lanes ready for an interaction before a wave runs that step
    int32_t mode;                  // 0 forward, 1 replay
    double* scalars;               // [kBS_COUNT] (forward)
    unsigned long long* events;    // [kBE_COUNT] or null
    int32_t nr;
    float dr;
    double* rd_r;                  // [nr] (forward)
    double* jac;                   // [nr][n_labels][2] sums of W L_j, W n_j (forward) or null
    int2* exit_rec;                // [n_packets] (bin or -1, W as float bits)
    const double* coef;            // [nr] replay weights g_b / (area_b N)
    double* grad;                  // [nx*ny*nz] d loss / d mu_a per voxel (replay)
    unsigned long long* mismatch;  // replayed exits that differ from the record
    unsigned long long* work;
    double* lines;                 // replay: cube-major line scratch (8 doubles per 2x2x2 cube) or null
    uint32_t nly, nlz;             // cubes along y and z
    uint32_t lbuf_stride;          // block size (see line_slot)
    int32_t win_x0, win_y0, win_z0;
    uint32_t win_nx, win_ny, win_nz, win_cells;
    uint32_t lds_tab, lds_scal, lds_ev, lds_hist, lds_jac, lds_lbuf, lds_stage, lds_win;
};

// LINES: the replay with line buffers, an instantiation of its own so that the forward keeps
// its register budget. SKIP: a forward without the per-voxel field that crosses the faces
// inside same-label cubes in one step, as the plain walk does (the same draws, merged
// arithmetic): the per-label sums need no voxel-by-voxel path, but its exit record is not
// the voxel-stepping replay's, so it is never replayed.
template <typename R, bool LINES, bool SKIP>
__global__ void __launch_bounds__(1024) __attribute__((amdgpu_waves_per_eu(4))) diff_kernel(const DiffParams P)
{
    using M = Mth<R>;
    extern __shared__ __align__(16) unsigned char smem[];
    LabelConst<R>* tab = reinterpret_cast<LabelConst<R>*>(smem + P.lds_tab);
    double* bscal = reinterpret_cast<double*>(smem + P.lds_scal);
    unsigned long long* bev = reinterpret_cast<unsigned long long*>(smem + P.lds_ev);
    double* hist = reinterpret_cast<double*>(smem + P.lds_hist);
    double* hjac = reinterpret_cast<double*>(smem + P.lds_jac);
    double* win = reinterpret_cast<double*>(smem + P.lds_win);
    double* lbuf = reinterpret_cast<double*>(smem + P.lds_lbuf);
    int32_t* stage = reinterpret_cast<int32_t*>(smem + P.lds_stage);

    const uint32_t end = P.n_labels;
    const uint32_t njac = P.jac ? P.n_labels : 0u;
    const bool replay = P.mode == 1;
    const bool use_lines = LINES && replay && P.lines != nullptr;
    for (uint32_t i = threadIdx.x; i <= end; i += blockDim.x) tab[i] = label_const<R>(P.optics[i]);
    for (uint32_t i = threadIdx.x; i < kBS_COUNT; i += blockDim.x) bscal[i] = 0.0;
    for (uint32_t i = threadIdx.x; i < kBE_COUNT; i += blockDim.x) bev[i] = 0ull;
    for (uint32_t i = threadIdx.x; i < (uint32_t)P.nr; i += blockDim.x) hist[i] = 0.0;
    for (uint32_t i = threadIdx.x; i < (uint32_t)P.nr * njac * 2u; i += blockDim.x) hjac[i] = 0.0;
    for (uint32_t i = threadIdx.x; i < P.win_cells; i += blockDim.x) win[i] = 0.0;
    if (use_lines)
        for (uint32_t i = threadIdx.x; i < 8u * P.lbuf_stride; i += blockDim.x) lbuf[i] = 0.0;
    __syncthreads();

    const uint32_t lane = threadIdx.x & 63u;
    const unsigned long long lt_mask = (lane == 0) ? 0ull : (~0ull >> (64u - lane));
    const R half = (R)P.pitch * (R)0.5;
    const R pitch = (R)P.pitch;
    const R ipitch = (R)1 / pitch;
    const bool normal_limit = (P.flags & 2u) != 0u;
    const bool field = P.mua_field != nullptr;

    bool alive = false;
    R px = 0, py = 0, pz = 0, vx = 0, vy = 0, vz = 0, ivx = 0, ivy = 0, ivz = 0;
    R w = 0, st = 0, mt = 0, ra = 0, ni = 0, u_roul = 0, u_th = 0, u_phi = 0;
    R fb0 = 0, fb1 = 0, fb2 = 0;
    R L0 = 0, L1 = 0, L2 = 0, L3 = 0, N0 = 0, N1 = 0, N2 = 0, N3 = 0;   // per-label path stats
    int32_t ax = 0, ay = 0, az = 0;
    int32_t rad = 0;                 // SKIP: uniform-cube radius around the voxel, a lower bound
    uint32_t lab = 0, inter = 0, fres = 0, nev = 0;
    unsigned long long pid = 0;
    PhiloxPacket pq{0u, 0u, 0u, 0u};  // the packet's part of its interaction draws
    double acc_abs = 0.0;
    // replay: pending gradient deposit and this packet's -c_b W. The slot is -(box cell)-1 in
    // the LDS box; outside it the voxel index, or with line buffers 8 x cube + element
    int32_t g_slot = kNoSlot;
    double g_len = 0.0, g_coef = 0.0;
    int32_t rec_bin = -1;
    float rec_w = 0.f;
    uint32_t n_cross = 0, n_req = 0;
    uint32_t n_pk = 0;               // packets this wave launched (wave-uniform)
    // line buffer: the cube it holds (-1 none) and a deposit into another cube waiting for
    // the buffer's hand-over at the top of the next iteration (p_line -1 none)
    int32_t cur_line = -1, p_line = -1;
    uint32_t p_e = 0;
    double p_dep = 0.0;

    unsigned long long b_next = 0, b_end = 0;
    bool exhausted = false;

    auto voxel_optics = [&](uint32_t l, int32_t x, int32_t y, int32_t z) {
        // mu_t and drop fraction of voxel (x, y, z) with label l
        if (field) {
            const uint32_t vi = (__umul24((uint32_t)x, P.ny) + (uint32_t)y) * P.nz + (uint32_t)z;
            const R ma = (R)P.mua_field[vi];
            mt = ma + tab[l].ms;                              // mu_a(voxel) + mu_s(label)
            ra = mt > (R)0 ? ma / mt : (R)0;
        } else {
            mt = tab[l].mt;
            ra = tab[l].ra;
        }
    };
    auto add_path = [&](uint32_t l, R len, R cnt) {
        L0 += l == 0 ? len : (R)0; N0 += l == 0 ? cnt : (R)0;
        L1 += l == 1 ? len : (R)0; N1 += l == 1 ? cnt : (R)0;
        L2 += l == 2 ? len : (R)0; N2 += l == 2 ? cnt : (R)0;
        L3 += l == 3 ? len : (R)0; N3 += l == 3 ? cnt : (R)0;
    };
    auto grad_emit = [&](int32_t slot, double v, bool last) {
        if (slot < 0) {
            lds_add(win + (-slot - 1), v);
        } else if (!use_lines) {
            gmem_add(P.grad + slot, v);
        } else {
            const int32_t li = slot >> 3;
            const uint32_t e = (uint32_t)slot & 7u;
            if (li == cur_line) lds_add(lbuf + line_slot(e, threadIdx.x, P.lbuf_stride), v);
            else if (!last) { p_line = li; p_e = e; p_dep = v; }
            else gmem_add(P.lines + slot, v);
        }
    };
    auto grad_segment = [&](R len) {
        // replay: length `len` travelled in the current voxel
        const uint32_t lx = (uint32_t)(ax - P.win_x0), ly = (uint32_t)(ay - P.win_y0), lz = (uint32_t)(az - P.win_z0);
        int32_t slot;
        if (lx < P.win_nx && ly < P.win_ny && lz < P.win_nz)
            slot = -(int32_t)((lx * P.win_ny + ly) * P.win_nz + lz) - 1;
        else if (use_lines)
            slot = (int32_t)((((__umul24((uint32_t)ax >> 1, P.nly) + ((uint32_t)ay >> 1)) * P.nlz + ((uint32_t)az >> 1)) << 3) |
                             (((uint32_t)ax & 1u) << 2) | (((uint32_t)ay & 1u) << 1) | ((uint32_t)az & 1u));
        else
            slot = (int32_t)((__umul24((uint32_t)ax, P.ny) + (uint32_t)ay) * P.nz + (uint32_t)az);
        if (slot == g_slot) {
            g_len += (double)len;
        } else {
            if (g_slot != kNoSlot) grad_emit(g_slot, g_coef * g_len, false);
            g_slot = slot;
            g_len = (double)len;
        }
    };
    auto grad_flush = [&]() {
        // the packet's last deposit: this iteration may already hold one for the hand-over,
        // so a deposit outside the lane's current cube goes to its line directly
        if (g_slot != kNoSlot) grad_emit(g_slot, g_coef * g_len, true);
        g_slot = kNoSlot;
        g_len = 0.0;
    };

    for (;;) {
        if (use_lines) {
            // line buffers (all lanes are here): hand over those whose lane waits to start another
            flush_lines(P.lines, P.lbuf_stride, lbuf, stage, p_line >= 0 && cur_line >= 0, cur_line, lt_mask, n_req);
            if (p_line >= 0) {
                cur_line = p_line;
                lds_add(lbuf + line_slot(p_e, threadIdx.x, P.lbuf_stride), p_dep);
                p_line = -1;
            }
        }
        const unsigned long long need = __ballot(!alive);
        if (need != 0ull && !exhausted) {
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
                    pid = b_next + rank;
                    bool skip = false;
                    if (replay) {
                        const int2 rec = P.exit_rec[pid];
                        rec_bin = rec.x;
                        rec_w = __int_as_float(rec.y);
                        const double c = rec_bin >= 0 && rec_bin < P.nr ? P.coef[rec_bin] : 0.0;
                        skip = (c == 0.0);
                        g_coef = -c * (double)rec_w;
                    }
                    if (!skip) {
                        const uint32_t plo = (uint32_t)pid, phi = (uint32_t)(pid >> 32);
                        ax = P.ix; ay = P.iy; az = P.iz;
                        px = 0; py = 0; pz = -half;
                        if (P.src_kind != 0)
                            launch_xy<R>(P.src_kind, P.radius, pitch, half, plo, phi, P.seed, P.stream, (int)P.nx,
                                         (int)P.ny, ax, ay, px, py);
                        vx = (R)P.dirx; vy = (R)P.diry; vz = (R)P.dirz;
                        ivx = M::rcp(M::fabs(vx)); ivy = M::rcp(M::fabs(vy)); ivz = M::rcp(M::fabs(vz));
                        w = (R)P.w0;
                        const uint32_t vi = ((uint32_t)ax * P.ny + (uint32_t)ay) * P.nz + (uint32_t)az;
                        const uint32_t l0 = P.labels[vi];
                        if (SKIP) rad = P.skip[vi];
                        lab = l0 < end ? l0 : end;
                        ni = tab[lab].n;
                        if (lab != end) voxel_optics(lab, ax, ay, az);
                        inter = 0; fres = 0; nev = 0;
                        L0 = L1 = L2 = L3 = N0 = N1 = N2 = N3 = (R)0;
                        pq = philox_packet(kTagWalk, plo, P.seed);
                        const uint4 r = philox_packet_block_sk(pq, 0u, phi, P.seed, P.stream);
                        st = M::neglog(u01<R>(r.x));
                        u_roul = u01<R>(r.y); u_th = u01<R>(r.z); u_phi = u01<R>(r.w);
                        alive = true;
                    }
                }
                const uint32_t launched = (uint32_t)__popcll(need) < avail ? (uint32_t)__popcll(need) : avail;
                b_next += launched;
                n_pk += launched;                 // wave-uniform
            }
        }
        if (__ballot(alive) == 0ull) break;

        // classify as the plain walk does: each live lane's next event is a face crossing or
        // the end of its hop; the wave runs one kind per step (interactions once `quorum`
        // lanes want one or none crosses). A lane not served waits unchanged, so every
        // packet's events and draws are those of any other order, and the replay retraces
        // the forward exactly.
        R db = (R)1000, smt = (R)0;
        int k = 0;
        bool want_int = false;
        const bool walking = alive && lab != end;
        if (walking) {
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
            want_int = !(st >= smt);
        }
        const unsigned long long m_int = __ballot(walking && want_int);
        const unsigned long long m_cross = __ballot(walking && !want_int);
        const bool run_int = m_cross == 0ull || (uint32_t)__popcll(m_int) >= P.quorum;
        if (!alive) continue;

        bool ended = false, killed = false;
        if (lab == end) {
            ended = true;
        } else if (want_int == run_int) {
            ++nev;
            if (!want_int && SKIP && rad > 0) {
                // to the uniform cube's face (as in the plain walk): one label throughout
                px += vx * db; py += vy * db; pz += vz * db;
                if (njac) add_path(lab, db, (R)0);
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
                ++n_cross;
                px += vx * db; py += vy * db; pz += vz * db;
                if (replay) grad_segment(db);
                else if (njac) add_path(lab, db, (R)0);
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
                    nxt = P.labels[nvi];
                    if (SKIP) nrad = P.skip[nvi];
                    nxt = nxt < end ? nxt : end;
                }
                const R nt = tab[nxt].n;
                bool transmit = true;
                if (ni != nt) {
                    R u;
                    if ((fres & 3u) == 0u) {
                        const uint4 r = philox4x32_10_sk(make_uint4(fres >> 2, kTagFresnel, (uint32_t)pid,
                                                                 (uint32_t)(pid >> 32)), P.seed, P.stream);
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
                        ended = true;
                    } else {
                        voxel_optics(lab, ax, ay, az);
                        ni = nt;
                    }
                } else {
                    if (k == 0) vx = -vx; else if (k == 1) vy = -vy; else vz = -vz;
                }
            } else {
                const R sm = M::div(st, mt);
                px += vx * sm; py += vy * sm; pz += vz * sm;
                if (replay) grad_segment(sm);
                else if (njac) add_path(lab, sm, (R)1);
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
                const R wn = w - w * ra;
                acc_abs += (double)w - (double)wn;
                w = wn;
                if (w <= (R)P.wth) {
                    if (!replay) lds_add(&bev[kBE_ROULETTE], 1ull);
                    if ((R)P.inv_m < u_roul) {
                        killed = true;
                    } else {
                        const R wm = w * (R)P.roulette_m;
                        if (!replay) lds_add(&bscal[kBS_GAIN], (double)wm - (double)w);
                        w = wm;
                    }
                }
                if (!killed) {
                    spin<R>(hg_cos<R>(tab[lab], u_th), u_phi, vx, vy, vz);
                    ivx = M::rcp(M::fabs(vx)); ivy = M::rcp(M::fabs(vy)); ivz = M::rcp(M::fabs(vz));
                    ++inter;
                    const uint4 r = philox_packet_block_sk(pq, inter, (uint32_t)(pid >> 32), P.seed, P.stream);
                    st = M::neglog(u01<R>(r.x));
                    u_roul = u01<R>(r.y); u_th = u01<R>(r.z); u_phi = u01<R>(r.w);
                } else {
                    ++inter;
                }
            }
            if (!ended && !killed && nev >= P.max_events) killed = true;
        }

        if (ended || killed) {
            // exit bin of a v_z < 0 escape (the detected signal), -1 otherwise
            int bin = -1;
            const float fvz = (float)vz;
            if (ended && fvz < 0.f && P.nr > 0) {
                const float x = P.pitch * (float)(ax - P.cx) + (float)px;
                const float y = P.pitch * (float)(ay - P.cy) + (float)py;
                bin = bin_of(__builtin_sqrtf(x * x + y * y), P.dr, P.nr);
            }
            if (replay) {
                grad_flush();
                if (bin != rec_bin || (float)w != rec_w) atomicAdd(P.mismatch, 1ull);
            } else {
                lds_add(&bev[kBE_INTERACT], (unsigned long long)inter);
                lds_add(&bev[kBE_FRESNEL], (unsigned long long)fres);
                if (killed) {
                    lds_add(&bev[kBE_KILLED], 1ull);
                    lds_add(&bscal[kBS_KILL], (double)w);
                } else {
                    lds_add(&bev[kBE_ESCAPED], 1ull);
                    lds_add(&bscal[fvz < 0.f ? kBS_RD : (fvz > 0.f ? kBS_TT : kBS_OTHER)], (double)w);
                }
                if (bin >= 0) {
                    const double wd = (double)w;
                    lds_add(&hist[bin], wd);
                    if (njac) {
                        double* h = hjac + (size_t)bin * njac * 2u;
                        const R Ls[4] = {L0, L1, L2, L3}, Ns[4] = {N0, N1, N2, N3};
#pragma unroll
                        for (uint32_t j = 0; j < (uint32_t)kMaxJacLabels; ++j) {
                            if (j < njac) {
                                lds_add(h + 2 * j, wd * (double)Ls[j]);
                                lds_add(h + 2 * j + 1, wd * (double)Ns[j]);
                            }
                        }
                    }
                }
                if (P.exit_rec) P.exit_rec[pid] = make_int2(bin, __float_as_int(bin >= 0 ? (float)w : 0.f));
            }
            alive = false;
        }
    }

    if (use_lines) {
        // every line buffer still held (no deposit waits: the loop's top placed the last)
        flush_lines(P.lines, P.lbuf_stride, lbuf, stage, cur_line >= 0, cur_line, lt_mask, n_req);
    }
    if (!replay) {
        lds_add(&bscal[kBS_ABS], acc_abs);
        if (lane == 0) {
            lds_add(&bscal[kBS_W0], (double)n_pk * (double)P.w0);
            lds_add(&bev[kBE_PACKETS], (unsigned long long)n_pk);
        }
        lds_add(&bev[kBE_CROSS], (unsigned long long)n_cross);
    }
    __syncthreads();
    if (!replay) {
        for (uint32_t i = threadIdx.x; i < kBS_COUNT; i += blockDim.x)
            if (bscal[i] != 0.0) gmem_add(P.scalars + i, bscal[i]);
        if (P.events)
            for (uint32_t i = threadIdx.x; i < kBE_COUNT; i += blockDim.x)
                if (bev[i] != 0ull) atomicAdd(P.events + i, bev[i]);
        for (uint32_t i = threadIdx.x; i < (uint32_t)P.nr; i += blockDim.x)
            if (hist[i] != 0.0) gmem_add(P.rd_r + i, hist[i]);
        for (uint32_t i = threadIdx.x; i < (uint32_t)P.nr * njac * 2u; i += blockDim.x)
            if (hjac[i] != 0.0) gmem_add(P.jac + i, hjac[i]);
    } else {
        for (uint32_t i = threadIdx.x; i < P.win_cells; i += blockDim.x) {
            const double v = win[i];
            if (v == 0.0) continue;
            const uint32_t lz = i % P.win_nz, t = i / P.win_nz;
            const uint32_t ly = t % P.win_ny, lx = t / P.win_ny;
            const uint32_t g =
                (((uint32_t)P.win_x0 + lx) * P.ny + (uint32_t)P.win_y0 + ly) * P.nz + (uint32_t)P.win_z0 + lz;
            gmem_add(P.grad + g, v);
        }
    }
}

}  // namespace pmo
