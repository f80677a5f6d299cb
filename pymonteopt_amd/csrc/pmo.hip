// libpmo.so: C ABI of the MI355X voxel photon walk (declarations and reference
// citations in include/pmo.h). Host code here validates arguments, sizes the persistent
// grid and the dynamic LDS layout, and launches on the caller's stream. It allocates
// nothing: the only per-call scratch is the caller's workspace.
#include <hip/hip_runtime.h>

#include <math.h>
#include <stdarg.h>
#include <stdio.h>
#include <string.h>

#include <mutex>

#include "pmo.h"
#include "adjoint.hpp"
#include "batch.hpp"
#include "obd.hpp"
#include "skip.hpp"
#include "turing.hpp"
#include "walk.hpp"

namespace {

thread_local char g_err[512] = "";

int fail(int code, const char* fmt, ...)
{
    va_list ap;
    va_start(ap, fmt);
    vsnprintf(g_err, sizeof(g_err), fmt, ap);
    va_end(ap);
    return code;
}

int hip_fail(hipError_t e, const char* what)
{
    return fail(PMO_ERR_HIP, "%s: %s", what, hipGetErrorString(e));
}

constexpr size_t kWorkspaceBytes = 256;
constexpr int kDefaultBlock = 256;
constexpr int kDefaultPacketsPerLane = 4;
constexpr uint32_t kMaxEvents = 1u << 24;
constexpr size_t kMaxLds = 64 * 1024;  // used when the device does not report its per-block LDS
constexpr int kDefaultWindow[3] = {16, 16, 8};
// Runs of at least kWideRunPackets packets with an absorbed map default to one 1024-lane
// block per CU holding a kWideWindow LDS window. One 28x28x24 box around the launch voxel
// takes 14% of the skin plate's deposits, against 5% for 16x16x8, and cuts the global FP64
// atomics per packet from 195 to 177.
constexpr int kWideBlock = 1024;
constexpr int kWideWindow[3] = {28, 28, 24};
constexpr uint64_t kWideRunPackets = 1ull << 20;
// A volume of at most kDefaultWindow's cell count is held whole in the window of a regular
// launch, whose blocks keep more waves per CU than one 1024-lane block: on the 12^3 padded
// 5 mm slab at 0.5 mm, 114 ms per 1e8 packets against 127 ms for the wide launch.
constexpr unsigned long long kWholeVolumeCells = (unsigned long long)kDefaultWindow[0] * kDefaultWindow[1] * kDefaultWindow[2];
// interactions run once 32 of a wave's 64 lanes want one. On the MI355X, against 48: +11%
// on the 256^3 bone phantom with either launch shape, +2% on the skin plate without the
// absorbed map, a tie with it, -2% on the slab
constexpr int kDefaultQuorum = 32;
// With the absorbed map through cube lines (absorbed_lines), a quorum of 40 and windows of at
// most 22x22x22 (wide launch: one 1024-lane block per CU, the window beside 64 KiB of line
// buffers) or 12x12x12 (256-lane blocks, as many per CU as the registers allow, 5, each
// window in its block's share of the LDS beside 16 KiB of line buffers). On the MI355X at 1e8
// packets: skin plate 1.56e8 photons/s wide (1.43e8 with 256-lane blocks; 1.28e8 with
// per-voxel atomics); 256^3 bone phantom 1.44e8 wide, 1.56e8 with 256-lane blocks and a
// 9x9x21 window (1.40e8 per voxel).
constexpr int kLinesWideWindow[3] = {22, 22, 22};
constexpr int kLinesWindow[3] = {12, 12, 12};
constexpr int kLinesQuorum = 40;
// idle lanes a wave gathers before it starts new packets, so the launch code runs for
// several lanes at once. On the MI355X: 512-plate sweep +3.5% (4 against 1; 16 loses 5%,
// 32 loses 30%), the skin plate without the absorbed map +1%, with it and the bone phantom
// unchanged.
constexpr int kDefaultRefill = 4;
// LDS a batched-walk block may spend on its waves' tallies of their current instances
// (kBS_COUNT + 2 nr doubles per wave): 93 doubles per wave with 40 radial bins, 11.9 KiB for
// a 1024-lane block. Past this (long Rd(r) tables) the waves add to global memory directly.
constexpr size_t kWaveAccBytes = 32768;
// LDS of a fluence-binning block's copy of the bins (8192 FP64 bins)
constexpr size_t kFluenceLdsBytes = 65536;

struct DevInfo {
    int n_cu = 0;
    size_t max_lds = kMaxLds;     // LDS one block may allocate (160 KiB on gfx950)
    size_t lds_per_cu = kMaxLds;  // LDS of one CU, shared by its resident blocks
    bool ok = false;
};

std::mutex g_mu;
DevInfo g_dev[64];

int device_info(DevInfo* out)
{
    int dev = 0;
    hipError_t e = hipGetDevice(&dev);
    if (e != hipSuccess) return hip_fail(e, "hipGetDevice");
    if (dev < 0 || dev >= 64) return fail(PMO_ERR_UNSUPPORTED, "device ordinal %d out of range", dev);
    std::lock_guard<std::mutex> lk(g_mu);
    if (!g_dev[dev].ok) {
        int n = 0;
        e = hipDeviceGetAttribute(&n, hipDeviceAttributeMultiprocessorCount, dev);
        if (e != hipSuccess) return hip_fail(e, "hipDeviceGetAttribute");
        g_dev[dev].n_cu = n;
        int lds = 0;
        e = hipDeviceGetAttribute(&lds, hipDeviceAttributeMaxSharedMemoryPerBlock, dev);
        if (e != hipSuccess) return hip_fail(e, "hipDeviceGetAttribute");
        if (lds > 0) g_dev[dev].max_lds = (size_t)lds;
        int lcu = 0;
        e = hipDeviceGetAttribute(&lcu, hipDeviceAttributeMaxSharedMemoryPerMultiprocessor, dev);
        if (e != hipSuccess) return hip_fail(e, "hipDeviceGetAttribute");
        g_dev[dev].lds_per_cu = lcu > 0 ? (size_t)lcu : g_dev[dev].max_lds;
        g_dev[dev].ok = true;
    }
    *out = g_dev[dev];
    return PMO_OK;
}

template <typename R>
int blocks_per_cu(int block, size_t lds, int* out, bool rz = false, bool skip = false, bool lines = false)
{
    int nb = 0;
    hipError_t e;
    if (lines)
        e = skip ? hipOccupancyMaxActiveBlocksPerMultiprocessor(&nb, pmo::walk_lines_kernel<true>, block, lds)
                 : hipOccupancyMaxActiveBlocksPerMultiprocessor(&nb, pmo::walk_lines_kernel<false>, block, lds);
    else if (skip)
        e = rz ? hipOccupancyMaxActiveBlocksPerMultiprocessor(&nb, pmo::walk_kernel<R, true, true>, block, lds)
               : hipOccupancyMaxActiveBlocksPerMultiprocessor(&nb, pmo::walk_kernel<R, false, true>, block, lds);
    else
        e = rz ? hipOccupancyMaxActiveBlocksPerMultiprocessor(&nb, pmo::walk_kernel<R, true, false>, block, lds)
               : hipOccupancyMaxActiveBlocksPerMultiprocessor(&nb, pmo::walk_kernel<R, false, false>, block, lds);
    if (e != hipSuccess) return hip_fail(e, "hipOccupancyMaxActiveBlocksPerMultiprocessor");
    *out = nb > 0 ? nb : 1;
    return PMO_OK;
}

size_t align_up(size_t x, size_t a) { return (x + a - 1) / a * a; }

}  // namespace

// Scratch of pmo_walk_noalloc: the walk's packet-claim counters in a static device array,
// so that entry point allocates nothing. Each device holds its own copy at its own
// address, looked up once per device with that device current; host code uses that
// address, never the array's name (a host-side stand-in).
static __device__ __align__(256) unsigned long long scratch[kWorkspaceBytes / sizeof(unsigned long long)];

namespace {

std::mutex g_scratch_mu;
void* g_scratch_ptr[64];

int scratch_address(void** out)
{
    int dev = 0;
    hipError_t e = hipGetDevice(&dev);
    if (e != hipSuccess) return hip_fail(e, "hipGetDevice");
    if (dev < 0 || dev >= 64) return fail(PMO_ERR_UNSUPPORTED, "device ordinal %d out of range", dev);
    std::lock_guard<std::mutex> lk(g_scratch_mu);
    if (!g_scratch_ptr[dev]) {
        void* p = nullptr;
        e = hipGetSymbolAddress(&p, HIP_SYMBOL(scratch));
        if (e != hipSuccess) return hip_fail(e, "hipGetSymbolAddress(scratch)");
        if (!p) return fail(PMO_ERR_HIP, "hipGetSymbolAddress(scratch) returned a null address");
        g_scratch_ptr[dev] = p;
    }
    *out = g_scratch_ptr[dev];
    return PMO_OK;
}

}  // namespace

extern "C" {

int pmo_abi_version(void) { return PMO_ABI_VERSION; }

const char* pmo_last_error(void) { return g_err; }

size_t pmo_workspace_bytes(void) { return kWorkspaceBytes; }

int pmo_device_info(int32_t* n_cu, int32_t* blocks_per_cu_walk, int32_t block_size)
{
    DevInfo d;
    int rc = device_info(&d);
    if (rc != PMO_OK) return rc;
    if (n_cu) *n_cu = d.n_cu;
    if (blocks_per_cu_walk) {
        const int bs = block_size > 0 ? block_size : kDefaultBlock;
        int nb = 0;
        rc = blocks_per_cu<float>(bs, 4096, &nb);
        if (rc != PMO_OK) return rc;
        *blocks_per_cu_walk = nb;
    }
    return PMO_OK;
}

int pmo_walk(const pmo_volume* vol, const pmo_source* src, const pmo_run* run, const pmo_tallies* tal,
             void* workspace, void* stream)
{
    g_err[0] = '\0';
    if (!vol || !src || !run || !tal) return fail(PMO_ERR_INVALID, "null argument");
    if (!vol->labels || !vol->optics) return fail(PMO_ERR_INVALID, "volume labels/optics pointer is null");
    if (!tal->scalars) return fail(PMO_ERR_INVALID, "tallies.scalars is required");
    if (!workspace) return fail(PMO_ERR_INVALID, "workspace is null");
    if (vol->nx < 3 || vol->ny < 3 || vol->nz < 3)
        return fail(PMO_ERR_INVALID, "volume extents (%d, %d, %d) must be >= 3", vol->nx, vol->ny, vol->nz);
    const unsigned long long nvox = (unsigned long long)vol->nx * vol->ny * vol->nz;
    if (nvox >= (1ull << 31)) return fail(PMO_ERR_UNSUPPORTED, "volume of %llu voxels exceeds 2^31", nvox);
    if (vol->nx >= (1 << 24) || vol->ny >= (1 << 24) || vol->nz >= (1 << 24))
        return fail(PMO_ERR_UNSUPPORTED, "volume extent >= 2^24");
    if (vol->n_labels < 1 || vol->n_labels > 255)
        return fail(PMO_ERR_LABEL, "n_labels = %d outside [1, 255]", vol->n_labels);
    if (!(vol->pitch > 0.f)) return fail(PMO_ERR_INVALID, "pitch must be > 0");
    if (src->kind < PMO_BEAM_PENCIL || src->kind > PMO_BEAM_STATES)
        return fail(PMO_ERR_INVALID, "unknown beam kind %d", src->kind);
    if (src->ix < 1 || src->ix > vol->nx - 2 || src->iy < 1 || src->iy > vol->ny - 2 || src->iz < 1 ||
        src->iz > vol->nz - 2)
        return fail(PMO_ERR_INVALID, "launch voxel (%d, %d, %d) is not inside the volume interior", src->ix,
                    src->iy, src->iz);
    if (src->kind == PMO_BEAM_STATES) {
        if (!src->add || !src->p || !src->v || !src->w)
            return fail(PMO_ERR_INVALID, "per-packet launch states need add, p, v and w");
        if (src->states_stride < run->n_packets)
            return fail(PMO_ERR_INVALID, "states_stride %llu < n_packets %llu",
                        (unsigned long long)src->states_stride, (unsigned long long)run->n_packets);
    } else {
        const float dn = src->dir[0] * src->dir[0] + src->dir[1] * src->dir[1] + src->dir[2] * src->dir[2];
        if (!(dn > 0.999f && dn < 1.001f)) return fail(PMO_ERR_INVALID, "launch direction is not a unit vector");
    }
    if (!(run->roulette_m >= 1.f)) return fail(PMO_ERR_INVALID, "roulette_m must be >= 1");
    if (vol->skip && !vol->label_radius)
        return fail(PMO_ERR_INVALID, "pmo_walk skips same-label faces from pmo_volume.label_radius: pack the labels "
                                     "with their radius (pmo_pack_label_radius) or pass skip = NULL");
    if (run->n_packets == 0) return PMO_OK;
    const bool whole = nvox <= kWholeVolumeCells;
    const bool fp64 = (run->flags & PMO_F_FP64) != 0u;
    // the absorbed map through cube lines (the FP32 walk only; the FP64 build and the r-z
    // fluence variant deposit voxel by voxel); a volume held whole in the window needs none
    const bool win_default = run->window[0] == 0 && run->window[1] == 0 && run->window[2] == 0;
    const bool lines = tal->absorbed != nullptr && tal->absorbed_lines != nullptr && !fp64 && !tal->fluence_rz &&
                       !(whole && win_default);
    const bool wide = run->n_packets >= kWideRunPackets && tal->absorbed != nullptr && !whole;
    const int bs = run->block_size > 0 ? run->block_size : (wide ? kWideBlock : kDefaultBlock);
    if (bs % 64 != 0 || bs > 1024) return fail(PMO_ERR_INVALID, "block_size %d must be a multiple of 64, <= 1024", bs);
    if (run->packets_per_lane > 4096) return fail(PMO_ERR_INVALID, "packets_per_lane %d too large", run->packets_per_lane);
    if (tal->nr < 0 || tal->na < 0 || tal->nl < 0 || tal->n_rings < 0)
        return fail(PMO_ERR_INVALID, "negative histogram size");
    if (tal->nr > 0 && !(tal->dr > 0.f)) return fail(PMO_ERR_INVALID, "dr must be > 0");
    if (tal->nl > 0 && !(tal->dl > 0.f)) return fail(PMO_ERR_INVALID, "dl must be > 0");
    if (tal->n_rings > 0 && !tal->rings) return fail(PMO_ERR_INVALID, "rings pointer is null");

    DevInfo d;
    int rc = device_info(&d);
    if (rc != PMO_OK) return rc;

    pmo::WalkParams P;
    memset(&P, 0, sizeof(P));
    P.labels = vol->labels;
    P.label_radius = vol->label_radius;
    P.optics = reinterpret_cast<const float4*>(vol->optics);
    P.nx = (uint32_t)vol->nx; P.ny = (uint32_t)vol->ny; P.nz = (uint32_t)vol->nz;
    P.nvox = (uint32_t)nvox;
    P.sx = P.ny * P.nz; P.sy = P.nz;
    P.n_labels = (uint32_t)vol->n_labels;
    P.cx = (vol->nx - 1) / 2; P.cy = (vol->ny - 1) / 2;
    P.pitch = vol->pitch;
    P.src_kind = src->kind; P.ix = src->ix; P.iy = src->iy; P.iz = src->iz;
    P.radius = src->radius; P.dirx = src->dir[0]; P.diry = src->dir[1]; P.dirz = src->dir[2];
    P.w0 = src->w0;
    if (src->kind == PMO_BEAM_STATES) {
        P.st_add = src->add; P.st_p = src->p; P.st_v = src->v; P.st_w = src->w;
        P.st_stride = src->states_stride;
    }
    P.n_packets = run->n_packets; P.first_packet = run->first_packet;
    P.seed = run->seed; P.stream = run->stream;
    P.wth = run->wth; P.roulette_m = run->roulette_m; P.inv_m = 1.f / run->roulette_m;
    P.flags = run->flags;
    P.batch = 64u;  // set below once the resident grid is known
    P.max_events = kMaxEvents;
    if (run->quorum < 0 || run->quorum > 64) return fail(PMO_ERR_INVALID, "quorum %d outside [0, 64]", run->quorum);
    // >= 1, so a wave whose lanes all wait for crossings still advances them
    P.quorum = run->quorum > 0 ? (uint32_t)run->quorum : (uint32_t)(lines ? kLinesQuorum : kDefaultQuorum);
    if (run->refill < 0 || run->refill > 64) return fail(PMO_ERR_INVALID, "refill %d outside [0, 64]", run->refill);
    P.refill = run->refill > 0 ? (uint32_t)run->refill : (uint32_t)kDefaultRefill;
    P.scalars = tal->scalars;
    P.events = reinterpret_cast<unsigned long long*>(tal->events);
    P.absorbed = tal->absorbed;
    P.nr = tal->nr; P.dr = tal->dr;
    P.rd_r = tal->nr > 0 ? tal->rd_r : nullptr;
    P.tt_r = tal->nr > 0 ? tal->tt_r : nullptr;
    P.rd_top_r = tal->nr > 0 ? tal->rd_top_r : nullptr;
    P.na = tal->na; P.da = tal->na > 0 ? 1.57079632679489662f / (float)tal->na : 1.f;
    P.rd_a = tal->na > 0 ? tal->rd_a : nullptr;
    P.tt_a = tal->na > 0 ? tal->tt_a : nullptr;
    P.nl = tal->nl; P.dl = tal->dl; P.line_y = tal->line_y;
    P.line_z = tal->nl > 0 ? tal->line_z : nullptr;
    P.n_rings = tal->n_rings; P.rings = tal->rings;
    P.ring_w = tal->n_rings > 0 ? tal->ring_w : nullptr;
    P.exit_state = tal->exit_state;
    P.exit_add = tal->exit_add;
    P.exit_p = tal->exit_p;
    if (tal->fluence_rz) {
        if (tal->rz_nr < 1 || tal->rz_nz < 1 || !(tal->rz_dr > 0.f) || !(tal->rz_dz > 0.f))
            return fail(PMO_ERR_INVALID, "r-z fluence needs rz_nr, rz_nz >= 1 and rz_dr, rz_dz > 0");
        if ((long long)tal->rz_nr * tal->rz_nz > (1 << 13))
            return fail(PMO_ERR_UNSUPPORTED, "r-z fluence of %d x %d bins exceeds 8192", tal->rz_nr, tal->rz_nz);
        P.rz_nr = tal->rz_nr; P.rz_nz = tal->rz_nz; P.rz_dr = tal->rz_dr; P.rz_dz = tal->rz_dz;
        P.fluence_rz = tal->fluence_rz;
    }
    P.work = reinterpret_cast<unsigned long long*>(workspace);
    P.key_dev = run->key_dev;

    // dynamic LDS: label table | block scalars | block events | histograms
    const size_t tab_bytes = (size_t)(vol->n_labels + 1) * (fp64 ? sizeof(pmo::LabelConst<double>)
                                                                  : sizeof(pmo::LabelConst<float>));
    P.lds_tab = 0;
    P.lds_scal = (uint32_t)align_up(tab_bytes, 16);
    P.lds_ev = P.lds_scal + (uint32_t)align_up(pmo::kBS_COUNT * sizeof(double), 16);
    P.lds_hist = P.lds_ev + (uint32_t)align_up(pmo::kBE_COUNT * sizeof(unsigned long long), 16);
    uint32_t off = 0;
    P.h_rd = off; off += P.rd_r ? (uint32_t)P.nr : 0u;
    P.h_tt = off; off += P.tt_r ? (uint32_t)P.nr : 0u;
    P.h_top = off; off += P.rd_top_r ? (uint32_t)P.nr : 0u;
    P.h_rda = off; off += P.rd_a ? (uint32_t)P.na : 0u;
    P.h_tta = off; off += P.tt_a ? (uint32_t)P.na : 0u;
    P.h_line = off; off += P.line_z ? (uint32_t)P.nl : 0u;
    P.h_ring = off; off += P.ring_w ? (uint32_t)P.n_rings : 0u;
    P.h_rz = off; off += P.fluence_rz ? (uint32_t)(P.rz_nr * P.rz_nz) : 0u;
    P.h_total = off;
    // line buffers (8 doubles per lane) and their hand-over stage (16 words per wave)
    P.lines = lines ? tal->absorbed_lines : nullptr;
    P.lbuf_stride = (uint32_t)bs;
    P.nly = (uint32_t)(vol->ny + 1) / 2u;
    P.nlz = (uint32_t)(vol->nz + 1) / 2u;
    P.lds_lbuf = (uint32_t)align_up(P.lds_hist + (size_t)P.h_total * sizeof(double), 16);
    P.lds_stage = P.lds_lbuf + (lines ? (uint32_t)(8u * (uint32_t)bs * sizeof(double)) : 0u);
    P.lds_win = P.lds_stage + (lines ? (uint32_t)((uint32_t)bs / 64u * 16u * sizeof(int32_t)) : 0u);
    if (P.lds_win > d.max_lds)
        return fail(PMO_ERR_UNSUPPORTED, "histograms and line buffers need %u B of LDS (limit %zu)", P.lds_win,
                    d.max_lds);

    // LDS window of the absorbed tally: the launch voxel region, where deposits collide. A
    // default window shrinks (depth first) to the LDS the histograms leave; one the caller
    // sizes must fit as given.
    const int* wdef = lines ? (bs >= kWideBlock ? kLinesWideWindow : kLinesWindow)
                            : (bs >= kWideBlock ? kWideWindow : kDefaultWindow);
    int wdim[3];
    const int ext[3] = {vol->nx, vol->ny, vol->nz};
    for (int a = 0; a < 3; ++a) {
        wdim[a] = run->window[a] == 0 ? (whole ? ext[a] : wdef[a]) : run->window[a];
        if (wdim[a] < 0 || !P.absorbed) wdim[a] = 0;
    }
    for (int a = 0; a < 3; ++a)
        if (wdim[a] > ext[a]) wdim[a] = ext[a];
    size_t win_budget = (d.max_lds - P.lds_win) / sizeof(double);
    if (lines && win_default) {
        // keep as many blocks per CU as the registers allow: each gets its share of the CU's
        // LDS, the line buffers first, the window from the rest
        int nb_reg = 0;
        rc = blocks_per_cu<float>(bs, 0, &nb_reg, false, P.label_radius != nullptr, true);
        if (rc != PMO_OK) return rc;
        const size_t share = d.lds_per_cu / (size_t)(nb_reg > 0 ? nb_reg : 1);
        const size_t cells = share > P.lds_win ? (share - P.lds_win) / sizeof(double) : 0;
        if (cells < win_budget) win_budget = cells;
    }
    if (win_default)
        while ((size_t)wdim[0] * wdim[1] * wdim[2] > win_budget) {
            if (wdim[2] > wdim[0] / 2 && wdim[2] > 1) --wdim[2];
            else if (wdim[0] > 1) { --wdim[0]; wdim[1] = wdim[1] > 1 ? wdim[1] - 1 : 1; }
            else { wdim[2] = 0; break; }
        }
    if (wdim[0] > 0 && wdim[1] > 0 && wdim[2] > 0) {
        const int ctr[3] = {src->ix, src->iy, src->iz};
        int org[3];
        for (int a = 0; a < 3; ++a) {
            org[a] = a < 2 ? ctr[a] - wdim[a] / 2 : ctr[a];
            if (org[a] < 0) org[a] = 0;
            if (org[a] + wdim[a] > ext[a]) org[a] = ext[a] - wdim[a];
        }
        P.win_x0 = org[0]; P.win_y0 = org[1]; P.win_z0 = org[2];
        P.win_nx = (uint32_t)wdim[0]; P.win_ny = (uint32_t)wdim[1]; P.win_nz = (uint32_t)wdim[2];
    } else {
        P.win_x0 = P.win_y0 = P.win_z0 = 0;
        P.win_nx = P.win_ny = P.win_nz = 0;
    }
    P.win_cells = P.win_nx * P.win_ny * P.win_nz;
    const size_t lds = P.lds_win + (size_t)P.win_cells * sizeof(double);
    if (lds > d.max_lds)
        return fail(PMO_ERR_UNSUPPORTED, "histograms and tally window need %zu B of LDS (limit %zu)", lds, d.max_lds);

    int nb = 0;
    const bool skip = P.label_radius != nullptr;
    rc = fp64 ? blocks_per_cu<double>(bs, lds, &nb, P.fluence_rz != nullptr, skip)
              : blocks_per_cu<float>(bs, lds, &nb, P.fluence_rz != nullptr, skip, lines);
    if (rc != PMO_OK) return rc;
    unsigned long long grid = run->grid_blocks > 0 ? (unsigned long long)run->grid_blocks
                                                   : (unsigned long long)nb * (unsigned long long)d.n_cu;
    // packets a wave claims per counter atomic: the caller's choice, else as many as keep at
    // least four claims per resident wave (so small runs still spread over the whole grid)
    int ppl = run->packets_per_lane;
    if (ppl <= 0) {
        const unsigned long long waves = grid * (unsigned long long)(bs / 64);
        const unsigned long long per = run->n_packets / (64ull * 4ull * (waves > 0 ? waves : 1ull));
        ppl = per < 1 ? 1 : (per > (unsigned long long)kDefaultPacketsPerLane ? kDefaultPacketsPerLane : (int)per);
    }
    P.batch = (uint32_t)(64 * ppl);
    const unsigned long long need_blocks = (run->n_packets + P.batch - 1) / P.batch;
    const unsigned long long waves_per_block = (unsigned long long)bs / 64;
    const unsigned long long useful = (need_blocks + waves_per_block - 1) / waves_per_block;
    if (grid > useful) grid = useful;
    if (grid < 1) grid = 1;

    hipStream_t s = reinterpret_cast<hipStream_t>(stream);
    hipLaunchKernelGGL(pmo::zero_u64_kernel, dim3(1), dim3(64), 0, s, P.work, 4);
    hipError_t e = hipGetLastError();
    if (e != hipSuccess) return hip_fail(e, "zero workspace launch");
    const unsigned long long n_cubes = (unsigned long long)((vol->nx + 1) / 2) * P.nly * P.nlz;
    const unsigned fill_grid = (unsigned)((n_cubes * 8ull + 255ull) / 256ull < 8192ull ? (n_cubes * 8ull + 255ull) / 256ull
                                                                                        : 8192ull);
    if (lines) {
        hipLaunchKernelGGL(pmo::zero_f64_kernel, dim3(fill_grid), dim3(256), 0, s, P.lines, n_cubes * 8ull);
        e = hipGetLastError();
        if (e != hipSuccess) return hip_fail(e, "zero line scratch launch");
    }
    const dim3 gd((unsigned)grid), bd(bs);
    const bool rz = P.fluence_rz != nullptr;
    if (fp64) {
        if (skip) {
            if (rz) hipLaunchKernelGGL((pmo::walk_kernel<double, true, true>), gd, bd, lds, s, P);
            else hipLaunchKernelGGL((pmo::walk_kernel<double, false, true>), gd, bd, lds, s, P);
        } else {
            if (rz) hipLaunchKernelGGL((pmo::walk_kernel<double, true, false>), gd, bd, lds, s, P);
            else hipLaunchKernelGGL((pmo::walk_kernel<double, false, false>), gd, bd, lds, s, P);
        }
    } else if (lines) {
        if (skip) hipLaunchKernelGGL(pmo::walk_lines_kernel<true>, gd, bd, lds, s, P);
        else hipLaunchKernelGGL(pmo::walk_lines_kernel<false>, gd, bd, lds, s, P);
    } else {
        if (skip) {
            if (rz) hipLaunchKernelGGL((pmo::walk_kernel<float, true, true>), gd, bd, lds, s, P);
            else hipLaunchKernelGGL((pmo::walk_kernel<float, false, true>), gd, bd, lds, s, P);
        } else {
            if (rz) hipLaunchKernelGGL((pmo::walk_kernel<float, true, false>), gd, bd, lds, s, P);
            else hipLaunchKernelGGL((pmo::walk_kernel<float, false, false>), gd, bd, lds, s, P);
        }
    }
    e = hipGetLastError();
    if (e != hipSuccess) return hip_fail(e, "walk launch");
    if (lines) {
        const unsigned g = (unsigned)((n_cubes + 255ull) / 256ull < 8192ull ? (n_cubes + 255ull) / 256ull : 8192ull);
        hipLaunchKernelGGL(pmo::lines_to_absorbed_kernel, dim3(g), dim3(256), 0, s, (const double*)P.lines, P.absorbed,
                           P.nx, P.ny, P.nz, P.nly, P.nlz, n_cubes);
        e = hipGetLastError();
        if (e != hipSuccess) return hip_fail(e, "line scratch transfer launch");
    }
    return PMO_OK;
}

int pmo_walk_noalloc(const pmo_volume* vol, const pmo_source* src, const pmo_run* run, const pmo_tallies* tal,
                     void* stream)
{
    g_err[0] = '\0';
    // the walk's scratch is its packet-claim counter block, the same 256 B at every size
    if (pmo_workspace_bytes() > sizeof(scratch))
        return fail(PMO_ERR_UNSUPPORTED, "the walk needs %zu B of scratch, the static scratch holds %zu B",
                    pmo_workspace_bytes(), sizeof(scratch));
    void* ws = nullptr;
    int rc = scratch_address(&ws);
    if (rc != PMO_OK) return rc;
    return pmo_walk(vol, src, run, tal, ws, stream);
}

int pmo_pack_label_radius(const uint8_t* labels, const uint8_t* radius, uint16_t* out, int64_t n, void* stream)
{
    g_err[0] = '\0';
    if (n < 0) return fail(PMO_ERR_INVALID, "n < 0");
    if (n == 0) return PMO_OK;
    if (!labels || !radius || !out) return fail(PMO_ERR_INVALID, "labels, radius and out must be non-null");
    const unsigned long long blocks = ((unsigned long long)n + 255ull) / 256ull;
    hipLaunchKernelGGL(pmo::pack_label_radius_kernel, dim3((unsigned)(blocks < 65536ull ? blocks : 65536ull)), dim3(256),
                       0, reinterpret_cast<hipStream_t>(stream), labels, radius, out, (unsigned long long)n);
    hipError_t e = hipGetLastError();
    return e == hipSuccess ? PMO_OK : hip_fail(e, "pack launch");
}

int pmo_skip_radius(const uint8_t* labels, int32_t nx, int32_t ny, int32_t nz, int32_t cap, uint8_t* radius,
                    uint8_t* scratch, void* stream)
{
    g_err[0] = '\0';
    if (!labels || !radius || !scratch) return fail(PMO_ERR_INVALID, "labels, radius and scratch must be non-null");
    if ((const void*)radius == (const void*)scratch || (const void*)labels == (const void*)radius ||
        (const void*)labels == (const void*)scratch)
        return fail(PMO_ERR_INVALID, "labels, radius and scratch must be distinct buffers");
    if (nx < 1 || ny < 1 || nz < 1) return fail(PMO_ERR_INVALID, "volume extents must be >= 1");
    const unsigned long long n = (unsigned long long)nx * ny * nz;
    if (n >= (1ull << 31)) return fail(PMO_ERR_UNSUPPORTED, "volume of %llu voxels exceeds 2^31", n);
    if (cap < 1 || cap > PMO_SKIP_MAX_CAP) return fail(PMO_ERR_INVALID, "cap %d outside [1, %d]", cap, PMO_SKIP_MAX_CAP);
    hipStream_t s = reinterpret_cast<hipStream_t>(stream);
    const uint32_t un = (uint32_t)n, c = (uint32_t)cap;
    const uint32_t X = (uint32_t)nx, Y = (uint32_t)ny, Z = (uint32_t)nz;
    const unsigned grid = (unsigned)((n + 255) / 256 < 65536 ? (n + 255) / 256 : 65536);
    const uint32_t bs = pmo::kSkipSweepBlock;
    // a sweep along an axis of length len and stride S over n_outer outer slices, in launches
    // of at most 65535 slices (the grid's y limit)
    auto sweep = [&](bool seed, const uint8_t* in, uint8_t* out, uint32_t n_outer, uint32_t len, uint32_t S,
                     uint32_t flag) {
        for (uint32_t o = 0; o < n_outer; o += 65535u) {
            const uint32_t m = n_outer - o < 65535u ? n_outer - o : 65535u;
            const dim3 g((S + bs - 1) / bs, m);
            if (seed)
                hipLaunchKernelGGL(pmo::skip_seed_sweep_kernel, g, dim3(bs), 0, s, in, out, o, len, S, c, flag);
            else
                hipLaunchKernelGGL(pmo::skip_minmax_sweep_kernel, g, dim3(bs), (size_t)(2u * c + 1u) * bs, s, in, out,
                                   o, len, S, c, flag);
        }
    };
    // B: w_z -> scratch; w_zy -> radius; b (from w_zyx) -> scratch
    hipLaunchKernelGGL(pmo::skip_seed_z_kernel, dim3(grid), dim3(256), 0, s, labels, un, Z, scratch);
    sweep(true, scratch, radius, X, Y, Z, 0u);
    sweep(true, radius, scratch, 1u, X, Y * Z, 1u);
    // min-max along z -> radius; along y -> scratch; along x -> radius (clamped)
    hipLaunchKernelGGL(pmo::skip_pass_kernel, dim3(grid), dim3(256), 0, s, (const uint8_t*)scratch, radius, un, Z, 1u,
                       c, 0u);
    sweep(false, radius, scratch, X, Y, Z, 0u);
    sweep(false, scratch, radius, 1u, X, Y * Z, 1u);
    hipError_t e = hipGetLastError();
    if (e != hipSuccess) return hip_fail(e, "skip radius launch");
    return PMO_OK;
}

int pmo_probe_atomics(void* buf, int64_t n, int32_t blocks, int32_t iters, int32_t kind, void* stream)
{
    g_err[0] = '\0';
    if (!buf || n < 1 || n >= (1ll << 32) || blocks < 1 || iters < 1) return fail(PMO_ERR_INVALID, "bad probe size");
    hipStream_t s = reinterpret_cast<hipStream_t>(stream);
    const uint32_t m = (uint32_t)n;
    switch (kind) {
    case PMO_PROBE_F64:
        hipLaunchKernelGGL(pmo::probe_atomics_kernel<double>, dim3(blocks), dim3(256), 0, s, (double*)buf, m, iters, 0);
        break;
    case PMO_PROBE_F32:
        hipLaunchKernelGGL(pmo::probe_atomics_kernel<float>, dim3(blocks), dim3(256), 0, s, (float*)buf, m, iters, 0);
        break;
    case PMO_PROBE_U64:
        hipLaunchKernelGGL(pmo::probe_atomics_kernel<unsigned long long>, dim3(blocks), dim3(256), 0, s,
                           (unsigned long long*)buf, m, iters, 0);
        break;
    case PMO_PROBE_U32:
        hipLaunchKernelGGL(pmo::probe_atomics_kernel<uint32_t>, dim3(blocks), dim3(256), 0, s, (uint32_t*)buf, m, iters, 0);
        break;
    case PMO_PROBE_F64_PER_XCD:
        hipLaunchKernelGGL(pmo::probe_atomics_kernel<double>, dim3(blocks), dim3(256), 0, s, (double*)buf, m, iters, 1);
        break;
    case PMO_PROBE_F64_PER_XCD_WG:
        hipLaunchKernelGGL((pmo::probe_atomics_kernel<double, true>), dim3(blocks), dim3(256), 0, s, (double*)buf, m,
                           iters, 1);
        break;
    case PMO_PROBE_F64_LINES:
    case PMO_PROBE_F64_BLOCKS_32B:
    case PMO_PROBE_F64_BLOCKS_64B:
    case PMO_PROBE_F64_BLOCKS_128B: {
        const uint32_t lg = kind == PMO_PROBE_F64_LINES ? 6u : (uint32_t)(kind - PMO_PROBE_F64_BLOCKS_32B) + 2u;
        if (m < (1u << lg)) return fail(PMO_ERR_INVALID, "this probe needs n >= %u", 1u << lg);
        hipLaunchKernelGGL(pmo::probe_atomics_lines_kernel, dim3(blocks), dim3(256), 0, s, (double*)buf, m, iters, lg);
        break;
    }
    default:
        return fail(PMO_ERR_INVALID, "unknown probe kind");
    }
    hipError_t e = hipGetLastError();
    return e == hipSuccess ? PMO_OK : hip_fail(e, "probe launch");
}

int pmo_probe_valu(float* out, int32_t blocks, int32_t iters, void* stream)
{
    g_err[0] = '\0';
    if (!out || blocks < 1 || iters < 1) return fail(PMO_ERR_INVALID, "bad probe size");
    hipLaunchKernelGGL(pmo::probe_valu_kernel, dim3(blocks), dim3(256), 0, reinterpret_cast<hipStream_t>(stream),
                       out, iters);
    hipError_t e = hipGetLastError();
    return e == hipSuccess ? PMO_OK : hip_fail(e, "probe launch");
}

int pmo_probe_philox(uint32_t* out, int32_t blocks, int32_t iters, void* stream)
{
    g_err[0] = '\0';
    if (!out || blocks < 1 || iters < 1) return fail(PMO_ERR_INVALID, "bad probe size");
    hipLaunchKernelGGL(pmo::probe_philox_kernel, dim3(blocks), dim3(256), 0, reinterpret_cast<hipStream_t>(stream),
                       out, iters, 1u);
    hipError_t e = hipGetLastError();
    return e == hipSuccess ? PMO_OK : hip_fail(e, "probe launch");
}

int pmo_philox(const uint32_t* ctr, uint32_t key0, uint32_t key1, uint32_t* out, int64_t n, void* stream)
{
    g_err[0] = '\0';
    if (n < 0) return fail(PMO_ERR_INVALID, "n < 0");
    if (n == 0) return PMO_OK;
    if (!ctr || !out) return fail(PMO_ERR_INVALID, "null pointer");
    const unsigned blocks = (unsigned)((n + 255) / 256);
    hipLaunchKernelGGL(pmo::philox_kernel, dim3(blocks), dim3(256), 0, reinterpret_cast<hipStream_t>(stream),
                       reinterpret_cast<const uint4*>(ctr), key0, key1, reinterpret_cast<uint4*>(out), n);
    hipError_t e = hipGetLastError();
    if (e != hipSuccess) return hip_fail(e, "philox launch");
    return PMO_OK;
}

}  // extern "C"

namespace {

int diff_launch(const pmo_volume* vol, const pmo_source* src, const pmo_run* run, double* scalars,
                uint64_t* events, const pmo_diff* diff, void* workspace, void* stream, int mode)
{
    g_err[0] = '\0';
    if (!vol || !src || !run || !diff || !workspace) return fail(PMO_ERR_INVALID, "null argument");
    if (!vol->labels || !vol->optics) return fail(PMO_ERR_INVALID, "volume labels/optics pointer is null");
    if (vol->nx < 3 || vol->ny < 3 || vol->nz < 3) return fail(PMO_ERR_INVALID, "volume extents must be >= 3");
    const unsigned long long nvox = (unsigned long long)vol->nx * vol->ny * vol->nz;
    if (nvox >= (1ull << 31) || vol->nx >= (1 << 24) || vol->ny >= (1 << 24))
        return fail(PMO_ERR_UNSUPPORTED, "volume too large");
    if (vol->n_labels < 1 || vol->n_labels > 255) return fail(PMO_ERR_LABEL, "n_labels outside [1, 255]");
    if (!(vol->pitch > 0.f)) return fail(PMO_ERR_INVALID, "pitch must be > 0");
    if (src->kind < PMO_BEAM_PENCIL || src->kind > PMO_BEAM_GAUSS) return fail(PMO_ERR_INVALID, "unknown beam kind");
    if (src->ix < 1 || src->ix > vol->nx - 2 || src->iy < 1 || src->iy > vol->ny - 2 || src->iz < 1 ||
        src->iz > vol->nz - 2)
        return fail(PMO_ERR_INVALID, "launch voxel is not inside the volume interior");
    if (!(run->roulette_m >= 1.f)) return fail(PMO_ERR_INVALID, "roulette_m must be >= 1");
    if ((run->flags & PMO_F_FP64) != 0u) return fail(PMO_ERR_UNSUPPORTED, "the differentiable walk runs in FP32");
    if (run->first_packet != 0) return fail(PMO_ERR_UNSUPPORTED, "first_packet must be 0");
    if (run->n_packets >= (1ull << 32)) return fail(PMO_ERR_UNSUPPORTED, "n_packets >= 2^32");
    if (diff->nr < 1 || !(diff->dr > 0.f)) return fail(PMO_ERR_INVALID, "need nr >= 1 radial bins and dr > 0");
    if (mode == 0 && (!scalars || !diff->rd_r)) return fail(PMO_ERR_INVALID, "forward needs scalars and rd_r");
    if (diff->jac && vol->n_labels > pmo::kMaxJacLabels)
        return fail(PMO_ERR_UNSUPPORTED, "per-label sums support at most %d labels", pmo::kMaxJacLabels);
    if (mode == 1 && (!diff->exit_rec || !diff->coef || !diff->grad_mua))
        return fail(PMO_ERR_INVALID, "adjoint needs exit_rec, coef and grad_mua");
    if (run->n_packets == 0) return PMO_OK;
    // the replay deposits the per-voxel gradient as the walk deposits the absorbed map, so
    // large replays take the same wide launch and LDS box
    const bool wide = mode == 1 && run->n_packets >= kWideRunPackets;
    const int bs = run->block_size > 0 ? run->block_size : (wide ? kWideBlock : kDefaultBlock);
    if (bs % 64 != 0 || bs > 1024) return fail(PMO_ERR_INVALID, "block_size must be a multiple of 64, <= 1024");
    DevInfo d;
    int rc = device_info(&d);
    if (rc != PMO_OK) return rc;

    pmo::DiffParams P;
    memset(&P, 0, sizeof(P));
    P.labels = vol->labels;
    // a forward without a per-voxel field may cross the faces inside same-label cubes in one
    // step (vol->skip); the replay and field forwards step voxel by voxel
    const bool skip = mode == 0 && vol->skip != nullptr && diff->mua_field == nullptr;
    P.skip = skip ? vol->skip : nullptr;
    P.optics = reinterpret_cast<const float4*>(vol->optics);
    P.mua_field = diff->mua_field;
    P.nx = (uint32_t)vol->nx; P.ny = (uint32_t)vol->ny; P.nz = (uint32_t)vol->nz;
    P.n_labels = (uint32_t)vol->n_labels;
    P.cx = (vol->nx - 1) / 2; P.cy = (vol->ny - 1) / 2;
    P.pitch = vol->pitch;
    P.src_kind = src->kind; P.ix = src->ix; P.iy = src->iy; P.iz = src->iz;
    P.radius = src->radius; P.dirx = src->dir[0]; P.diry = src->dir[1]; P.dirz = src->dir[2];
    P.w0 = src->w0;
    P.n_packets = run->n_packets;
    P.seed = run->seed; P.stream = run->stream;
    P.wth = run->wth; P.roulette_m = run->roulette_m; P.inv_m = 1.f / run->roulette_m;
    P.flags = run->flags;
    P.max_events = kMaxEvents;
    if (run->quorum < 0 || run->quorum > 64) return fail(PMO_ERR_INVALID, "quorum %d outside [0, 64]", run->quorum);
    P.quorum = run->quorum > 0 ? (uint32_t)run->quorum : (uint32_t)kDefaultQuorum;
    P.mode = mode;
    P.scalars = scalars;
    P.events = reinterpret_cast<unsigned long long*>(events);
    P.nr = diff->nr; P.dr = diff->dr;
    P.rd_r = diff->rd_r;
    P.jac = mode == 0 ? diff->jac : nullptr;
    P.exit_rec = reinterpret_cast<int2*>(diff->exit_rec);
    P.coef = diff->coef;
    P.grad = diff->grad_mua;
    P.mismatch = reinterpret_cast<unsigned long long*>(diff->mismatch);
    P.work = reinterpret_cast<unsigned long long*>(workspace);
    if (mode == 1 && !P.mismatch) P.mismatch = P.work + 2;   // scratch counter in the workspace
    // replay deposits through cube lines: a deposit's slot is 8 x cube + element, which must
    // stay below 2^31 (else the per-voxel atomics)
    P.nly = (uint32_t)((vol->ny + 1) / 2);
    P.nlz = (uint32_t)((vol->nz + 1) / 2);
    const unsigned long long n_cubes = (unsigned long long)((vol->nx + 1) / 2) * P.nly * P.nlz;
    const bool lines = mode == 1 && diff->grad_lines != nullptr && n_cubes * 8ull < (1ull << 31);
    P.lines = lines ? diff->grad_lines : nullptr;
    P.lbuf_stride = (uint32_t)bs;
    const uint32_t njac = P.jac ? P.n_labels : 0u;
    P.lds_tab = 0;
    P.lds_scal = (uint32_t)align_up((size_t)(vol->n_labels + 1) * sizeof(pmo::LabelConst<float>), 16);
    P.lds_ev = P.lds_scal + (uint32_t)align_up(pmo::kBS_COUNT * sizeof(double), 16);
    P.lds_hist = P.lds_ev + (uint32_t)align_up(pmo::kBE_COUNT * sizeof(unsigned long long), 16);
    P.lds_jac = P.lds_hist + (uint32_t)align_up((size_t)P.nr * sizeof(double), 16);
    P.lds_lbuf = P.lds_jac + (uint32_t)align_up((size_t)P.nr * njac * 2u * sizeof(double), 16);
    P.lds_stage = P.lds_lbuf + (lines ? (uint32_t)(8u * (uint32_t)bs * sizeof(double)) : 0u);
    P.lds_win = P.lds_stage + (lines ? (uint32_t)((uint32_t)bs / 64u * 16u * sizeof(int32_t)) : 0u);
    if (P.lds_win > d.max_lds) return fail(PMO_ERR_UNSUPPORTED, "differentiable walk needs %u B of LDS", P.lds_win);
    // replay: LDS box of the gradient map around the launch voxel (as for the absorbed
    // tally), shrunk depth first to the LDS left
    if (mode == 1) {
        const int* wdef = bs >= kWideBlock ? (lines ? kLinesWideWindow : kWideWindow) : kDefaultWindow;
        int wdim[3] = {wdef[0], wdef[1], wdef[2]};
        const int ext[3] = {vol->nx, vol->ny, vol->nz};
        const int ctr[3] = {src->ix, src->iy, src->iz};
        for (int a = 0; a < 3; ++a)
            if (wdim[a] > ext[a]) wdim[a] = ext[a];
        const size_t budget = (d.max_lds - P.lds_win) / sizeof(double);
        while ((size_t)wdim[0] * wdim[1] * wdim[2] > budget) {
            if (wdim[2] > wdim[0] / 2 && wdim[2] > 1) --wdim[2];
            else if (wdim[0] > 1) { --wdim[0]; wdim[1] = wdim[1] > 1 ? wdim[1] - 1 : 1; }
            else { wdim[2] = 0; break; }
        }
        if (wdim[2] > 0) {
            int org[3];
            for (int a = 0; a < 3; ++a) {
                org[a] = a < 2 ? ctr[a] - wdim[a] / 2 : ctr[a];
                if (org[a] < 0) org[a] = 0;
                if (org[a] + wdim[a] > ext[a]) org[a] = ext[a] - wdim[a];
            }
            P.win_x0 = org[0]; P.win_y0 = org[1]; P.win_z0 = org[2];
            P.win_nx = (uint32_t)wdim[0]; P.win_ny = (uint32_t)wdim[1]; P.win_nz = (uint32_t)wdim[2];
        }
    }
    P.win_cells = P.win_nx * P.win_ny * P.win_nz;
    const size_t lds = P.lds_win + (size_t)P.win_cells * sizeof(double);
    if (lds > d.max_lds) return fail(PMO_ERR_UNSUPPORTED, "differentiable walk needs %zu B of LDS", lds);

    int nb = 0;
    hipError_t e =
        lines  ? hipOccupancyMaxActiveBlocksPerMultiprocessor(&nb, pmo::diff_kernel<float, true, false>, bs, lds)
        : skip ? hipOccupancyMaxActiveBlocksPerMultiprocessor(&nb, pmo::diff_kernel<float, false, true>, bs, lds)
               : hipOccupancyMaxActiveBlocksPerMultiprocessor(&nb, pmo::diff_kernel<float, false, false>, bs, lds);
    if (e != hipSuccess) return hip_fail(e, "hipOccupancyMaxActiveBlocksPerMultiprocessor");
    if (nb < 1) nb = 1;
    unsigned long long grid = run->grid_blocks > 0 ? (unsigned long long)run->grid_blocks
                                                   : (unsigned long long)nb * (unsigned long long)d.n_cu;
    int ppl = run->packets_per_lane;
    if (ppl <= 0) {
        const unsigned long long waves = grid * (unsigned long long)(bs / 64);
        const unsigned long long per = run->n_packets / (64ull * 4ull * (waves > 0 ? waves : 1ull));
        ppl = per < 1 ? 1 : (per > (unsigned long long)kDefaultPacketsPerLane ? kDefaultPacketsPerLane : (int)per);
    }
    P.batch = (uint32_t)(64 * ppl);
    const unsigned long long need_blocks = (run->n_packets + P.batch - 1) / P.batch;
    const unsigned long long wpb = (unsigned long long)bs / 64;
    const unsigned long long useful = (need_blocks + wpb - 1) / wpb;
    if (grid > useful) grid = useful;
    if (grid < 1) grid = 1;
    hipStream_t s = reinterpret_cast<hipStream_t>(stream);
    hipLaunchKernelGGL(pmo::zero_u64_kernel, dim3(1), dim3(64), 0, s, P.work, 4);
    e = hipGetLastError();
    if (e != hipSuccess) return hip_fail(e, "zero workspace launch");
    const unsigned cube_grid = (unsigned)((n_cubes + 255ull) / 256ull < 8192ull ? (n_cubes + 255ull) / 256ull : 8192ull);
    if (lines) {
        hipLaunchKernelGGL(pmo::zero_f64_kernel, dim3(cube_grid), dim3(256), 0, s, P.lines, n_cubes * 8ull);
        e = hipGetLastError();
        if (e != hipSuccess) return hip_fail(e, "zero line scratch launch");
    }
    if (lines) hipLaunchKernelGGL((pmo::diff_kernel<float, true, false>), dim3((unsigned)grid), dim3(bs), lds, s, P);
    else if (skip) hipLaunchKernelGGL((pmo::diff_kernel<float, false, true>), dim3((unsigned)grid), dim3(bs), lds, s, P);
    else hipLaunchKernelGGL((pmo::diff_kernel<float, false, false>), dim3((unsigned)grid), dim3(bs), lds, s, P);
    e = hipGetLastError();
    if (e != hipSuccess) return hip_fail(e, "differentiable walk launch");
    if (lines) {
        hipLaunchKernelGGL(pmo::lines_to_absorbed_kernel, dim3(cube_grid), dim3(256), 0, s, (const double*)P.lines,
                           P.grad, P.nx, P.ny, P.nz, P.nly, P.nlz, n_cubes);
        e = hipGetLastError();
        if (e != hipSuccess) return hip_fail(e, "line scratch transfer launch");
    }
    return PMO_OK;
}

}  // namespace

extern "C" {

int pmo_walk_forward_diff(const pmo_volume* vol, const pmo_source* src, const pmo_run* run, double* scalars,
                          uint64_t* events, const pmo_diff* diff, void* workspace, void* stream)
{
    return diff_launch(vol, src, run, scalars, events, diff, workspace, stream, 0);
}

int pmo_walk_adjoint(const pmo_volume* vol, const pmo_source* src, const pmo_run* run, const pmo_diff* diff,
                     void* workspace, void* stream)
{
    return diff_launch(vol, src, run, nullptr, nullptr, diff, workspace, stream, 1);
}

int64_t pmo_plate_schedule(const uint64_t* n_packets, int32_t n_inst, int32_t batch, uint32_t* sched, int64_t cap)
{
    g_err[0] = '\0';
    if (!n_packets || n_inst < 1) return fail(PMO_ERR_INVALID, "need at least one instance");
    if (batch < 1) return fail(PMO_ERR_INVALID, "batch must be >= 1");
    if (cap > 0 && !sched) return fail(PMO_ERR_INVALID, "sched is null");
    int64_t total = 0;
    uint64_t max_rounds = 0;
    for (int32_t i = 0; i < n_inst; ++i) {
        if (n_packets[i] >= (1ull << 32)) return fail(PMO_ERR_UNSUPPORTED, "instance %d budget >= 2^32", i);
        const uint64_t r = (n_packets[i] + (uint64_t)batch - 1) / (uint64_t)batch;
        total += (int64_t)r;
        if (r > max_rounds) max_rounds = r;
    }
    if (cap <= 0) return total;
    int64_t m = 0;
    for (uint64_t round = 0; round < max_rounds && m < cap; ++round) {
        for (int32_t i = 0; i < n_inst && m < cap; ++i) {
            const uint64_t first = round * (uint64_t)batch;
            if (first >= n_packets[i]) continue;
            const uint64_t left = n_packets[i] - first;
            sched[4 * m + 0] = (uint32_t)i;
            sched[4 * m + 1] = (uint32_t)first;
            sched[4 * m + 2] = (uint32_t)(left < (uint64_t)batch ? left : (uint64_t)batch);
            sched[4 * m + 3] = 0u;
            ++m;
        }
    }
    return total;
}

int pmo_walk_plates(const pmo_plate_instance* inst, int32_t n_inst, const uint32_t* schedule, int64_t n_batches,
                    const pmo_run* run, const pmo_batch_tallies* tal, void* workspace, void* stream)
{
    g_err[0] = '\0';
    static_assert(sizeof(pmo_plate_instance) == sizeof(pmo::PlateInstance), "instance layout");
    if (!inst || !schedule || !run || !tal || !workspace) return fail(PMO_ERR_INVALID, "null argument");
    if (n_inst < 1) return fail(PMO_ERR_INVALID, "n_inst must be >= 1");
    if (n_batches < 0) return fail(PMO_ERR_INVALID, "n_batches < 0");
    if (!tal->scalars) return fail(PMO_ERR_INVALID, "tallies.scalars is required");
    if (tal->nr < 0 || (tal->nr > 0 && !(tal->dr > 0.f))) return fail(PMO_ERR_INVALID, "bad radial bins");
    if (!(run->roulette_m >= 1.f)) return fail(PMO_ERR_INVALID, "roulette_m must be >= 1");
    if ((run->flags & PMO_F_FP64) != 0u) return fail(PMO_ERR_UNSUPPORTED, "the batched walk runs in FP32 only");
    if (n_batches == 0) return PMO_OK;
    const int bs = run->block_size > 0 ? run->block_size : kDefaultBlock;
    if (bs % 64 != 0 || bs > 1024) return fail(PMO_ERR_INVALID, "block_size %d must be a multiple of 64, <= 1024", bs);
    DevInfo d;
    int rc = device_info(&d);
    if (rc != PMO_OK) return rc;

    pmo::BatchParams P;
    memset(&P, 0, sizeof(P));
    P.inst = reinterpret_cast<const pmo::PlateInstance*>(inst);
    P.n_inst = n_inst;
    P.schedule = reinterpret_cast<const uint4*>(schedule);
    P.n_batches = (unsigned long long)n_batches;
    P.seed = run->seed; P.stream = run->stream;
    P.key_dev = run->key_dev;
    P.wth = run->wth; P.roulette_m = run->roulette_m; P.inv_m = 1.f / run->roulette_m;
    P.flags = run->flags;
    P.max_events = kMaxEvents;
    if (run->quorum < 0 || run->quorum > 64) return fail(PMO_ERR_INVALID, "quorum %d outside [0, 64]", run->quorum);
    P.quorum = run->quorum > 0 ? (uint32_t)run->quorum : (uint32_t)kDefaultQuorum;
    if (run->refill < 0 || run->refill > 64) return fail(PMO_ERR_INVALID, "refill %d outside [0, 64]", run->refill);
    P.refill = run->refill > 0 ? (uint32_t)run->refill : (uint32_t)kDefaultRefill;
    P.scalars = tal->scalars;
    P.events = reinterpret_cast<unsigned long long*>(tal->events);
    P.nr = tal->nr; P.dr = tal->dr;
    P.rd_r = tal->nr > 0 ? tal->rd_r : nullptr;
    P.tt_r = tal->nr > 0 ? tal->tt_r : nullptr;
    P.work = reinterpret_cast<unsigned long long*>(workspace);
    // each wave's LDS tally of its current instance, while it stays small beside the walk
    const size_t acc = (size_t)pmo::kBS_COUNT + 2u * (size_t)(tal->nr > 0 ? tal->nr : 0);
    P.wave_acc = acc * sizeof(double) * (size_t)(bs / 64) <= kWaveAccBytes ? (uint32_t)acc : 0u;
    const size_t lds = (size_t)P.wave_acc * sizeof(double) * (size_t)(bs / 64);

    const bool layer = (run->flags & PMO_F_LAYER_STEP) != 0u;
    int nb = 0;
    hipError_t e = layer ? hipOccupancyMaxActiveBlocksPerMultiprocessor(&nb, pmo::plate_batch_kernel<float, true>, bs, lds)
                         : hipOccupancyMaxActiveBlocksPerMultiprocessor(&nb, pmo::plate_batch_kernel<float, false>, bs, lds);
    if (e != hipSuccess) return hip_fail(e, "hipOccupancyMaxActiveBlocksPerMultiprocessor");
    if (nb < 1) nb = 1;
    unsigned long long grid = run->grid_blocks > 0 ? (unsigned long long)run->grid_blocks
                                                   : (unsigned long long)nb * (unsigned long long)d.n_cu;
    const unsigned long long wpb = (unsigned long long)bs / 64;
    const unsigned long long useful = ((unsigned long long)n_batches + wpb - 1) / wpb;
    if (grid > useful) grid = useful;
    hipStream_t s = reinterpret_cast<hipStream_t>(stream);
    hipLaunchKernelGGL(pmo::zero_u64_kernel, dim3(1), dim3(64), 0, s, P.work, 4);
    e = hipGetLastError();
    if (e != hipSuccess) return hip_fail(e, "zero workspace launch");
    if (layer)
        hipLaunchKernelGGL((pmo::plate_batch_kernel<float, true>), dim3((unsigned)grid), dim3(bs), lds, s, P);
    else
        hipLaunchKernelGGL((pmo::plate_batch_kernel<float, false>), dim3((unsigned)grid), dim3(bs), lds, s, P);
    e = hipGetLastError();
    if (e != hipSuccess) return hip_fail(e, "batched walk launch");
    return PMO_OK;
}

}  // extern "C"

struct pmo_turing {
    int n;
    hipfftHandle r2c, c2r;
    size_t work_bytes, spec_bytes;
};

namespace {

unsigned grid_for(unsigned long long n, int n_cu)
{
    unsigned long long b = (n + 255) / 256;
    const unsigned long long cap = (unsigned long long)n_cu * 16ull;
    return (unsigned)(b < 1 ? 1 : (b > cap ? cap : b));
}

int fft_fail(hipfftResult r, const char* what) { return fail(PMO_ERR_HIP, "%s: hipFFT error %d", what, (int)r); }

}  // namespace

extern "C" {

int pmo_turing_create(int32_t n, pmo_turing** out)
{
    g_err[0] = '\0';
    if (!out) return fail(PMO_ERR_INVALID, "null output");
    *out = nullptr;
    if (n < 2 || n > 1024) return fail(PMO_ERR_INVALID, "grid size %d outside [2, 1024]", n);
    pmo_turing* h = new pmo_turing();
    h->n = n;
    size_t w1 = 0, w2 = 0;
    hipfftResult r = hipfftCreate(&h->r2c);
    if (r == HIPFFT_SUCCESS) r = hipfftCreate(&h->c2r);
    if (r == HIPFFT_SUCCESS) r = hipfftSetAutoAllocation(h->r2c, 0);
    if (r == HIPFFT_SUCCESS) r = hipfftSetAutoAllocation(h->c2r, 0);
    if (r == HIPFFT_SUCCESS) r = hipfftMakePlan3d(h->r2c, n, n, n, HIPFFT_R2C, &w1);
    if (r == HIPFFT_SUCCESS) r = hipfftMakePlan3d(h->c2r, n, n, n, HIPFFT_C2R, &w2);
    if (r != HIPFFT_SUCCESS) {
        pmo_turing_destroy(h);
        return fft_fail(r, "plan");
    }
    h->work_bytes = align_up(w1 > w2 ? w1 : w2, 256);
    h->spec_bytes = align_up((size_t)n * n * (n / 2 + 1) * sizeof(float2), 256);
    *out = h;
    return PMO_OK;
}

int64_t pmo_turing_workspace_bytes(const pmo_turing* h)
{
    if (!h) return fail(PMO_ERR_INVALID, "null handle");
    return (int64_t)(2 * h->spec_bytes + h->work_bytes);
}

int pmo_turing_destroy(pmo_turing* h)
{
    if (!h) return PMO_OK;
    if (h->r2c) hipfftDestroy(h->r2c);
    if (h->c2r) hipfftDestroy(h->c2r);
    delete h;
    return PMO_OK;
}

int pmo_turing_step(pmo_turing* h, const pmo_turing_params* prm, float* u, float* v, void* workspace,
                    int32_t steps, void* stream)
{
    g_err[0] = '\0';
    if (!h || !prm || !u || !v || !workspace) return fail(PMO_ERR_INVALID, "null argument");
    if (steps < 0) return fail(PMO_ERR_INVALID, "steps < 0");
    if (!(prm->dx > 0.f) || !(prm->dt > 0.f) || prm->du < 0.f || prm->dv < 0.f)
        return fail(PMO_ERR_INVALID, "need dx, dt > 0 and du, dv >= 0");
    DevInfo d;
    int rc = device_info(&d);
    if (rc != PMO_OK) return rc;
    hipStream_t s = reinterpret_cast<hipStream_t>(stream);
    const int n = h->n;
    const unsigned long long cells = (unsigned long long)n * n * n;
    float2* U = reinterpret_cast<float2*>(workspace);
    float2* V = reinterpret_cast<float2*>(reinterpret_cast<char*>(workspace) + h->spec_bytes);
    void* work = reinterpret_cast<char*>(workspace) + 2 * h->spec_bytes;
    hipfftResult r = hipfftSetStream(h->r2c, s);
    if (r == HIPFFT_SUCCESS) r = hipfftSetStream(h->c2r, s);
    if (r == HIPFFT_SUCCESS) r = hipfftSetWorkArea(h->r2c, work);
    if (r == HIPFFT_SUCCESS) r = hipfftSetWorkArea(h->c2r, work);
    if (r != HIPFFT_SUCCESS) return fft_fail(r, "stream / work area");
    const pmo::TuringCoef c{prm->dt, prm->a1, prm->a2, prm->a3, prm->b1, prm->b2};
    const float cu = prm->dt * prm->du / (prm->dx * prm->dx), cv = prm->dt * prm->dv / (prm->dx * prm->dx);
    const float scale = (float)(1.0 / (double)cells);
    const unsigned gr = grid_for(cells, d.n_cu);
    const unsigned gs = grid_for((unsigned long long)n * n * (n / 2 + 1), d.n_cu);
    for (int it = 0; it < steps; ++it) {
        hipLaunchKernelGGL(pmo::turing_reaction_kernel, dim3(gr), dim3(256), 0, s, u, v, cells, c);
        hipError_t e = hipGetLastError();
        if (e != hipSuccess) return hip_fail(e, "reaction launch");
        r = hipfftExecR2C(h->r2c, u, reinterpret_cast<hipfftComplex*>(U));
        if (r == HIPFFT_SUCCESS) r = hipfftExecR2C(h->r2c, v, reinterpret_cast<hipfftComplex*>(V));
        if (r != HIPFFT_SUCCESS) return fft_fail(r, "forward transform");
        hipLaunchKernelGGL(pmo::turing_spectrum_kernel, dim3(gs), dim3(256), (size_t)n * sizeof(float), s, U, V, n,
                           cu, cv, scale);
        e = hipGetLastError();
        if (e != hipSuccess) return hip_fail(e, "spectrum launch");
        r = hipfftExecC2R(h->c2r, reinterpret_cast<hipfftComplex*>(U), u);
        if (r == HIPFFT_SUCCESS) r = hipfftExecC2R(h->c2r, reinterpret_cast<hipfftComplex*>(V), v);
        if (r != HIPFFT_SUCCESS) return fft_fail(r, "inverse transform");
    }
    return PMO_OK;
}

int pmo_threshold_count(float* u, int64_t n, float thr, uint64_t* n_below, void* stream)
{
    g_err[0] = '\0';
    if (!u || !n_below || n < 0) return fail(PMO_ERR_INVALID, "bad argument");
    if (n == 0) return PMO_OK;
    DevInfo d;
    int rc = device_info(&d);
    if (rc != PMO_OK) return rc;
    hipLaunchKernelGGL(pmo::threshold_kernel, dim3(grid_for((unsigned long long)n, d.n_cu)), dim3(256), 0,
                       reinterpret_cast<hipStream_t>(stream), u, (unsigned long long)n, thr,
                       reinterpret_cast<unsigned long long*>(n_below));
    hipError_t e = hipGetLastError();
    return e == hipSuccess ? PMO_OK : hip_fail(e, "threshold launch");
}

int pmo_count_peaks(const float* u, int64_t rows, int32_t len, int32_t order, uint64_t* count, void* stream)
{
    g_err[0] = '\0';
    if (!u || !count || rows < 0 || len < 1 || order < 1) return fail(PMO_ERR_INVALID, "bad argument");
    if (rows == 0) return PMO_OK;
    DevInfo d;
    int rc = device_info(&d);
    if (rc != PMO_OK) return rc;
    const unsigned long long total = (unsigned long long)rows * (unsigned long long)len;
    hipLaunchKernelGGL(pmo::count_peaks_kernel, dim3(grid_for(total, d.n_cu)), dim3(256), 0,
                       reinterpret_cast<hipStream_t>(stream), u, (unsigned long long)rows, len, order,
                       reinterpret_cast<unsigned long long*>(count));
    hipError_t e = hipGetLastError();
    return e == hipSuccess ? PMO_OK : hip_fail(e, "peak count launch");
}

int pmo_binarize(const float* u, int64_t n, int8_t* labels, void* stream)
{
    g_err[0] = '\0';
    if (!u || !labels || n < 0) return fail(PMO_ERR_INVALID, "bad argument");
    if (n == 0) return PMO_OK;
    DevInfo d;
    int rc = device_info(&d);
    if (rc != PMO_OK) return rc;
    hipLaunchKernelGGL(pmo::binarize_kernel, dim3(grid_for((unsigned long long)n, d.n_cu)), dim3(256), 0,
                       reinterpret_cast<hipStream_t>(stream), u, (unsigned long long)n, labels);
    hipError_t e = hipGetLastError();
    return e == hipSuccess ? PMO_OK : hip_fail(e, "binarize launch");
}

int pmo_fluence_bin(const float* p, int64_t stride, const float* w, int64_t n, int32_t dims, int32_t nr, float dr,
                    int32_t nz, float dz, double* out, void* stream)
{
    g_err[0] = '\0';
    if (!out) return fail(PMO_ERR_INVALID, "null output");
    if (n < 0 || stride < n) return fail(PMO_ERR_INVALID, "need 0 <= n <= stride");
    if (n > 0 && (!p || !w)) return fail(PMO_ERR_INVALID, "null point or weight array");
    if (dims != 2 && dims != 3) return fail(PMO_ERR_INVALID, "dims must be 2 or 3");
    if (nr < 1 || nz < 1 || !(dr > 0.f) || !(dz > 0.f)) return fail(PMO_ERR_INVALID, "bad bins");
    if (n == 0) return PMO_OK;
    DevInfo d;
    int rc = device_info(&d);
    if (rc != PMO_OK) return rc;
    // bins that fit a block's LDS are summed there first (one 1024-lane block per CU, so
    // each bin takes at most one global add per CU); larger tables take one add per point
    const unsigned long long nbins = dims == 2 ? (unsigned long long)nr * nz : 4ull * nr * nr * nz;
    const bool in_lds = nbins * sizeof(double) <= kFluenceLdsBytes;
    const unsigned bs = in_lds ? 1024u : 256u;
    unsigned long long blocks = ((unsigned long long)n + bs - 1) / bs;
    const unsigned long long cap = (unsigned long long)d.n_cu * (in_lds ? 1ull : 8ull);
    if (blocks > cap) blocks = cap;
    const uint32_t lds_bins = in_lds ? (uint32_t)nbins : 0u;
    hipLaunchKernelGGL(pmo::fluence_bin_kernel, dim3((unsigned)blocks), dim3(bs), (size_t)lds_bins * sizeof(double),
                       reinterpret_cast<hipStream_t>(stream), p, (unsigned long long)stride, w, (unsigned long long)n,
                       dims, nr, dr, nz, dz, out, lds_bins);
    hipError_t e = hipGetLastError();
    return e == hipSuccess ? PMO_OK : hip_fail(e, "fluence binning launch");
}

int pmo_obd_trace(const pmo_obd* cfg, const float* p, const float* v, const float* w, int64_t n, int64_t stride,
                  const float* steps, int32_t n_steps, double* intensity, void* stream)
{
    g_err[0] = '\0';
    static_assert(sizeof(pmo::ObdStep) == PMO_OBD_STEP_FIELDS * sizeof(float), "OBD step layout");
    if (!cfg || !steps || !intensity) return fail(PMO_ERR_INVALID, "null argument");
    if (n < 0 || stride < n) return fail(PMO_ERR_INVALID, "need 0 <= n <= stride");
    if (n > 0 && (!p || !v || !w)) return fail(PMO_ERR_INVALID, "null exit-state array");
    if (n_steps < 1 || n_steps > 1024) return fail(PMO_ERR_INVALID, "n_steps %d outside [1, 1024]", n_steps);
    if (!(cfg->n_1 > 1.f) || !(cfg->n_2 > 1.f)) return fail(PMO_ERR_INVALID, "lens indices must exceed 1");
    if (!(cfg->r_1 > 0.f) || !(cfg->r_2 > 0.f) || !(cfg->outer_r1 > 0.f) || !(cfg->outer_r2 > 0.f) ||
        cfg->outer_r1 >= cfg->r_1 || cfg->outer_r2 >= cfg->r_2)
        return fail(PMO_ERR_INVALID, "lens radii: need 0 < aperture < surface radius");
    if (!(cfg->slit_r_out > cfg->slit_r_in) || !(cfg->pd_r > 0.f) || cfg->slit_thickness < 0.f)
        return fail(PMO_ERR_INVALID, "bad slit or photodiode size");
    if (n == 0) return PMO_OK;
    DevInfo d;
    int rc = device_info(&d);
    if (rc != PMO_OK) return rc;

    pmo::ObdParams P;
    memset(&P, 0, sizeof(P));
    P.p = p; P.v = v; P.w = w;
    P.n = (unsigned long long)n; P.stride = (unsigned long long)stride;
    P.steps = reinterpret_cast<const pmo::ObdStep*>(steps);
    P.n_steps = n_steps;
    P.r1 = cfg->r_1; P.n1 = cfg->n_1; P.outer1_sq = cfg->outer_r1 * cfg->outer_r1;
    P.r2 = cfg->r_2; P.n2 = cfg->n_2; P.outer2_sq = cfg->outer_r2 * cfg->outer_r2;
    P.d_in = cfg->slit_r_in; P.d_out = cfg->slit_r_out; P.slit_th = cfg->slit_thickness;
    P.pd_r_sq = cfg->pd_r * cfg->pd_r;
    P.ld_fix = cfg->ld_fix_part;
    // LD mount half-planes (LdFixingPart, _classes.py:337-350)
    const double ang = cfg->lfp_angle, wd = cfg->lfp_width;
    P.lfp_a1 = (float)(1.0 / tan(ang) - 1.0 / (sin(ang) * cos(ang)));
    P.lfp_b1 = (float)(wd / cos(ang));
    P.lfp_a2 = (float)((cos(ang) + 1.0) / sin(ang));
    P.lfp_b2 = (float)(-wd);
    P.lfp_b3 = (float)(-wd);
    P.lfp_th = cfg->lfp_thickness;
    P.inversion = cfg->inversion;
    P.side = cfg->side;
    P.expect = cfg->expect;
    P.seed = cfg->seed; P.stream = cfg->stream;
    P.out = intensity;

    const int bs = 256;
    // step table | FP64 I(Z) bins | one queue of uint16 (lane, step) pairs per wave
    const size_t lds = (size_t)n_steps * (sizeof(pmo::ObdStep) + sizeof(double)) +
                       (size_t)(bs / 64) * pmo::kObdQueue * sizeof(uint16_t);
    unsigned long long blocks = ((unsigned long long)n + bs - 1) / bs;
    const unsigned long long cap = (unsigned long long)d.n_cu * 8ull;
    if (blocks > cap) blocks = cap;
    hipLaunchKernelGGL(pmo::obd_kernel, dim3((unsigned)blocks), dim3(bs), lds, reinterpret_cast<hipStream_t>(stream),
                       P);
    hipError_t e = hipGetLastError();
    return e == hipSuccess ? PMO_OK : hip_fail(e, "OBD launch");
}

}  // extern "C"
