/*
 * pmo.h -- C ABI of libpmo.so, the MI355X (gfx950) voxel Monte Carlo photon walk.
 *
 * The library replaces the walk kernels of pyMonteOpt's voxel models:
 *   - the Numba CPU walk  vmc_kernel(add, p, v, w, ma, ms, n, g, voxel_model, l, nPh, end_point)
 *     (pymopt/voxel/_kernel.py:23-172), and
 *   - the CuPy RawKernel  cuVMC(add, p, v, w, ma, ms, n, g, voxel_model, l, M, L, nPh,
 *     end_point, rand_seed)  (pymopt/voxel_gpu/_cukernel.py:27-227), launched from
 *     BaseVoxelMonteCarlo.start (pymopt/voxel_gpu/_classes.py:75-129).
 *
 * Differences in shape from those interfaces, all deliberate:
 *   - packets are generated on the device from the source description and their id
 *     (no 28 B/packet host state, no H2D/D2H copies of add/p/v/w);
 *   - the walk accumulates FP64 tallies on the device (Rd, Tt, absorbed weight per voxel,
 *     radial/angular/line/annular exit tallies, an exact energy ledger, event counts);
 *   - the per-packet exit state (the reference's add/p/v/w after the walk) is available
 *     as an optional dump for small runs;
 *   - every call is asynchronous on the caller's HIP stream and returns an error code.
 *
 * Conventions: all pointers marked "device" are HIP device pointers owned by the caller;
 * the library never frees them and keeps no device state between calls except its code
 * objects. Tally buffers are accumulated into (caller zeroes them). The walk entry points
 * allocate nothing: per-call scratch lives in the caller-provided workspace of
 * pmo_workspace_bytes() bytes, so they may be captured into a HIP graph. The one exception
 * is pmo_turing_create, whose hipFFT plans hold hipFFT's own tables until
 * pmo_turing_destroy. tests/c/walk_slab.c is a complete C caller.
 * A `stream` argument is a hipStream_t (NULL = the default stream).
 * Return value: PMO_OK or a negative PMO_ERR_* code; pmo_last_error() gives the message
 * (thread-local).
 */
#ifndef PMO_H
#define PMO_H

#include <stddef.h>
#include <stdint.h>

#ifdef __cplusplus
extern "C" {
#endif

#define PMO_ABI_VERSION 11

/* ---- error codes ---------------------------------------------------------------- */
#define PMO_OK 0
#define PMO_ERR_INVALID (-1)     /* bad argument: null pointer, size or range        */
#define PMO_ERR_LABEL (-2)       /* label volume / optics table inconsistency         */
#define PMO_ERR_HIP (-3)         /* HIP runtime error (launch, device query)          */
#define PMO_ERR_UNSUPPORTED (-4) /* request outside what this library implements      */

/* ---- beam kinds ------------------------------------------------------------------ */
#define PMO_BEAM_PENCIL 0 /* all packets start at the centre of the launch voxel's top face  */
#define PMO_BEAM_FLAT 1   /* (x, y) uniform in a disc of radius `radius` (mm)               */
#define PMO_BEAM_GAUSS 2  /* TEM00: x, y ~ N(0, radius/2), radius = beam waist w (mm)       */
#define PMO_BEAM_STATES 3 /* per-packet launch states from the caller (pmo_source.add/p/v/w),
                             the arrays the reference kernels take (pymopt/voxel/_kernel.py:23-28,
                             pymopt/voxel_gpu/_cukernel.py:27-31); pmo_walk only             */

/* ---- run flags ------------------------------------------------------------------- */
#define PMO_F_FP64 (1u << 0)           /* walk in double precision (checking build)        */
#define PMO_F_NORMAL_FRESNEL (1u << 1) /* at exactly normal incidence use the Fresnel limit
                                          ((ni-nt)/(ni+nt))^2; default keeps R = 0 there,
                                          as pymopt/voxel/_kernel.py:89-93 does            */
#define PMO_F_LAYER_STEP (1u << 2)     /* pmo_walk_plates only: step from layer boundary to
                                          layer boundary and lateral wall in continuous
                                          coordinates instead of from voxel face to voxel
                                          face. In a plate the skipped faces change neither
                                          label nor index, so the walk is the same in exact
                                          arithmetic and uses the same draws; results equal
                                          the voxel walk's up to FP32 rounding. Needs every
                                          instance's layer table.                           */

/* ---- scalar tally slots (double) ---------------------------------------------------- */
enum {
    PMO_S_RD = 0,        /* weight escaping with v_z < 0 (reference getRdTtRate Rd)           */
    PMO_S_TT = 1,        /* weight escaping with v_z > 0 (reference Tt)                       */
    PMO_S_OTHER = 2,     /* weight escaping with v_z == 0                                     */
    PMO_S_ABS = 3,       /* absorbed weight (sum of all deposits)                             */
    PMO_S_KILL = 4,      /* weight removed by Russian roulette                                */
    PMO_S_GAIN = 5,      /* weight added by roulette survival (m-1)*w                         */
    PMO_S_W0 = 6,        /* launched weight                                                   */
    PMO_S_FACE = 7,      /* 6 slots: escaped weight through x-, x+, y-, y+, z- (top), z+      */
    PMO_S_COUNT = 13
};

/* ---- event counters (uint64) -------------------------------------------------------- */
enum {
    PMO_EV_PACKETS = 0,  /* packets launched                                                  */
    PMO_EV_INTERACT = 1, /* hops that ended inside the medium (drop + roulette test + spin)   */
    PMO_EV_CROSS = 2,    /* voxel-face crossings (including reflections)                      */
    PMO_EV_FRESNEL = 3,  /* crossings with a refractive-index change (Fresnel draw)           */
    PMO_EV_ROULETTE = 4, /* roulette tests (w <= wth)                                         */
    PMO_EV_KILLED = 5,   /* packets ended by roulette                                         */
    PMO_EV_ESCAPED = 6,  /* packets that entered an end (ambient) voxel                       */
    PMO_EV_GLOBAL_DEPOSITS = 7, /* global atomic requests for the absorbed-weight map: one per
                                   voxel deposit outside the LDS window, or with absorbed_lines
                                   one per 64 B cube line handed over                        */
    PMO_EV_COUNT = 8
};

/* ---- geometry store ----------------------------------------------------------------- */
typedef struct pmo_volume {
    const uint8_t* labels; /* device, C order [nx][ny][nz] (z fastest); the 1-voxel outer shell
                              must hold the end label n_labels                                */
    int32_t nx, ny, nz;    /* padded extents (reference voxel_model.shape)                     */
    int32_t n_labels;      /* tissue labels 0..n_labels-1; label n_labels is the end label     */
    float pitch;           /* voxel edge (mm), the reference's voxel_space / l                 */
    const float* optics;   /* device, float4 per label, n_labels+1 entries: (mu_a, mu_s, g, n);
                              entry n_labels holds n_ambient (mu_a, mu_s, g unused)            */
    const uint8_t* skip;   /* device [nx][ny][nz], pmo_skip_radius of these labels, or NULL.
                              pmo_walk_forward_diff (without mua_field) crosses the faces
                              inside same-label regions in one step with it; pmo_walk reads the
                              packed form below instead and rejects skip without it.         */
    const uint16_t* label_radius; /* device [nx][ny][nz], labels[v] | radius[v] << 8
                              (pmo_pack_label_radius of the labels and their pmo_skip_radius),
                              or NULL. With it pmo_walk crosses the faces inside same-label
                              regions in one step (the same draws; merged arithmetic, rounded
                              once), reading a voxel's label and radius with one 2-byte load;
                              NULL steps voxel by voxel as the reference does. It must hold
                              exactly `labels`. No reference counterpart.                      */
} pmo_volume;

typedef struct pmo_source {
    int32_t kind;          /* PMO_BEAM_*                                                       */
    int32_t ix, iy, iz;    /* launch voxel = beam centre (reference _set_inital_add)           */
    float radius;          /* disc radius / Gaussian waist (mm), unused for the pencil beam    */
    float dir[3];          /* launch direction, unit vector (reference _set_inital_vector)     */
    float w0;              /* launch weight, 1 - Rsp (reference _set_inital_w)                 */
    /* PMO_BEAM_STATES only (else NULL/0): device arrays indexed by packet id - first_packet,
       component c of packet i at [c * states_stride + i]. add is the voxel address, p the
       position in the voxel's frame ([-pitch/2, pitch/2] per axis; outside values are clamped),
       v the direction, w the weight. An address outside the volume leaves at once. ix, iy, iz
       still give the centre of the LDS tally window (and must be an interior voxel).          */
    const int32_t* add;    /* [3][states_stride]                                                */
    const float* p;        /* [3][states_stride]                                                */
    const float* v;        /* [3][states_stride]                                                */
    const float* w;        /* [states_stride]                                                   */
    uint64_t states_stride; /* >= n_packets                                                     */
} pmo_source;

typedef struct pmo_run {
    uint64_t n_packets;     /* packets in this call                                            */
    uint64_t first_packet;  /* packet ids are [first_packet, first_packet + n_packets)         */
    uint32_t seed;          /* Philox key word 0 (the reference's rand_seed)                   */
    uint32_t stream;        /* Philox key word 1 (independent streams for one seed)            */
    float wth;              /* roulette threshold (reference 1e-4)                             */
    float roulette_m;       /* roulette survival 1/m, weight multiplier m (reference 10)       */
    uint32_t flags;         /* PMO_F_*                                                         */
    int32_t block_size;     /* threads per block, multiple of 64, <= 1024; 0 = default (1024 for
                               runs of >= 2^20 packets with an absorbed map over more than
                               16*16*8 voxels, else 256)                                       */
    int32_t grid_blocks;    /* persistent blocks; 0 = fill the device at full occupancy        */
    int32_t packets_per_lane; /* a wave claims 64*this packet ids per work-counter atomic; 0 = default */
    int32_t window[3];      /* extents (x, y, z) of the box of the absorbed tally that each block
                               accumulates in LDS, centred laterally on the launch voxel and
                               starting at its z; {-1,-1,-1} = none; {0,0,0} = default: the
                               whole volume when it has at most 16*16*8 voxels, else 28x28x24
                               with 1024-lane blocks (the default block size for runs of >= 2^20
                               packets with an absorbed map), else 16x16x8; shrunk to the LDS the
                               histograms leave                                                 */
    const uint32_t* key_dev; /* device (seed, stream) read by the kernel instead of seed/stream when
                                non-NULL, so a captured HIP graph can be replayed with new keys */
    int32_t quorum;          /* a wave runs its interaction step once this many of its 64 lanes are
                                ready for one (else it advances face crossings); 0 = default
                                (32; 40 with absorbed_lines)                                    */
    int32_t refill;          /* pmo_walk, pmo_walk_plates: a wave starts new packets once this many
                                of its lanes are idle (or all are), so the launch code runs for
                                several lanes at a time; 1 = every idle lane at once; 0 = default
                                (4). Packets and their histories do not change.                 */
} pmo_run;

typedef struct pmo_tallies {
    double* scalars;        /* device [PMO_S_COUNT], required                                  */
    uint64_t* events;       /* device [PMO_EV_COUNT] or NULL                                   */
    double* absorbed;       /* device [nx*ny*nz] deposited weight per voxel (padded index) or NULL */
    int32_t nr;             /* radial bins (i*dr, (i+1)*dr], r = hypot(x, y) of the exit point */
    float dr;
    double* rd_r;           /* device [nr], v_z < 0 exits (reference radialDistance on Rd_index) or NULL */
    double* tt_r;           /* device [nr], v_z > 0 exits or NULL                               */
    double* rd_top_r;       /* device [nr], top-face exits (virtualOBD.py:165 detector) or NULL */
    int32_t na;             /* angular bins of width pi/(2 na) over acos|v_z|                  */
    double* rd_a;           /* device [na] or NULL (reference angularyResolved on Rd)          */
    double* tt_a;           /* device [na] or NULL                                             */
    int32_t nl;             /* side line bins (i*dl, (i+1)*dl] over exit z                     */
    float dl;
    float line_y;           /* |y| < line_y selection (reference lineDistance y_range)         */
    double* line_z;         /* device [nl], v_x < 0 exits (virtualOBD.py:171-172) or NULL      */
    int32_t n_rings;        /* annular top-face detectors                                      */
    const float* rings;     /* device [2*n_rings]: (r_in, r_out], mm                           */
    double* ring_w;         /* device [n_rings] or NULL                                        */
    float* exit_state;      /* device [7][n_packets] (x, y, z, vx, vy, vz, w) in the reference's
                               get_result encoding (z rounded to 6 decimals as np.round(.., 6)
                               there), indexed by packet id - first_packet, or NULL; z also
                               selects the top-face (z < 1e-8) and side-line detectors        */
    int32_t rz_nr, rz_nz;   /* internal fluence bins r in [i*rz_dr, (i+1)*rz_dr), z in
                               [k*rz_dz, (k+1)*rz_dz), r = hypot(x, y) and z of the interaction point */
    float rz_dr, rz_dz;
    double* fluence_rz;     /* device [rz_nr][rz_nz]: sum of deposit / mu_a at each interaction,
                               the r-z fluence pymopt/fluence/intarnal_fluence.py Fluence2D intends
                               (pymopt/voxel_unverified/_classes.py:394-399), or NULL. Held in LDS
                               per block: rz_nr*rz_nz*8 B counts against the block's LDS (the
                               device's per-block limit, 160 KiB on gfx950)                     */
    int32_t* exit_add;      /* device [3][n_packets] or NULL: the exit state the reference kernel
                               returns (pymopt/voxel/_kernel.py:110-116,172), the end-label voxel
                               each packet entered; 0 for packets ended by roulette (:134-137)   */
    float* exit_p;          /* device [3][n_packets] or NULL: the position in that voxel's frame
                               ([-pitch/2, pitch/2] per axis), 0 after roulette; l*(add - centre) + p
                               is the exit point (_encooder, voxel/_classes.py:261-269)           */
    double* absorbed_lines; /* device scratch of PMO_ABSORBED_LINES_DOUBLES(nx, ny, nz) doubles, or
                               NULL. With it pmo_walk gathers the absorbed weight outside its LDS
                               window per 2x2x2-voxel cube (one 64 B line of FP64) in per-lane LDS
                               buffers and adds each cube's line with one atomic request, then adds
                               the scratch into `absorbed`; the library clears it. NULL: one FP64
                               atomic per voxel deposit. Either way `absorbed` gets the same sums
                               (up to FP64 summation order).                                    */
} pmo_tallies;

/* doubles of pmo_tallies.absorbed_lines for an nx x ny x nz volume: 8 per 2x2x2 cube */
#define PMO_ABSORBED_LINES_DOUBLES(nx, ny, nz) \
    (8ull * (unsigned long long)(((nx) + 1) / 2) * (unsigned long long)(((ny) + 1) / 2) * \
     (unsigned long long)(((nz) + 1) / 2))

/* ---- batched plates ----------------------------------------------------------------- */
/* One instance of a batched sweep: a plate (pymopt/voxel/_classes.py:352-424 PlateModel
 * geometry) of nx x ny x nz padded voxels whose interior label depends on z only. */
typedef struct pmo_plate_instance {
    const uint8_t* z_labels; /* device [nz]: label of each z slice; slices 0 and nz-1 hold the end
                                label, as does the lateral 1-voxel shell                          */
    const float* optics;     /* device float4 [n_labels+1] (mu_a, mu_s, g, n), as pmo_volume      */
    int32_t nx, ny, nz;
    int32_t n_labels;
    float pitch;
    int32_t src_kind, ix, iy, iz; /* launch (pmo_source fields)                                 */
    float radius, dir[3], w0;
    uint64_t n_packets;      /* this instance's budget (< 2^32)                                    */
    const float* layers;     /* device float4 [n_layers]: (z_top, z_bottom, label, 0) of each run of
                                equal labels in z_labels[1 .. nz-2], top to bottom, z in mm below the
                                top face (slice i spans pitch*(i-1) .. pitch*i); read with
                                PMO_F_LAYER_STEP, else may be NULL                                  */
    int32_t n_layers;
    int32_t reserved;        /* 0                                                                  */
} pmo_plate_instance;

typedef struct pmo_batch_tallies {
    double* scalars;         /* device [n_inst][PMO_S_COUNT], required                            */
    uint64_t* events;        /* device [PMO_EV_COUNT] summed over instances, or NULL             */
    int32_t nr;              /* radial bins per instance (i*dr, (i+1)*dr]                         */
    float dr;
    double* rd_r;            /* device [n_inst][nr] v_z < 0 exits, or NULL                       */
    double* tt_r;            /* device [n_inst][nr] v_z > 0 exits, or NULL                       */
} pmo_batch_tallies;

/* ---- differentiable walk ------------------------------------------------------------ */
typedef struct pmo_diff {
    const float* mua_field;  /* device [nx*ny*nz] per-voxel mu_a replacing the label's, or NULL   */
    int32_t nr;              /* radial bins of the detected signal Rd(r), (i*dr, (i+1)*dr]          */
    float dr;
    double* rd_r;            /* forward: device [nr] weight of v_z < 0 exits per bin               */
    double* jac;             /* forward: device [nr][n_labels][2] sums of W L_j and W n_j of those
                                exits (L_j path length, n_j interactions in label j), or NULL;
                                needs n_labels <= 4                                                 */
    int32_t* exit_rec;       /* device [2*n_packets] per packet (bin or -1, weight as float bits):
                                written by the forward pass, read by the adjoint                    */
    const double* coef;      /* adjoint: device [nr], d loss / d Rd_r[b] divided by (area_b * N)    */
    double* grad_mua;        /* adjoint: device [nx*ny*nz] accumulates d loss / d mu_a per voxel    */
    uint64_t* mismatch;      /* adjoint: device counter of replayed exits that differ from the
                                forward record (0 when replay is exact), or NULL                    */
    double* grad_lines;      /* adjoint: device scratch of PMO_ABSORBED_LINES_DOUBLES(nx, ny, nz)
                                doubles, or NULL. With it the replay gathers its deposits outside
                                the LDS box per 2x2x2-voxel cube in per-lane LDS buffers (one
                                atomic request per 64 B cube line handed over) and then adds the
                                scratch into grad_mua; the library clears it. NULL: one FP64
                                atomic per voxel deposit. Same sums either way (up to FP64
                                summation order).                                                   */
} pmo_diff;

/* ---- library ------------------------------------------------------------------------ */
int pmo_abi_version(void);
const char* pmo_last_error(void);
size_t pmo_workspace_bytes(void);
/* number of compute units and the persistent grid the walk would use for a block size */
int pmo_device_info(int32_t* n_cu, int32_t* blocks_per_cu_walk, int32_t block_size);

/* One walk of run->n_packets packets through `vol`.
 * Replaces cuVMC (pymopt/voxel_gpu/_cukernel.py:27-227) plus the host tallies built on its
 * output: getRdTtRate (pymopt/voxel_gpu/_classes.py:326-334), radialDistance /
 * angularyResolved / lineDistance (pymopt/metrics/_scoring.py:16-51) and the virtualOBD
 * detectors (pymopt/virtualOBD.py:158-172). */
int pmo_walk(const pmo_volume* vol, const pmo_source* src, const pmo_run* run,
             const pmo_tallies* tal, void* workspace, void* stream);

/* pmo_walk without a workspace argument, for callers that cannot allocate device memory
 * (code under HIP graph capture, tight serving loops, harnesses that forbid hipMalloc). No
 * counterpart in the reference. Its scratch, the walk's 256 B packet-claim counter block, is
 * a static device array of libpmo.so (one per GPU, held until the process ends); that need is
 * the same at every volume and packet count, so no size is refused for it. The call makes no
 * hipMalloc/hipFree; the counters are cleared by a kernel on `stream`. Every call on one GPU
 * shares that GPU's array: calls that overlap on one GPU (two streams, two threads) race and
 * must be serialised by the caller. Arguments, results and errors are pmo_walk's. */
int pmo_walk_noalloc(const pmo_volume* vol, const pmo_source* src, const pmo_run* run,
                     const pmo_tallies* tal, void* stream);

/* out[v] = labels[v] | radius[v] << 8 for v < n (device buffers; out holds n uint16): the
 * packed form of a label volume and its pmo_skip_radius that pmo_volume.label_radius takes.
 * No reference counterpart. */
int pmo_pack_label_radius(const uint8_t* labels, const uint8_t* radius, uint16_t* out, int64_t n, void* stream);

/* Largest radius pmo_skip_radius accepts */
#define PMO_SKIP_MAX_CAP 127

/* radius[v] = min(cap, Chebyshev distance from voxel v to the nearest voxel that lies on the
 * volume's outer face, has a 26-neighbour of another label or holds label 255): the cube of
 * half-width radius[v] voxels around v holds v's label only. labels is the walk's volume; radius and
 * scratch are distinct device buffers of nx*ny*nz bytes owned by the caller (scratch is
 * overwritten).
 * 1 <= cap <= PMO_SKIP_MAX_CAP. Feeds pmo_volume.skip; no reference counterpart: the
 * reference steps every voxel face (pymopt/voxel/_kernel.py:62-125). */
int pmo_skip_radius(const uint8_t* labels, int32_t nx, int32_t ny, int32_t nz, int32_t cap, uint8_t* radius,
                    uint8_t* scratch, void* stream);

/* Diagnostic: out[4*i .. 4*i+3] = Philox4x32-10(ctr[4*i .. 4*i+3], key) on the device, the
 * generator every walk draw comes from (no reference counterpart: the reference draws from
 * curand XORWOW, pymopt/voxel_gpu/_cukernel.py:42-43). ctr and out are device arrays of
 * 4*n uint32. */
int pmo_philox(const uint32_t* ctr, uint32_t key0, uint32_t key1, uint32_t* out, int64_t n,
               void* stream);

/* Forward pass of the differentiable walk: the walk of pmo_walk restricted to the scalar
 * tallies (into `scalars` [PMO_S_COUNT], `events` [PMO_EV_COUNT] or NULL) and the detected
 * signal Rd(r), plus the perturbation-MC sums of pmo_diff.jac and the per-packet exit
 * record. With vol->skip and no mua_field it crosses the faces inside same-label cubes in one
 * step, as pmo_walk does (the same draws; the per-label sums change by rounding only); its
 * exit record is then not the voxel-stepping replay's and must not be passed to
 * pmo_walk_adjoint (which would count mismatches). No reference counterpart: the reference
 * estimates optical properties by a parameter sweep
 * (.ipynb_checkpoints/si_in_光学特性推定-checkpoint.ipynb:13945-13967). */
int pmo_walk_forward_diff(const pmo_volume* vol, const pmo_source* src, const pmo_run* run, double* scalars,
                          uint64_t* events, const pmo_diff* diff, void* workspace, void* stream);

/* Adjoint by path replay: regenerates from its Philox counter the path of every packet whose
 * recorded exit bin has coef != 0 and accumulates -coef[bin] * W * (length in voxel v) into
 * grad_mua[v]. Needs the same volume, source, run (seed, stream, packets) and mua_field as
 * the forward pass that wrote exit_rec. */
int pmo_walk_adjoint(const pmo_volume* vol, const pmo_source* src, const pmo_run* run, const pmo_diff* diff,
                     void* workspace, void* stream);

/* Diagnostics that re-measure the two peaks the walk is measured against (no reference
 * counterpart). The caller times them on `stream` (e.g. with HIP events):
 * pmo_probe_atomics issues blocks*256*iters device-scope atomic adds (no return) of 1 to
 * pseudo-random addresses of the device array buf[n], whose element type `kind` names; with
 * PMO_PROBE_F64_PER_XCD, buf holds 8 copies of n doubles and each workgroup adds into the
 * copy of the XCD it runs on (PMO_PROBE_F64_PER_XCD_WG: the same at workgroup scope);
 * with PMO_PROBE_F64_LINES each wave adds to 64 consecutive doubles (eight 64 B lines) of a
 * pseudo-random 512 B block of buf[n], n >= 64; with PMO_PROBE_F64_BLOCKS_32B / _64B / _128B
 * the lanes in groups of 4 / 8 / 16 each add to the consecutive doubles of their own
 * pseudo-random 32 / 64 / 128 B-aligned block (n >= 16). pmo_probe_valu runs blocks*256 lanes that
 * each issue iters x 32 independent FP32 FMAs and write one float to out[blocks*256].
 * pmo_probe_philox runs blocks*256 lanes that each compute iters Philox4x32-10 blocks (the
 * walk's draws, one block per interaction) and write one word to out[blocks*256]. */
enum {
    PMO_PROBE_F64 = 0,
    PMO_PROBE_F32 = 1,
    PMO_PROBE_U64 = 2,
    PMO_PROBE_U32 = 3,
    PMO_PROBE_F64_PER_XCD = 4,
    PMO_PROBE_F64_PER_XCD_WG = 5,
    PMO_PROBE_F64_LINES = 6,
    PMO_PROBE_F64_BLOCKS_32B = 7,
    PMO_PROBE_F64_BLOCKS_64B = 8,
    PMO_PROBE_F64_BLOCKS_128B = 9
};
int pmo_probe_atomics(void* buf, int64_t n, int32_t blocks, int32_t iters, int32_t kind, void* stream);
int pmo_probe_valu(float* out, int32_t blocks, int32_t iters, void* stream);
int pmo_probe_philox(uint32_t* out, int32_t blocks, int32_t iters, void* stream);

/* Host helper: the round-robin batch schedule of a sweep. Writes up to `cap` batches
 * (instance, first local packet id, count, 0) into `sched` (host, 4*cap uint32), cycling over
 * the instances with at most `batch` packets per entry. Returns the number of batches the
 * sweep needs (call with cap = 0 to size the buffer) or a negative error code. */
int64_t pmo_plate_schedule(const uint64_t* n_packets, int32_t n_inst, int32_t batch, uint32_t* sched,
                           int64_t cap);

/* ---- OBD receiver ------------------------------------------------------------------- */
/* The receiver of pymopt/optics/_classes.py:381-567 (class OBD; SideOBD through x_shift):
 * lens 1 (plano face toward the sample), slit 1, the LD mount, slit 2, lens 2 (convex face
 * toward the sample) and the photodiode, scanned over focus offsets Z. Lengths in mm. */
#define PMO_OBD_STEP_FIELDS 12  /* floats per row of `steps`: z_lens1, c_lens1 (lens-1 sphere
                                   centre), z_slit1, z_lfp (LD mount front), z_slit2, z_lens2,
                                   c_lens2, z_pd, x_shift (SideOBD lateral shift), 3 unused;
                                   computed as OBD.opticalUnit does (:454-460)              */
typedef struct pmo_obd {
    float r_1, n_1, outer_r1;      /* lens 1: surface radius r_1, index, aperture outerD_1/2    */
    float r_2, n_2, outer_r2;      /* lens 2                                                    */
    float slit_r_in, slit_r_out;   /* annulus slit_D/2 -/+ slit_width/2 (Slit.d_in, d_out)      */
    float slit_thickness;
    float pd_r;                    /* photodiode radius d_pd/2                                  */
    int32_t ld_fix_part;           /* apply the LD mount cut                                    */
    float lfp_angle;               /* LdFixingPart.angle as the reference writes it: 68/180     */
    float lfp_thickness, lfp_width;/* 5 and 6 mm in the reference                               */
    int32_t inversion;             /* trace v_z > 0 exits, mirrored in y (params 'inversion')   */
    float side;                    /* side view of a block this wide, 0 = off (params 'side')   */
    int32_t expect;                /* 1: weight by the Fresnel transmission instead of drawing  */
    uint32_t seed, stream;         /* Philox key of the Fresnel draws                           */
} pmo_obd;

/* Adds to intensity[s] the weight of the exit packets that reach the photodiode at focus
 * step s (OBD.opticalAnalysisMeth, :491-525). p and v are device [3][stride] rows x, y, z
 * (the layout of pmo_tallies.exit_state rows 0-5), w is device [n]; the packets taken are
 * those with v_z < 0 (v_z > 0 with inversion). steps is device [n_steps][PMO_OBD_STEP_FIELDS],
 * 1 <= n_steps <= 1024, and intensity device [n_steps] (FP64, zeroed by the caller). */
int pmo_obd_trace(const pmo_obd* cfg, const float* p, const float* v, const float* w, int64_t n, int64_t stride,
                  const float* steps, int32_t n_steps, double* intensity, void* stream);

/* ---- fluence binning ---------------------------------------------------------------- */
/* Fluence2D / Fluence3D.saveFluesnce of pymopt/fluence/intarnal_fluence.py (:12-152) with
 * each weight kept with its own point (the reference sorts the points but not the weights,
 * :35-37 and :113-115). Adds w[j] to the bin of point (p[0][j], p[1][j], p[2][j]); p is
 * device [3][stride], w device [n]. dims = 2: out [nr][nz], r = hypot(x, y) in
 * [i*dr, (i+1)*dr), z in [k*dz, (k+1)*dz). dims = 3: out [2nr][2nr][nz], x and y in
 * [(i-nr)*dr, (i-nr+1)*dr). Points outside every bin are skipped. out is FP64 (accumulated). */
int pmo_fluence_bin(const float* p, int64_t stride, const float* w, int64_t n, int32_t dims, int32_t nr, float dr,
                    int32_t nz, float dz, double* out, void* stream);

/* ---- Turing-pattern bone model ------------------------------------------------------ */
/* TuringPattern of pymopt/modeling_gpu/_classes.py:25-176: a reaction-diffusion system on
 * a periodic n^3 grid whose thresholded activator is the trabecular-bone label volume. */
typedef struct pmo_turing pmo_turing;   /* the FFT plans of one grid size (opaque)             */
typedef struct pmo_turing_params {
    float du, dv;          /* diffusion coefficients of the activator u and inhibitor v          */
    float dx, dt;          /* grid spacing and time step                                         */
    float a1, a2, a3;      /* f(u, v) = u + dt (a1 u + a2 v + a3 u^3); 0.6, -1, -1 (:51)          */
    float b1, b2;          /* g(u, v) = v + dt (b1 u + b2 v); 1.5, -2 (:52)                       */
} pmo_turing_params;

/* Plans the R2C and C2R transforms of an n^3 grid (2 <= n <= 1024). hipFFT keeps its own
 * small tables; the transforms' work area and the two spectra come from the caller's
 * workspace of pmo_turing_workspace_bytes(). */
int pmo_turing_create(int32_t n, pmo_turing** out);
int64_t pmo_turing_workspace_bytes(const pmo_turing* h);
int pmo_turing_destroy(pmo_turing* h);
/* `steps` steps of _calc_onestep (:112-115) on device fields u, v [n][n][n] in place:
 * u <- Re F^-1[Re F[f(u, v)] / lambda_u], v <- Re F^-1[Re F[g(u, v)] / lambda_v], lambda
 * the symbol of the implicit 7-point diffusion stencil (:83-104). */
int pmo_turing_step(pmo_turing* h, const pmo_turing_params* prm, float* u, float* v, void* workspace,
                    int32_t steps, void* stream);
/* u <= thr -> 0, adding the number of such cells to *n_below (device) (:134-138). */
int pmo_threshold_count(float* u, int64_t n, float thr, uint64_t* n_below, void* stream);
/* Adds to *count (device) the strict local maxima of u along rows of length len against
 * `order` neighbours each side, edges clipped (scipy argrelmax; :140-160). */
int pmo_count_peaks(const float* u, int64_t rows, int32_t len, int32_t order, uint64_t* count, void* stream);
/* labels = u > 0 (:162-165). */
int pmo_binarize(const float* u, int64_t n, int8_t* labels, void* stream);

/* Many plates in one persistent launch. Packet j of instance i draws from Philox key
 * (seed, stream + i) with packet id j, i.e. exactly the draws of a plain pmo_walk of that
 * plate with stream + i, so the sweep reproduces the sequential runs it replaces: one
 * VoxelPlateModel build/start per parameter set, as the reference's optical-property sweep
 * does (.ipynb_checkpoints/si_in_光学特性推定-checkpoint.ipynb:13945-13967).
 * `inst` is a device array of n_inst descriptors, `schedule` the device copy of
 * pmo_plate_schedule's output. run->n_packets, first_packet and window are ignored. */
int pmo_walk_plates(const pmo_plate_instance* inst, int32_t n_inst, const uint32_t* schedule,
                    int64_t n_batches, const pmo_run* run, const pmo_batch_tallies* tal, void* workspace,
                    void* stream);

#ifdef __cplusplus
}
#endif

#endif /* PMO_H */
