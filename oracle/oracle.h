/*
 * oracle.h -- TEST INFRASTRUCTURE ONLY.
 *
 * CPU restatement of pyMonteOpt's voxel photon walk, used as the parity checker of the
 * HIP walk and as the timed CPU baseline. Only tests/, __graft_entry__.smoke() and
 * bench.py's cpu_baseline leg may load it, and only as the checker: nothing in the
 * product path (pymonteopt_amd/) links, imports or calls it.
 *
 * It follows pymopt/voxel/_kernel.py:23-172 (the Numba walk) statement by statement,
 * with the literal acos/asin/sin/tan Fresnel of :84-93, and adds what the reference does
 * not have: on-line FP64 tallies (same definitions as include/pmo.h), an exact energy
 * ledger and event counters. The random numbers come from the same counter-based
 * Philox4x32-10 assignment as the HIP walk (packet id + draw index), so the two can be
 * compared packet-for-packet as well as statistically.
 *
 * Two builds of one source: _f64 (double state, the oracle proper) and _f32 (float
 * state and arithmetic, the reference's float32 semantics of
 * pymopt/voxel_gpu/_cukernel.py:27-227).
 */
#ifndef PMO_ORACLE_H
#define PMO_ORACLE_H

#include <stdint.h>

#ifdef __cplusplus
extern "C" {
#endif

typedef struct og_volume {
    const uint8_t* labels; /* C order [nx][ny][nz]                                   */
    int32_t nx, ny, nz;
    int32_t n_labels;      /* end label = n_labels                                   */
    double pitch;
    const double* mua;     /* [n_labels]                                             */
    const double* mus;     /* [n_labels]                                             */
    const double* g;       /* [n_labels]                                             */
    const double* n;       /* [n_labels + 1], n[n_labels] = ambient                  */
} og_volume;

typedef struct og_source {
    int32_t kind;          /* 0 pencil, 1 flat disc, 2 gaussian, 3 per-packet states */
    int32_t ix, iy, iz;
    double radius;
    double dir[3];
    double w0;
    /* kind 3: the reference kernel's launch arrays (pymopt/voxel/_kernel.py:23-28), packet
       i's component c at [c * states_stride + i], i = packet id - first_packet */
    const int32_t* add;
    const float* p;
    const float* v;
    const float* w;
    uint64_t states_stride;
} og_source;

typedef struct og_run {
    uint64_t n_packets;
    uint64_t first_packet;
    uint32_t seed, stream;
    double wth, roulette_m;
    uint32_t flags;        /* bit 1: Fresnel limit at normal incidence               */
    int32_t n_threads;     /* OpenMP threads, 0 = library default                    */
} og_run;

typedef struct og_tallies {
    double scalars[13];    /* PMO_S_* layout                                         */
    uint64_t events[7];    /* PMO_EV_* layout                                        */
    double* absorbed;      /* [nx*ny*nz] or NULL                                     */
    int32_t nr;
    double dr;
    double* rd_r;
    double* tt_r;
    double* rd_top_r;
    int32_t na;
    double* rd_a;
    double* tt_a;
    int32_t nl;
    double dl, line_y;
    double* line_z;
    int32_t n_rings;
    const double* rings;   /* [2*n_rings]                                            */
    double* ring_w;
    float* exit_state;     /* [7][n_packets] or NULL                                 */
    /* per-label path statistics of packets escaping with v_z < 0, binned like rd_r:
       sum of w*L_j and w*n_j (L_j path length, n_j interactions in label j) */
    double* rd_wl;         /* [nr][n_labels] or NULL                                 */
    double* rd_wn;         /* [nr][n_labels] or NULL                                 */
    /* internal fluence in r-z bins [i*rz_dr, (i+1)*rz_dr) x [k*rz_dz, (k+1)*rz_dz): sum of
       deposit / mu_a at each interaction's position (the intent of
       pymopt/fluence/intarnal_fluence.py Fluence2D, pymopt/voxel_unverified/_classes.py:394-399) */
    int32_t rz_nr, rz_nz;
    double rz_dr, rz_dz;
    double* fluence_rz;    /* [rz_nr][rz_nz] or NULL                                 */
} og_tallies;

/* event trace of one packet (type, voxel, position, direction, weight) */
typedef struct og_event {
    int32_t type;          /* 0 launch, 1 transmit, 2 reflect, 3 interaction, 4 escape, 5 kill */
    int32_t ax, ay, az;
    double p[3];
    double v[3];
    double w;
} og_event;

int og_walk_f64(const og_volume* vol, const og_source* src, const og_run* run, og_tallies* tal);
int og_walk_f32(const og_volume* vol, const og_source* src, const og_run* run, og_tallies* tal);
/* returns the number of events written (<= cap) for packet `pid` */
int og_trace_f64(const og_volume* vol, const og_source* src, const og_run* run, uint64_t pid,
                 og_event* ev, int32_t cap);

/* primitives, for pinning each piece against closed forms */
void og_philox4x32_10(const uint32_t ctr[4], const uint32_t key[2], uint32_t out[4]);
double og_u01(uint32_t x);
double og_fresnel(double ni, double nt, double abs_vk, double* cos_t);  /* literal reference form */
double og_hg_cos(double g, double u);
void og_spin(double v[3], double cos_th, double phi);                   /* reference rotation      */
double og_boundary(const double p[3], const double v[3], double half, int32_t* k);

#ifdef __cplusplus
}
#endif

#endif
