/*
 * asan_check.c -- TEST INFRASTRUCTURE ONLY: a C driver of the oracle built with
 * AddressSanitizer and UndefinedBehaviorSanitizer (make -C oracle asan). It walks small
 * volumes through every source kind and every tally of og_walk_f64 / og_walk_f32, with
 * several threads and packet-id offsets, traces packets with og_trace_f64, and checks the
 * energy ledger of each run. Any out-of-bounds access, use after free or undefined
 * behaviour aborts the program with the sanitizer's report; a ledger miss exits with 1.
 */
#include <math.h>
#include <stdio.h>
#include <stdlib.h>
#include <string.h>

#include "oracle.h"

static int failures = 0;

static void check_ledger(const char* what, const double* s)
{
    const double bal = s[0] + s[1] + s[2] + s[3] + s[4] - s[5];
    if (!(fabs(bal - s[6]) <= 1e-9 * s[6])) {
        fprintf(stderr, "%s: ledger w0 %.17g, balance %.17g\n", what, s[6], bal);
        failures++;
    }
}

/* labels: end shell, interior layers or a pseudo-random mix of n_labels labels */
static uint8_t* make_volume(int nx, int ny, int nz, int n_labels, int mixed, uint32_t seed)
{
    uint8_t* v = (uint8_t*)malloc((size_t)nx * ny * nz);
    uint32_t x = seed * 2654435761u + 1u;
    for (int i = 0; i < nx; i++)
        for (int j = 0; j < ny; j++)
            for (int k = 0; k < nz; k++) {
                uint8_t lab;
                if (i == 0 || j == 0 || k == 0 || i == nx - 1 || j == ny - 1 || k == nz - 1) lab = (uint8_t)n_labels;
                else if (mixed) { x ^= x << 13; x ^= x >> 17; x ^= x << 5; lab = (uint8_t)(x % (uint32_t)n_labels); }
                else lab = (uint8_t)(((k - 1) * n_labels) / (nz - 2));
                v[((size_t)i * ny + j) * nz + k] = lab;
            }
    return v;
}

static void run_case(const char* name, int nx, int ny, int nz, int n_labels, int mixed, int kind, int f32,
                     uint64_t n_packets, uint64_t first, int threads, uint32_t flags)
{
    double mua[4] = {0.1, 0.5, 0.02, 1.0}, mus[4] = {10.0, 30.0, 12.0, 5.0}, g[4] = {0.9, 0.8, 0.0, 0.95};
    double n[5] = {1.37, 1.44, 1.40, 1.55, 1.0};
    n[n_labels] = 1.0;
    uint8_t* lab = make_volume(nx, ny, nz, n_labels, mixed, (uint32_t)n_packets);
    og_volume V = {lab, nx, ny, nz, n_labels, 0.1, mua, mus, g, n};
    og_source S;
    memset(&S, 0, sizeof(S));
    S.kind = kind == 3 ? 3 : kind;
    S.ix = (nx - 1) / 2; S.iy = (ny - 1) / 2; S.iz = 1;
    S.radius = 0.3; S.dir[2] = 1.0; S.w0 = 0.97;
    int32_t* add = NULL; float *p = NULL, *v = NULL, *w = NULL;
    if (kind == 3) {
        /* per-packet launch states, some outside the volume and some on voxel faces */
        add = (int32_t*)malloc(3 * n_packets * sizeof(int32_t));
        p = (float*)malloc(3 * n_packets * sizeof(float));
        v = (float*)malloc(3 * n_packets * sizeof(float));
        w = (float*)malloc(n_packets * sizeof(float));
        for (uint64_t i = 0; i < n_packets; i++) {
            add[i] = (int32_t)(i % 7 == 0 ? -1 : S.ix + (int)(i % 3) - 1);
            add[n_packets + i] = S.iy; add[2 * n_packets + i] = 1 + (int32_t)(i % 2);
            p[i] = (float)((i % 5) * 0.025 - 0.05); p[n_packets + i] = 0.05f; p[2 * n_packets + i] = -0.05f;
            const double th = 0.1 * (double)(i % 9);
            v[i] = (float)sin(th); v[n_packets + i] = 0.f; v[2 * n_packets + i] = (float)cos(th);
            w[i] = 0.9f;
        }
        S.add = add; S.p = p; S.v = v; S.w = w; S.states_stride = n_packets;
    }
    og_run R = {n_packets, first, 7u, 3u, 1e-4, 10.0, flags, threads};
    og_tallies T;
    memset(&T, 0, sizeof(T));
    const size_t nvox = (size_t)nx * ny * nz;
    double rings[4] = {0.1, 0.3, 0.3, 0.6};
    T.absorbed = (double*)calloc(nvox, sizeof(double));
    T.nr = 12; T.dr = 0.05;
    T.rd_r = (double*)calloc(12, sizeof(double)); T.tt_r = (double*)calloc(12, sizeof(double));
    T.rd_top_r = (double*)calloc(12, sizeof(double));
    T.na = 9; T.rd_a = (double*)calloc(9, sizeof(double)); T.tt_a = (double*)calloc(9, sizeof(double));
    T.nl = 10; T.dl = 0.1; T.line_y = 0.5; T.line_z = (double*)calloc(10, sizeof(double));
    T.n_rings = 2; T.rings = rings; T.ring_w = (double*)calloc(2, sizeof(double));
    T.exit_state = (float*)calloc(7 * n_packets, sizeof(float));
    T.rd_wl = (double*)calloc(12 * (size_t)n_labels, sizeof(double));
    T.rd_wn = (double*)calloc(12 * (size_t)n_labels, sizeof(double));
    T.rz_nr = 6; T.rz_nz = 7; T.rz_dr = 0.1; T.rz_dz = 0.1;
    T.fluence_rz = (double*)calloc(42, sizeof(double));
    const int rc = f32 ? og_walk_f32(&V, &S, &R, &T) : og_walk_f64(&V, &S, &R, &T);
    if (rc != 0) { fprintf(stderr, "%s: og_walk returned %d\n", name, rc); failures++; }
    check_ledger(name, T.scalars);
    if (T.events[0] != n_packets) { fprintf(stderr, "%s: %llu packets\n", name, (unsigned long long)T.events[0]); failures++; }
    /* traces of a few packets into a short buffer (the cap is honoured) */
    og_event ev[64];
    for (uint64_t i = 0; i < 4 && i < n_packets; i++) {
        const int m = og_trace_f64(&V, &S, &R, first + i, ev, 64);
        if (m < 1 || m > 64) { fprintf(stderr, "%s: trace returned %d\n", name, m); failures++; }
    }
    free(T.absorbed); free(T.rd_r); free(T.tt_r); free(T.rd_top_r); free(T.rd_a); free(T.tt_a); free(T.line_z);
    free(T.ring_w); free(T.exit_state); free(T.rd_wl); free(T.rd_wn); free(T.fluence_rz);
    free(add); free(p); free(v); free(w);
    free(lab);
    printf("%-28s ok\n", name);
}

int main(void)
{
    /* primitives on edge inputs */
    double ct;
    (void)og_fresnel(1.0, 1.5, 1.0, &ct);
    (void)og_fresnel(1.5, 1.0, 0.1, &ct);      /* total internal reflection */
    (void)og_hg_cos(0.0, 0.5); (void)og_hg_cos(0.99, 1.0); (void)og_hg_cos(0.9, 1e-7);
    double vv[3] = {0, 0, 1};
    og_spin(vv, 0.3, 1.0);
    double pp[3] = {0.05, 0, -0.05}, dd[3] = {1, 0, 0};
    int32_t k;
    (void)og_boundary(pp, dd, 0.05, &k);

    run_case("slab pencil f64", 9, 9, 8, 1, 0, 0, 0, 3000, 0, 2, 0);
    run_case("3 layers flat f64", 11, 11, 12, 3, 0, 1, 0, 3000, 100, 3, 0);
    run_case("3 layers gauss f32", 11, 11, 12, 3, 0, 2, 1, 3000, 0, 2, 2u);
    run_case("mixed 4 labels f64", 10, 12, 9, 4, 1, 0, 0, 4000, 7, 4, 0);
    run_case("mixed 4 labels f32", 10, 12, 9, 4, 1, 0, 1, 4000, 0, 1, 2u);
    run_case("launch states f64", 9, 9, 8, 2, 0, 3, 0, 700, 0, 2, 0);
    run_case("launch states f32", 9, 9, 8, 2, 1, 3, 1, 700, 0, 2, 0);
    run_case("3x3x3 volume f64", 3, 3, 3, 1, 0, 0, 0, 200, 0, 1, 0);
    if (failures) {
        fprintf(stderr, "%d failures\n", failures);
        return 1;
    }
    printf("all cases ok\n");
    return 0;
}
