/*
 * A C program that calls libpmo.so through include/pmo.h only: no Python, no PyTorch.
 *
 * It walks the homogeneous slab (101 x 101 x 50 voxels of 0.1 mm, mu_a 0.1 /mm, mu_s 10 /mm,
 * g 0.9, n 1.37, in air; pencil beam at the top-face centre) with FP64 tallies of Rd, Tt,
 * the absorbed map and Rd(r), Tt(r) in 50 bins of 0.1 mm, and prints the raw sums as one
 * JSON line. The geometry is the reference's PlateModel (pymopt/voxel/_classes.py:352-424):
 * a 1-voxel shell of the end label around the interior.
 *
 * usage: walk_slab N_PACKETS SEED
 */
#include <stdio.h>
#include <stdlib.h>
#include <string.h>

#include <hip/hip_runtime_api.h>

#include "pmo.h"

#define NX 103
#define NY 103
#define NZ 52
#define NR 50

#define HIP_OK(call)                                                                   \
    do {                                                                               \
        hipError_t e_ = (call);                                                        \
        if (e_ != hipSuccess) {                                                        \
            fprintf(stderr, "%s: %s\n", #call, hipGetErrorString(e_));                 \
            return 2;                                                                  \
        }                                                                              \
    } while (0)

int main(int argc, char** argv)
{
    if (argc != 3) {
        fprintf(stderr, "usage: %s N_PACKETS SEED\n", argv[0]);
        return 1;
    }
    const unsigned long long n = strtoull(argv[1], NULL, 10);
    const unsigned seed = (unsigned)strtoul(argv[2], NULL, 10);
    if (pmo_abi_version() != PMO_ABI_VERSION) {
        fprintf(stderr, "libpmo ABI %d, header %d\n", pmo_abi_version(), PMO_ABI_VERSION);
        return 1;
    }

    /* label volume: interior label 0, the outer shell the end label 1 */
    const size_t nvox = (size_t)NX * NY * NZ;
    unsigned char* lab = (unsigned char*)malloc(nvox);
    for (int x = 0; x < NX; ++x)
        for (int y = 0; y < NY; ++y)
            for (int z = 0; z < NZ; ++z) {
                const int shell = x == 0 || y == 0 || z == 0 || x == NX - 1 || y == NY - 1 || z == NZ - 1;
                lab[((size_t)x * NY + y) * NZ + z] = (unsigned char)(shell ? 1 : 0);
            }
    const float optics[8] = {0.1f, 10.f, 0.9f, 1.37f, /* ambient */ 0.f, 0.f, 0.f, 1.0f};

    unsigned char* d_lab;
    float* d_opt;
    double *d_scal, *d_abs, *d_rd, *d_tt;
    unsigned long long* d_ev;
    void* d_ws;
    HIP_OK(hipMalloc((void**)&d_lab, nvox));
    HIP_OK(hipMalloc((void**)&d_opt, sizeof(optics)));
    HIP_OK(hipMalloc((void**)&d_scal, PMO_S_COUNT * sizeof(double)));
    HIP_OK(hipMalloc((void**)&d_ev, PMO_EV_COUNT * sizeof(unsigned long long)));
    HIP_OK(hipMalloc((void**)&d_abs, nvox * sizeof(double)));
    HIP_OK(hipMalloc((void**)&d_rd, NR * sizeof(double)));
    HIP_OK(hipMalloc((void**)&d_tt, NR * sizeof(double)));
    HIP_OK(hipMalloc(&d_ws, pmo_workspace_bytes()));
    HIP_OK(hipMemcpy(d_lab, lab, nvox, hipMemcpyHostToDevice));
    HIP_OK(hipMemcpy(d_opt, optics, sizeof(optics), hipMemcpyHostToDevice));
    HIP_OK(hipMemset(d_scal, 0, PMO_S_COUNT * sizeof(double)));
    HIP_OK(hipMemset(d_ev, 0, PMO_EV_COUNT * sizeof(unsigned long long)));
    HIP_OK(hipMemset(d_abs, 0, nvox * sizeof(double)));
    HIP_OK(hipMemset(d_rd, 0, NR * sizeof(double)));
    HIP_OK(hipMemset(d_tt, 0, NR * sizeof(double)));

    pmo_volume vol = {d_lab, NX, NY, NZ, 1, 0.1f, d_opt};
    /* w0 = 1 - Rsp at normal incidence from air (pymopt/voxel/_classes.py:163-175) */
    const double r = (1.0 - 1.37) / (1.0 + 1.37);
    pmo_source src = {PMO_BEAM_PENCIL, (NX - 1) / 2, (NY - 1) / 2, 1, 0.f, {0.f, 0.f, 1.f}, 1.0f - (float)(r * r)};
    pmo_run run;
    memset(&run, 0, sizeof(run));
    run.n_packets = n;
    run.seed = seed;
    run.wth = 1e-4f;
    run.roulette_m = 10.f;
    pmo_tallies tal;
    memset(&tal, 0, sizeof(tal));
    tal.scalars = d_scal;
    tal.events = (uint64_t*)d_ev;
    tal.absorbed = d_abs;
    tal.nr = NR;
    tal.dr = 0.1f;
    tal.rd_r = d_rd;
    tal.tt_r = d_tt;

    if (pmo_walk(&vol, &src, &run, &tal, d_ws, NULL) != PMO_OK) {
        fprintf(stderr, "pmo_walk: %s\n", pmo_last_error());
        return 3;
    }
    HIP_OK(hipDeviceSynchronize());

    double scal[PMO_S_COUNT], rd[NR];
    unsigned long long ev[PMO_EV_COUNT];
    double* absorbed = (double*)malloc(nvox * sizeof(double));
    HIP_OK(hipMemcpy(scal, d_scal, sizeof(scal), hipMemcpyDeviceToHost));
    HIP_OK(hipMemcpy(ev, d_ev, sizeof(ev), hipMemcpyDeviceToHost));
    HIP_OK(hipMemcpy(rd, d_rd, sizeof(rd), hipMemcpyDeviceToHost));
    HIP_OK(hipMemcpy(absorbed, d_abs, nvox * sizeof(double), hipMemcpyDeviceToHost));
    double abs_sum = 0.0;
    for (size_t i = 0; i < nvox; ++i) abs_sum += absorbed[i];

    printf("{\"n\": %llu, \"seed\": %u, \"Rd\": %.17g, \"Tt\": %.17g, \"A\": %.17g, \"w0\": %.17g, "
           "\"absorbed_sum\": %.17g, \"interactions\": %llu, \"crossings\": %llu, \"rd_r\": [",
           n, seed, scal[PMO_S_RD], scal[PMO_S_TT], scal[PMO_S_ABS], scal[PMO_S_W0], abs_sum,
           ev[PMO_EV_INTERACT], ev[PMO_EV_CROSS]);
    for (int i = 0; i < NR; ++i) printf("%s%.17g", i ? ", " : "", rd[i]);
    printf("]}\n");

    hipFree(d_lab); hipFree(d_opt); hipFree(d_scal); hipFree(d_ev); hipFree(d_abs);
    hipFree(d_rd); hipFree(d_tt); hipFree(d_ws);
    free(lab); free(absorbed);
    return 0;
}
