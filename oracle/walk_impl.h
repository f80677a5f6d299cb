/*
 * walk_impl.h -- TEST INFRASTRUCTURE ONLY (see oracle.h).
 * Included twice by oracle.c with REAL = double / float and SFX = f64 / f32.
 *
 * Statement-level restatement of pymopt/voxel/_kernel.py:36-165. Line references in the
 * comments below are to that file unless stated otherwise.
 */

#define CAT2(a, b) a##_##b
#define CAT(a, b) CAT2(a, b)
#define FN(name) CAT(name, SFX)

typedef struct FN(ctx) {
    double scal[13];
    uint64_t ev[7];
    double* absorbed;
    double* rd_r;
    double* tt_r;
    double* rd_top_r;
    double* rd_a;
    double* tt_a;
    double* line_z;
    double* ring_w;
    double* rd_wl;
    double* rd_wn;
    double* frz;
} FN(ctx);

static inline REAL FN(rsgn)(REAL x) { return x > 0 ? (REAL)1 : (x < 0 ? (REAL)-1 : (REAL)0); }

/* Fresnel reflectance exactly as _kernel.py:84-93 writes it (acos/asin/sin/tan). */
static REAL FN(fresnel_lit)(REAL ni, REAL nt, REAL avk, REAL* at_out, int normal_limit)
{
    REAL ai = (REAL)acos((double)avk); /* :84 */
    REAL at = 0, ra;
    if ((ni > nt) && (ai >= (REAL)asin((double)(nt / ni)))) { /* :85 total internal reflection */
        ra = 1;
    } else {
        at = (REAL)asin((double)((ni / nt) * (REAL)sin((double)ai))); /* :88 */
        if (ai != 0) {                                               /* :89-91 */
            REAL a = (REAL)sin((double)(ai - at)) / (REAL)sin((double)(ai + at));
            REAL b = (REAL)tan((double)(ai - at)) / (REAL)tan((double)(ai + at));
            ra = (a * a + b * b) / 2;
        } else {                                                     /* :92-93 reference: R = 0 */
            ra = normal_limit ? ((ni - nt) / (ni + nt)) * ((ni - nt) / (ni + nt)) : 0;
        }
    }
    *at_out = at;
    return ra;
}

/* Henyey-Greenstein cos(theta), _kernel.py:12-21; the clamp is applied on both sides
   (the reference clamps below only, a rounding overshoot above 1 gives NaN there). */
static inline REAL FN(hg)(REAL g, REAL u)
{
    REAL th;
    if (g != 0) {
        REAL t = (1 - g * g) / (1 - g + 2 * g * u);
        th = (1 + g * g - t * t) / (2 * g);
        if (th < -1) th = -1;
        if (th > 1) th = 1;
    } else {
        th = 2 * u - 1;
    }
    return th;
}

/* direction update of _kernel.py:150-165 */
static inline void FN(spin)(REAL v[3], REAL cos_th, REAL phi)
{
    REAL sin_th = (REAL)sqrt((double)(1 - cos_th * cos_th));
    REAL cos_fi = (REAL)cos((double)phi), sin_fi = (REAL)sin((double)phi);
    if ((REAL)0.99999 < (REAL)fabs((double)v[2])) {
        v[0] = sin_th * cos_fi;
        v[1] = sin_th * sin_fi;
        v[2] = FN(rsgn)(v[2]) * cos_th;
    } else {
        REAL valf = (REAL)sqrt((double)(1 - v[2] * v[2]));
        REAL v0 = v[0], v1 = v[1], v2 = v[2];
        v[0] = sin_th * (v0 * v2 * cos_fi - v1 * sin_fi) / valf + v0 * cos_th;
        v[1] = sin_th * (v1 * v2 * cos_fi + v0 * sin_fi) / valf + v1 * cos_th;
        v[2] = -sin_th * cos_fi * valf + v2 * cos_th;
    }
    REAL nv = (REAL)sqrt((double)(v[0] * v[0] + v[1] * v[1] + v[2] * v[2]));
    v[0] /= nv; v[1] /= nv; v[2] /= nv;
}

/* distance to the exit face of the current voxel, _kernel.py:63-68 (x < y < z on ties) */
static inline REAL FN(boundary)(const REAL p[3], const REAL v[3], REAL half, int* k)
{
    REAL db = 1000;
    *k = 0;
    for (int i = 0; i < 3; i++) {
        if ((REAL)fabs((double)v[i]) > 0) {
            REAL valf = (half - FN(rsgn)(v[i]) * p[i]) / (REAL)fabs((double)v[i]);
            if (valf < db) { *k = i; db = valf; }
        }
    }
    return db;
}

/* Philox "A" block i of packet pid: the four uniforms used by hop i and interaction i */
static inline void FN(ablock)(const og_run* R, uint64_t pid, uint32_t i, REAL u[4])
{
    uint32_t c[4] = {i, 0u, (uint32_t)pid, (uint32_t)(pid >> 32)}, k[2] = {R->seed, R->stream}, o[4];
    og_philox4x32_10(c, k, o);
    for (int j = 0; j < 4; j++) u[j] = (REAL)og_u01(o[j]);
}

static inline REAL FN(bdraw)(const og_run* R, uint64_t pid, uint32_t f)
{
    uint32_t c[4] = {f >> 2, 1u, (uint32_t)pid, (uint32_t)(pid >> 32)}, k[2] = {R->seed, R->stream}, o[4];
    og_philox4x32_10(c, k, o);
    return (REAL)og_u01(o[f & 3u]);
}

static inline int64_t FN(lin)(const og_volume* V, int ax, int ay, int az)
{
    return ((int64_t)ax * V->ny + ay) * V->nz + az;
}

static inline int FN(bin)(double x, double width, int nbins)
{
    /* bins (i*width, (i+1)*width]; x == 0 falls in none (reference alpha[i] < r test) */
    if (!(x > 0)) return -1;
    double q = x / width;
    int64_t b = (int64_t)ceil(q) - 1;
    if (b < 0 || b >= nbins) return -1;
    return (int)b;
}

static void FN(record_exit)(const og_volume* V, const og_tallies* T, FN(ctx)* C, uint64_t slot,
                            uint64_t n_packets, const int a[3], const REAL p[3], const REAL v[3],
                            REAL w, int k, int s, const double* Lj, const double* nj)
{
    double l = V->pitch;
    int cx = (V->nx - 1) / 2, cy = (V->ny - 1) / 2;       /* _get_center_add, voxel/_classes.py:124-128 */
    double x = l * (a[0] - cx) + (double)p[0];              /* _encooder, voxel/_classes.py:261-269 */
    double y = l * (a[1] - cy) + (double)p[1];
    /* np.round(.., 6) of the reference's z: exits through a face lie exactly on it */
    double z = nearbyint((l * (a[2] - 1) + (double)p[2] + l / 2) * 1e6) / 1e6;
    double wd = (double)w;
    if (v[2] < 0) C->scal[0] += wd;                         /* getRdTtRate, voxel/_classes.py:285-293 */
    else if (v[2] > 0) C->scal[1] += wd;
    else C->scal[2] += wd;
    C->scal[7 + 2 * k + (s > 0 ? 1 : 0)] += wd;
    /* top-face detectors: virtualOBD.py:165 selects v_z < 0 and z < 1e-8 on the encoded exit */
    int top = (v[2] < 0 && z < 1e-8);
    double r = sqrt(x * x + y * y);
    if (T->nr > 0) {
        int b = FN(bin)(r, T->dr, T->nr);                   /* radialDistance, metrics/_scoring.py:16-25 */
        if (b >= 0) {
            if (v[2] < 0 && C->rd_r) C->rd_r[b] += wd;
            if (v[2] > 0 && C->tt_r) C->tt_r[b] += wd;
            if (top && C->rd_top_r) C->rd_top_r[b] += wd;
            if (v[2] < 0 && C->rd_wl) {
                for (int j = 0; j < V->n_labels; j++) {
                    C->rd_wl[(int64_t)b * V->n_labels + j] += wd * Lj[j];
                    C->rd_wn[(int64_t)b * V->n_labels + j] += wd * nj[j];
                }
            }
        }
    }
    if (T->na > 0) {
        double ang = acos(fabs((double)v[2]));              /* angularyResolved, _scoring.py:27-38 */
        int b = FN(bin)(ang, M_PI / (2.0 * T->na), T->na);
        if (b >= 0) {
            if (v[2] < 0 && C->rd_a) C->rd_a[b] += wd;
            if (v[2] > 0 && C->tt_a) C->tt_a[b] += wd;
        }
    }
    if (T->nl > 0 && C->line_z && v[0] < 0 && fabs(y) < T->line_y) { /* virtualOBD.py:171-172 */
        int b = FN(bin)(z, T->dl, T->nl);
        if (b >= 0) C->line_z[b] += wd;
    }
    if (top && T->n_rings > 0 && C->ring_w) {
        for (int i = 0; i < T->n_rings; i++)
            if (r > T->rings[2 * i] && r <= T->rings[2 * i + 1]) C->ring_w[i] += wd;
    }
    if (T->exit_state) {
        float* e = T->exit_state;
        e[0 * n_packets + slot] = (float)x; e[1 * n_packets + slot] = (float)y;
        e[2 * n_packets + slot] = (float)z; e[3 * n_packets + slot] = (float)v[0];
        e[4 * n_packets + slot] = (float)v[1]; e[5 * n_packets + slot] = (float)v[2];
        e[6 * n_packets + slot] = (float)w;
    }
}

/* One packet. When `trace` is non-NULL, events are written there and tallies untouched. */
static int FN(packet)(const og_volume* V, const og_source* S, const og_run* R, const og_tallies* T,
                      FN(ctx)* C, uint64_t pid, og_event* trace, int cap)
{
    const int end = V->n_labels;
    const REAL l = (REAL)V->pitch, half = l / 2;
    const uint64_t slot = pid - R->first_packet;
    int ntr = 0;
    int a[3] = {S->ix, S->iy, S->iz};
    REAL p[3] = {0, 0, -half}, v[3] = {(REAL)S->dir[0], (REAL)S->dir[1], (REAL)S->dir[2]};
    double Lj[256], nj[256];
    const int want_path = (C && C->rd_wl) ? 1 : 0;
    if (want_path) for (int j = 0; j < V->n_labels; j++) { Lj[j] = 0; nj[j] = 0; }

    if (S->kind == 3) {                                  /* the caller's launch state */
        const uint64_t sd = S->states_stride;
        for (int c = 0; c < 3; c++) {
            a[c] = S->add[c * sd + slot];
            p[c] = (REAL)S->p[c * sd + slot];
            if (p[c] < -half) p[c] = -half;
            if (p[c] > half) p[c] = half;
            v[c] = (REAL)S->v[c * sd + slot];
        }
    } else if (S->kind == 1 || S->kind == 2) {           /* flat disc / TEM00 lateral position */
        uint32_t c[4] = {0u, 2u, (uint32_t)pid, (uint32_t)(pid >> 32)}, k[2] = {R->seed, R->stream}, o[4];
        og_philox4x32_10(c, k, o);
        REAL u0 = (REAL)og_u01(o[0]), u1 = (REAL)og_u01(o[1]);
        REAL rad, x, y;
        if (S->kind == 1) rad = (REAL)S->radius * (REAL)sqrt((double)u0);
        else rad = (REAL)(S->radius / 2) * (REAL)sqrt(-2.0 * log((double)u0));
        REAL ang = (REAL)(2.0 * M_PI) * u1;
        x = rad * (REAL)cos((double)ang);
        y = rad * (REAL)sin((double)ang);
        int dx = (int)floor((double)(x / l + (REAL)0.5)), dy = (int)floor((double)(y / l + (REAL)0.5));
        a[0] += dx; a[1] += dy;
        p[0] = x - l * (REAL)dx;
        p[1] = y - l * (REAL)dy;
        /* rounding of x / l can put the position a hair outside the voxel: clamp */
        if (p[0] < -half) p[0] = -half;
        if (p[0] > half) p[0] = half;
        if (p[1] < -half) p[1] = -half;
        if (p[1] > half) p[1] = half;
    }
    REAL w = S->kind == 3 ? (REAL)S->w[slot] : (REAL)S->w0;
    if (C) { C->scal[6] += (double)w; C->ev[0]++; }
    if (trace && ntr < cap) {
        og_event* e = &trace[ntr++];
        e->type = 0; e->ax = a[0]; e->ay = a[1]; e->az = a[2];
        for (int j = 0; j < 3; j++) { e->p[j] = p[j]; e->v[j] = v[j]; }
        e->w = w;
    }

    /* :44; a caller's address outside the volume leaves at once (as in libpmo) */
    const int inside = a[0] >= 0 && a[0] < V->nx && a[1] >= 0 && a[1] < V->ny && a[2] >= 0 && a[2] < V->nz;
    int lab = inside ? V->labels[FN(lin)(V, a[0], a[1], a[2])] : end;
    if (lab >= end) {                    /* launched in an end voxel: leaves along v_z at once */
        if (C) {
            C->ev[6]++;
            FN(record_exit)(V, T, C, slot, R->n_packets, a, p, v, w, 2, v[2] > 0 ? 1 : -1, Lj, nj);
        }
        return ntr;
    }
    REAL ni = (REAL)V->n[lab];                              /* :54 */
    REAL mt = (REAL)V->mua[lab] + (REAL)V->mus[lab];        /* :56 */
    uint32_t inter = 0, fres = 0;
    REAL u[4];
    FN(ablock)(R, pid, 0u, u);

    for (;;) {                                              /* :60 interaction loop */
        REAL st = -(REAL)log((double)u[0]);                 /* :61 */
        for (;;) {                                          /* :62 crossing loop */
            int k;
            REAL db = FN(boundary)(p, v, half, &k);         /* :63-68 */
            if (st >= db * mt) {                            /* :69 */
                for (int i = 0; i < 3; i++) p[i] += v[i] * db;   /* :70-71 */
                if (want_path) Lj[lab] += (double)db;
                p[k] = half * FN(rsgn)(v[k]);               /* :72 */
                st -= db * mt;                              /* :73 */
                int s = v[k] > 0 ? 1 : -1;
                int b[3] = {a[0], a[1], a[2]};
                b[k] += s;                                  /* :74-76 */
                int nxt = V->labels[FN(lin)(V, b[0], b[1], b[2])]; /* :77 */
                REAL nt = (REAL)V->n[nxt];                  /* :79 */
                int transmit = 1;
                if (C) C->ev[2]++;
                if (ni != nt) {                             /* :82 */
                    REAL at;
                    REAL ra = FN(fresnel_lit)(ni, nt, (REAL)fabs((double)v[k]), &at, (R->flags >> 1) & 1u);
                    REAL rnd = FN(bdraw)(R, pid, fres);     /* :94 */
                    fres++;
                    if (C) C->ev[3]++;
                    if (ra < rnd) {                         /* :96-106 transmit with refraction */
                        for (int i = 0; i < 3; i++) {
                            if (i == k) v[i] = FN(rsgn)(v[i]) * (REAL)cos((double)at);
                            else v[i] = v[i] * ni / nt;
                        }
                        REAL nv = (REAL)sqrt((double)(v[0] * v[0] + v[1] * v[1] + v[2] * v[2]));
                        for (int i = 0; i < 3; i++) v[i] /= nv;
                    } else {
                        transmit = 0;                       /* :107-108 */
                    }
                }
                if (transmit) {                             /* :110-118 */
                    a[k] += s;
                    p[k] = -p[k];
                    lab = nxt;
                    if (trace && ntr < cap) {
                        og_event* e = &trace[ntr++];
                        e->type = (lab == end) ? 4 : 1; e->ax = a[0]; e->ay = a[1]; e->az = a[2];
                        for (int j = 0; j < 3; j++) { e->p[j] = p[j]; e->v[j] = v[j]; }
                        e->w = w;
                    }
                    if (lab == end) {
                        if (C) {
                            C->ev[6]++;
                            FN(record_exit)(V, T, C, slot, R->n_packets, a, p, v, w, k, s, Lj, nj);
                        }
                        return ntr;
                    }
                    mt = (REAL)V->mua[lab] + (REAL)V->mus[lab];
                    ni = nt;
                } else {
                    v[k] = -v[k];                           /* :119-120 Fresnel reflection */
                    if (trace && ntr < cap) {
                        og_event* e = &trace[ntr++];
                        e->type = 2; e->ax = a[0]; e->ay = a[1]; e->az = a[2];
                        for (int j = 0; j < 3; j++) { e->p[j] = p[j]; e->v[j] = v[j]; }
                        e->w = w;
                    }
                }
            } else {                                        /* :121-125 hop ends in the voxel */
                for (int i = 0; i < 3; i++) p[i] += v[i] * st / mt;
                if (want_path) { Lj[lab] += (double)(st / mt); nj[lab] += 1.0; }
                st = 0;
                break;
            }
        }
        /* drop, :130; deposit = exact difference of the two float weights */
        REAL wn = w - w * (REAL)V->mua[lab] / mt;
        double dep = (double)w - (double)wn;
        w = wn;
        if (C) {
            C->ev[1]++;
            C->scal[3] += dep;
            if (C->absorbed) C->absorbed[FN(lin)(V, a[0], a[1], a[2])] += dep;
            if (C->frz && V->mua[lab] > 0) {                 /* r-z fluence at the interaction */
                double l64 = V->pitch;
                int cx = (V->nx - 1) / 2, cy = (V->ny - 1) / 2;
                double x = l64 * (a[0] - cx) + (double)p[0];
                double y = l64 * (a[1] - cy) + (double)p[1];
                double z = l64 * (a[2] - 1) + (double)p[2] + l64 / 2;
                int64_t ir = (int64_t)floor(sqrt(x * x + y * y) / T->rz_dr);
                int64_t iz = (int64_t)floor(z / T->rz_dz);
                if (ir >= 0 && ir < T->rz_nr && iz >= 0 && iz < T->rz_nz)
                    C->frz[ir * T->rz_nz + iz] += dep / V->mua[lab];
            }
        }
        if (w <= (REAL)R->wth) {                            /* roulette, :132-141 */
            if (C) C->ev[4]++;
            if ((REAL)(1.0 / R->roulette_m) < u[1]) {
                if (C) {
                    C->scal[4] += (double)w;
                    C->ev[5]++;
                    if (T->exit_state) {                   /* reference zeroes add/p/v/w of a killed packet */
                        double l64 = V->pitch;
                        int cx = (V->nx - 1) / 2, cy = (V->ny - 1) / 2;
                        float* e = T->exit_state;
                        uint64_t n = R->n_packets;
                        e[0 * n + slot] = (float)(-l64 * cx); e[1 * n + slot] = (float)(-l64 * cy);
                        e[2 * n + slot] = (float)(-l64 / 2); e[3 * n + slot] = 0.f;
                        e[4 * n + slot] = 0.f; e[5 * n + slot] = 0.f; e[6 * n + slot] = 0.f;
                    }
                }
                if (trace && ntr < cap) {
                    og_event* e = &trace[ntr++];
                    e->type = 5; e->ax = a[0]; e->ay = a[1]; e->az = a[2];
                    for (int j = 0; j < 3; j++) { e->p[j] = p[j]; e->v[j] = v[j]; }
                    e->w = w;
                }
                return ntr;
            } else {
                REAL wm = w * (REAL)R->roulette_m;
                if (C) C->scal[5] += (double)wm - (double)w;
                w = wm;
            }
        }
        REAL cos_th = FN(hg)((REAL)V->g[lab], u[2]);       /* :143 */
        REAL fi = 2 * (REAL)3.1415927 * u[3];               /* :146 */
        FN(spin)(v, cos_th, fi);                            /* :144-165 */
        if (trace && ntr < cap) {                           /* interaction: post-spin state */
            og_event* e = &trace[ntr++];
            e->type = 3; e->ax = a[0]; e->ay = a[1]; e->az = a[2];
            for (int j = 0; j < 3; j++) { e->p[j] = p[j]; e->v[j] = v[j]; }
            e->w = w;
        }
        inter++;
        FN(ablock)(R, pid, inter, u);
    }
}

int FN(og_walk)(const og_volume* V, const og_source* S, const og_run* R, og_tallies* T)
{
    if (!V || !S || !R || !T || !V->labels) return -1;
    if (V->n_labels < 1 || V->n_labels > 255) return -2;
    const int64_t nvox = (int64_t)V->nx * V->ny * V->nz;
    int nthreads = R->n_threads > 0 ? R->n_threads : omp_get_max_threads();
    FN(ctx)* ctxs = (FN(ctx)*)calloc((size_t)nthreads, sizeof(FN(ctx)));
    if (!ctxs) return -3;
    const int nl = V->n_labels;
    for (int t = 0; t < nthreads; t++) {
        FN(ctx)* C = &ctxs[t];
        C->absorbed = T->absorbed ? (double*)calloc((size_t)nvox, sizeof(double)) : NULL;
        C->rd_r = (T->nr > 0 && T->rd_r) ? (double*)calloc((size_t)T->nr, sizeof(double)) : NULL;
        C->tt_r = (T->nr > 0 && T->tt_r) ? (double*)calloc((size_t)T->nr, sizeof(double)) : NULL;
        C->rd_top_r = (T->nr > 0 && T->rd_top_r) ? (double*)calloc((size_t)T->nr, sizeof(double)) : NULL;
        C->rd_a = (T->na > 0 && T->rd_a) ? (double*)calloc((size_t)T->na, sizeof(double)) : NULL;
        C->tt_a = (T->na > 0 && T->tt_a) ? (double*)calloc((size_t)T->na, sizeof(double)) : NULL;
        C->line_z = (T->nl > 0 && T->line_z) ? (double*)calloc((size_t)T->nl, sizeof(double)) : NULL;
        C->ring_w = (T->n_rings > 0 && T->ring_w) ? (double*)calloc((size_t)T->n_rings, sizeof(double)) : NULL;
        C->rd_wl = (T->nr > 0 && T->rd_wl) ? (double*)calloc((size_t)T->nr * nl, sizeof(double)) : NULL;
        C->rd_wn = (T->nr > 0 && T->rd_wn) ? (double*)calloc((size_t)T->nr * nl, sizeof(double)) : NULL;
        C->frz = (T->rz_nr > 0 && T->rz_nz > 0 && T->fluence_rz)
                     ? (double*)calloc((size_t)T->rz_nr * T->rz_nz, sizeof(double)) : NULL;
    }
    const int64_t n = (int64_t)R->n_packets;
#pragma omp parallel for num_threads(nthreads) schedule(dynamic, 64)
    for (int64_t i = 0; i < n; i++) {
        FN(ctx)* C = &ctxs[omp_get_thread_num()];
        FN(packet)(V, S, R, T, C, R->first_packet + (uint64_t)i, NULL, 0);
    }
    /* deterministic merge in thread order */
    for (int t = 0; t < nthreads; t++) {
        FN(ctx)* C = &ctxs[t];
        for (int j = 0; j < 13; j++) T->scalars[j] += C->scal[j];
        for (int j = 0; j < 7; j++) T->events[j] += C->ev[j];
#define MERGE(buf, cnt) if (C->buf) { for (int64_t q = 0; q < (int64_t)(cnt); q++) T->buf[q] += C->buf[q]; free(C->buf); }
        MERGE(absorbed, nvox)
        MERGE(rd_r, T->nr)
        MERGE(tt_r, T->nr)
        MERGE(rd_top_r, T->nr)
        MERGE(rd_a, T->na)
        MERGE(tt_a, T->na)
        MERGE(line_z, T->nl)
        MERGE(ring_w, T->n_rings)
        MERGE(rd_wl, (int64_t)T->nr * nl)
        MERGE(rd_wn, (int64_t)T->nr * nl)
        if (C->frz) {
            for (int64_t q = 0; q < (int64_t)T->rz_nr * T->rz_nz; q++) T->fluence_rz[q] += C->frz[q];
            free(C->frz);
        }
#undef MERGE
    }
    free(ctxs);
    return 0;
}

#undef FN
#undef CAT
#undef CAT2
