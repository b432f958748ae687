/*
 * ktm_oracle.c -- CPU oracle (test infrastructure) for the Kirchhoff
 * prestack time-migration path of Kirchhoff.c.  See ktm_oracle.h for the
 * contract and the list of checks that pin each function.
 *
 * Build with -ffp-contract=off and without -ffast-math: every expression
 * below is written so that its promotions and rounding points are those of
 * the corresponding line of Kirchhoff.c under strict IEEE evaluation.
 */
#include "ktm_oracle.h"

#include <math.h>
#include <stdlib.h>
#include <string.h>

#define ORC_PI 3.14159265359 /* Kirchhoff.c:97 (not M_PI) */

void orc_grid_init(const orc_params *p, orc_grid *g)
{
    /* Kirchhoff.c:266-277 */
    int Nx = 1, Ny = 1;
    if (p->inc_iline != 0.0)
        Nx = (int)ceil((p->end_iline - p->start_iline) / p->inc_iline);
    if (p->inc_xline != 0.0)
        Ny = (int)ceil((p->end_xline - p->start_xline) / p->inc_xline);
    if (Nx == 0) Nx = 1;
    if (Ny == 0) Ny = 1;
    /* Kirchhoff.c:280-300 */
    g->no_job = p->jobs_x * p->jobs_y;
    g->nx = (int)ceil(1.0 * Nx / p->jobs_x);
    g->ny = (int)ceil(1.0 * Ny / p->jobs_y);
    if (g->nx == 0) g->nx = 1;
    if (g->ny == 0) g->ny = 1;
    g->Nx = Nx;
    g->Ny = Ny;
    /* Kirchhoff.c:312-313 */
    g->sinA = (float)sin(p->angle_e_to_n * ORC_PI / 180.0);
    g->cosA = (float)cos(p->angle_e_to_n * ORC_PI / 180.0);
}

/* ------------------------------------------------------------------ A4 -- */
float orc_ibm_to_float(const uint8_t be[4])
{
    /* ntohf (Kirchhoff.c:882-895) reverses the four file bytes, so the int
     * view of the swapped float is the big-endian value of the word. */
    uint32_t u = ((uint32_t)be[0] << 24) | ((uint32_t)be[1] << 16) |
                 ((uint32_t)be[2] << 8) | (uint32_t)be[3];
    int32_t w;
    memcpy(&w, &u, 4);
    /* Kirchhoff.c:1287-1289, evaluated left to right: int * float * float */
    int sgn = (((w >> 31) & 0x01) * (-2)) + 1;
    float frac = (float)(w & 0x00ffffff) / 0x1000000;
    float scale = (float)pow(16, ((w >> 24) & 0x7f) - 64);
    return sgn * frac * scale;
}

void orc_ibm_block(const uint8_t *bytes, int64_t nwords, float *out)
{
    for (int64_t i = 0; i < nwords; ++i)
        out[i] = orc_ibm_to_float(bytes + 4 * i);
}

/* ------------------------------------------------------------------ A2 -- */
int orc_alias_filter(const orc_params *p, int ns, float dt, float *anti_alias)
{
    /* Kirchhoff.c:387-393 */
    float DIP = 45.0f, V_max = 1500.0f;
    float tanA = (float)tan(DIP * ORC_PI / 180.0);
    float mm = (float)(p->iline_distance <= p->xline_distance ? p->iline_distance
                                                             : p->xline_distance);
    float fmax = (float)(V_max / 4.0 / mm / tanA);
    int trange_max = (int)floor(1.0 / 2.0 / fmax / dt);
    /* Kirchhoff.c:401-423; the array starts zeroed like alloc1d's calloc,
     * which is what the mirror loop reads at i == ns/2 for odd ns. */
    for (int i = 0; i < ns; ++i) anti_alias[i] = 0.0f;
    float fs = (float)(2.0 / dt / ns);
    fmax = (float)p->anti_alias_cutoff;
    for (int i = 0; i <= ns / 2; ++i) anti_alias[i] = 1.0f;
    fmax = (float)(fmax * 3.0);
    for (int i = 0; i <= ns / 2; ++i) {
        float f = i * fs;
        float q = f * f * f * f / fmax / fmax / fmax / fmax;
        anti_alias[i] = (float)(1.0 / sqrt(1.0 + q));
    }
    for (int i = ns / 2; i < ns; ++i) anti_alias[i] = anti_alias[ns - i];
    return trange_max;
}

/* ------------------------------------------------------------------ A6 -- */
static void t2_table(int ns, float dt, float *t2)
{
    for (int t = 0; t < ns; ++t) /* Kirchhoff.c:1246-1247 */
        t2[t] = (float)(0.25 * t * t * dt * dt);
}

void orc_trace_prep_fp32(const float *tr, int ns, float dt, int tm, float *P)
{
    float *t2 = (float *)malloc(sizeof(float) * (size_t)ns);
    t2_table(ns, dt, t2);
    /* integral 1, Kirchhoff.c:1314-1316: float + float/double -> double */
    P[tm] = tr[0];
    for (int t = 1; t < ns; ++t)
        P[t + tm] = (float)((double)P[t - 1 + tm] + tr[t] / (4.0 * t2[t]));
    /* tail, Kirchhoff.c:1319-1320: all float */
    for (int t = ns; t < ns + tm; ++t)
        P[t + tm] = P[t - 1 + tm] + tr[ns - 1] / dt / dt / t / t;
    /* integral 2, Kirchhoff.c:1323-1324 */
    for (int t = 1; t < ns + tm; ++t)
        P[t + tm] = P[t - 1 + tm] + P[t + tm];
    /* mirror pad, Kirchhoff.c:1327-1328 */
    for (int t = 0; t < tm; ++t)
        P[t] = P[2 * tm - t];
    free(t2);
}

void orc_trace_prep_fp64(const float *tr, int ns, float dt, int tm, double *P)
{
    const double dtd = dt;
    P[tm] = tr[0];
    for (int t = 1; t < ns; ++t) {
        double t2 = 0.25 * t * t * dtd * dtd;
        P[t + tm] = P[t - 1 + tm] + tr[t] / (4.0 * t2);
    }
    for (int t = ns; t < ns + tm; ++t)
        P[t + tm] = P[t - 1 + tm] + (double)tr[ns - 1] / dtd / dtd / t / t;
    for (int t = 1; t < ns + tm; ++t)
        P[t + tm] = P[t - 1 + tm] + P[t + tm];
    for (int t = 0; t < tm; ++t)
        P[t] = P[2 * tm - t];
}

/* ------------------------------------------------------------------ A5 -- */
int orc_trace_geometry(const orc_params *p, const orc_grid *g,
                       const int32_t c[4], float geo[6])
{
    const float cosA = g->cosA, sinA = g->sinA;
    float sxt, syt, gxt, gyt;
    int i, j;
    if (p->scale_xy == 0.0) {
        /* Kirchhoff.c:1297-1304: htonl() is uint32_t, minus a double origin,
         * truncated into an int */
        i = (int)((double)(uint32_t)c[0] - p->ox);
        j = (int)((double)(uint32_t)c[1] - p->oy);
        sxt = cosA * i + sinA * j;
        syt = -sinA * i + cosA * j;
        i = (int)((double)(uint32_t)c[2] - p->ox);
        j = (int)((double)(uint32_t)c[3] - p->oy);
        gxt = cosA * i + sinA * j;
        gyt = -sinA * i + cosA * j;
    } else {
        /* Kirchhoff_GHTest1.c:1302-1315 */
        const double s = p->scale_xy;
        i = (int)((double)(uint32_t)c[0] - p->ox / s);
        j = (int)((double)(uint32_t)c[1] - p->oy / s);
        sxt = (float)(cosA * i * s + sinA * j * s);
        syt = (float)(-sinA * i * s + cosA * j * s);
        i = (int)((double)(uint32_t)c[2] - p->ox / s);
        j = (int)((double)(uint32_t)c[3] - p->oy / s);
        gxt = (float)(cosA * i * s + sinA * j * s);
        gyt = (float)(-sinA * i * s + cosA * j * s);
    }
    /* Kirchhoff.c:1307-1310 */
    float h2 = (float)(0.25 * ((sxt - gxt) * (sxt - gxt) + (syt - gyt) * (syt - gyt)));
    float cmpX = (float)(0.5 * (sxt + gxt));
    float cmpY = (float)(0.5 * (syt + gyt));
    int off = (int)floor((2.0 * sqrt(h2) - p->oh) / p->dh);
    geo[0] = sxt; geo[1] = syt; geo[2] = gxt; geo[3] = gyt;
    geo[4] = cmpX; geo[5] = cmpY;
    return off;
}

/* ----------------------------------------------------------- A7-A10 -- */
typedef struct job_ctx {
    const orc_params *p;
    orc_grid g;
    int ns, nsm, tm, job, xslice, yslice;
    float dt, dtm;
    /* output addressing: column (xc,yc) of the job lands at
     * [(x0+xc)*NY + (y0+yc)][nh][ns] */
    int x0, y0, NY;
} job_ctx;

static float image_x(const job_ctx *c, double colx)
{
    const orc_params *p = c->p;
    double start = p->image_origin_zero ? 0.0 : (double)p->start_iline;
    return (float)((start + (colx * p->inc_iline)) * p->iline_distance);
}

static float image_y(const job_ctx *c, double coly)
{
    const orc_params *p = c->p;
    double start = p->image_origin_zero ? 0.0 : (double)p->start_xline;
    return (float)((start + (coly * p->inc_xline)) * p->xline_distance);
}

/* Kirchhoff.c:1357-1372; returns -1 where the reference would index the
 * velocity model below zero. */
static int vel_index(const orc_params *p, int col, int axis)
{
    int start = axis == 0 ? p->start_iline : p->start_xline;
    int ovel = axis == 0 ? p->ovel_ilines : p->ovel_xlines;
    int dvel = axis == 0 ? p->dvel_ilines : p->dvel_xlines;
    int nvel = axis == 0 ? p->nvel_ilines : p->nvel_xlines;
    double inc = axis == 0 ? p->inc_iline : p->inc_xline;
    int origin = p->image_origin_zero ? 0 : (start - ovel) / dvel;
    int idx = (int)round(origin + col * (inc / dvel));
    if (idx > nvel - 1) idx = nvel - 1;
    if (p->vel_lower_clamp && idx < 0) idx = 0;
    return idx < 0 ? -1 : idx;
}

/* Slab prefilter of Kirchhoff.c:1332-1338. */
static int slab_keep(const job_ctx *c, float cmpX, float cmpY)
{
    const orc_params *p = c->p;
    const int nx = c->g.nx, ny = c->g.ny;
    float IX = image_x(c, nx / 2.0 + c->xslice * nx);
    float IY = image_y(c, ny / 2.0 + c->yslice * ny);
    float dist = (float)sqrt((IX - cmpX) * (IX - cmpX) + (IY - cmpY) * (IY - cmpY));
    float add_on = (float)(0.5 * sqrt(nx * p->inc_iline * p->iline_distance * nx *
                                          p->inc_iline * p->iline_distance +
                                      ny * p->inc_xline * p->xline_distance * ny *
                                          p->inc_xline * p->xline_distance));
    return dist < p->aperture + add_on;
}

static int64_t migrate_job_fp32(const job_ctx *c, const float *traces,
                                const int32_t *coords, int64_t ntr, const float *vel2,
                                float *km, float *fold)
{
    const orc_params *p = c->p;
    const int ns = c->ns, tm = c->tm, nsm = c->nsm, nh = p->nh;
    const int nx = c->g.nx, ny = c->g.ny;
    const float dt = c->dt, dt_ratio = c->dt / c->dtm;
    float *t2 = (float *)malloc(sizeof(float) * (size_t)ns);
    float *P = (float *)malloc(sizeof(float) * (size_t)(ns + 2 * tm));
    int *t1tab = (int *)malloc(sizeof(int) * (size_t)ns);
    int64_t ncontrib = 0;
    t2_table(ns, dt, t2);
    for (int t = 0; t < ns; ++t) { /* Kirchhoff.c:1376-1378 */
        int t1 = (int)round(t * dt_ratio);
        t1tab[t] = t1 > nsm - 1 ? nsm - 1 : t1;
    }
    for (int64_t it = 0; it < ntr; ++it) {
        float geo[6];
        const int off = orc_trace_geometry(p, &c->g, coords + 4 * it, geo);
        const float sxt = geo[0], syt = geo[1], gxt = geo[2], gyt = geo[3];
        const float cmpX = geo[4], cmpY = geo[5];
        orc_trace_prep_fp32(traces + (size_t)it * ns, ns, dt, tm, P);
        if (!slab_keep(c, cmpX, cmpY)) continue;
        for (int xc = 0; xc < nx; ++xc) {
            for (int yc = 0; yc < ny; ++yc) {
                const int colx = xc + c->xslice * nx, coly = yc + c->yslice * ny;
                const float IX = image_x(c, colx), IY = image_y(c, coly);
                const float dIS2 = (IX - sxt) * (IX - sxt) + (IY - syt) * (IY - syt);
                const float dIR2 = (IX - gxt) * (IX - gxt) + (IY - gyt) * (IY - gyt);
                const float dist =
                    (float)sqrt((IX - cmpX) * (IX - cmpX) + (IY - cmpY) * (IY - cmpY));
                if (!(dist < p->aperture)) continue;
                const int xc1 = vel_index(p, colx, 0), yc1 = vel_index(p, coly, 1);
                if (xc1 < 0 || yc1 < 0) { ncontrib = -1; goto done; }
                const float *v2 = vel2 + ((size_t)xc1 * p->nvel_xlines + yc1) * nsm;
                ncontrib += ns;
                if (off >= nh || off < 0) continue; /* Kirchhoff.c:1387, trace-uniform */
                const size_t base =
                    (((size_t)(c->x0 + xc) * c->NY + (c->y0 + yc)) * nh + off) * ns;
                for (int t = 0; t < ns; ++t) { /* Kirchhoff.c:1374-1399 */
                    const float v = v2[t1tab[t]];
                    const float sqrt1 = t2[t] + dIS2 / v;
                    const float sqrt2 = t2[t] + dIR2 / v;
                    const float dsr = (float)(sqrt(sqrt1) + sqrt(sqrt2));
                    const int T = (int)floor(dsr / dt);
                    if (T < ns) {
                        const float F = (float)(-2.0 * P[T + tm] + P[T + tm - 2] + P[T + tm + 2]);
                        km[base + t] = km[base + t] + F;
                        fold[base + t] = (float)(fold[base + t] + 1.0);
                    }
                }
            }
        }
    }
done:
    free(t2);
    free(P);
    free(t1tab);
    return ncontrib;
}

static int64_t migrate_job_fp64(const job_ctx *c, const float *traces,
                                const int32_t *coords, int64_t ntr, const float *vel,
                                int full, double *km, double *fold)
{
    const orc_params *p = c->p;
    const int ns = c->ns, tm = c->tm, nsm = c->nsm, nh = p->nh;
    const int nx = c->g.nx, ny = c->g.ny;
    const float dt = c->dt, dt_ratio = c->dt / c->dtm;
    float *t2 = (float *)malloc(sizeof(float) * (size_t)ns);
    double *P = (double *)malloc(sizeof(double) * (size_t)(ns + 2 * tm));
    int *t1tab = (int *)malloc(sizeof(int) * (size_t)ns);
    int64_t ncontrib = 0;
    t2_table(ns, dt, t2);
    for (int t = 0; t < ns; ++t) {
        int t1 = (int)round(t * dt_ratio);
        t1tab[t] = t1 > nsm - 1 ? nsm - 1 : t1;
    }
    for (int64_t it = 0; it < ntr; ++it) {
        float geo[6];
        const int off = orc_trace_geometry(p, &c->g, coords + 4 * it, geo);
        orc_trace_prep_fp64(traces + (size_t)it * ns, ns, dt, tm, P);
        if (!slab_keep(c, geo[4], geo[5])) continue;
        for (int xc = 0; xc < nx; ++xc) {
            for (int yc = 0; yc < ny; ++yc) {
                const int colx = xc + c->xslice * nx, coly = yc + c->yslice * ny;
                const float IX = image_x(c, colx), IY = image_y(c, coly);
                const float dIS2 = (IX - geo[0]) * (IX - geo[0]) + (IY - geo[1]) * (IY - geo[1]);
                const float dIR2 = (IX - geo[2]) * (IX - geo[2]) + (IY - geo[3]) * (IY - geo[3]);
                const float dist = (float)sqrt((IX - geo[4]) * (IX - geo[4]) +
                                               (IY - geo[5]) * (IY - geo[5]));
                if (!(dist < p->aperture)) continue;
                const int xc1 = vel_index(p, colx, 0), yc1 = vel_index(p, coly, 1);
                if (xc1 < 0 || yc1 < 0) { ncontrib = -1; goto done; }
                const float *vv = vel + ((size_t)xc1 * p->nvel_xlines + yc1) * nsm;
                ncontrib += ns;
                if (off >= nh || off < 0) continue;
                const size_t base =
                    (((size_t)(c->x0 + xc) * c->NY + (c->y0 + yc)) * nh + off) * ns;
                /* widened geometry for the full-FP64 mode */
                const double dIS2d = ((double)IX - geo[0]) * ((double)IX - geo[0]) +
                                     ((double)IY - geo[1]) * ((double)IY - geo[1]);
                const double dIR2d = ((double)IX - geo[2]) * ((double)IX - geo[2]) +
                                     ((double)IY - geo[3]) * ((double)IY - geo[3]);
                for (int t = 0; t < ns; ++t) {
                    const float vf = vv[t1tab[t]];
                    int T;
                    if (full) {
                        const double v2 = (double)vf * (double)vf;
                        const double t2d = 0.25 * t * t * (double)dt * (double)dt;
                        const double dsr = sqrt(t2d + dIS2d / v2) + sqrt(t2d + dIR2d / v2);
                        T = (int)floor(dsr / (double)dt);
                    } else {
                        const float v2 = vf * vf;
                        const float sqrt1 = t2[t] + dIS2 / v2;
                        const float sqrt2 = t2[t] + dIR2 / v2;
                        const float dsr = (float)(sqrt(sqrt1) + sqrt(sqrt2));
                        T = (int)floor(dsr / dt);
                    }
                    if (T < ns) {
                        km[base + t] += -2.0 * P[T + tm] + P[T + tm - 2] + P[T + tm + 2];
                        fold[base + t] += 1.0;
                    }
                }
            }
        }
    }
done:
    free(t2);
    free(P);
    free(t1tab);
    return ncontrib;
}

static int64_t run_job(const orc_params *p, int ns, float dt, int nsm, float dtm, int job,
                       const float *traces, const int32_t *coords, int64_t ntr,
                       const float *vel, int mode, void *km, void *fold, int whole)
{
    job_ctx c;
    memset(&c, 0, sizeof c);
    c.p = p;
    orc_grid_init(p, &c.g);
    if (job < 0 || job >= c.g.no_job || ns <= 0 || nsm <= 0) return -1;
    float *aa = (float *)malloc(sizeof(float) * (size_t)ns);
    c.tm = orc_alias_filter(p, ns, dt, aa);
    free(aa);
    if (c.tm < 2) return -1; /* the trange=2 taps would leave dataP */
    c.ns = ns; c.dt = dt; c.nsm = nsm; c.dtm = dtm; c.job = job;
    c.xslice = job / p->jobs_y; /* Kirchhoff.c:287-295 */
    c.yslice = job % p->jobs_y;
    if (whole) {
        c.x0 = c.xslice * c.g.nx;
        c.y0 = c.yslice * c.g.ny;
        c.NY = c.g.ny * p->jobs_y;
    } else {
        c.x0 = 0; c.y0 = 0; c.NY = c.g.ny;
    }
    /* accumulators start at km=0, fold=1 (Kirchhoff.c:1262-1274) */
    const size_t slab = (size_t)p->nh * ns;
    for (int xc = 0; xc < c.g.nx; ++xc)
        for (int yc = 0; yc < c.g.ny; ++yc) {
            size_t b = ((size_t)(c.x0 + xc) * c.NY + (c.y0 + yc)) * slab;
            for (size_t k = 0; k < slab; ++k) {
                if (mode == ORC_FP32) {
                    ((float *)km)[b + k] = 0.0f;
                    ((float *)fold)[b + k] = 1.0f;
                } else {
                    ((double *)km)[b + k] = 0.0;
                    ((double *)fold)[b + k] = 1.0;
                }
            }
        }
    if (mode == ORC_FP32) {
        const size_t nv = (size_t)p->nvel_ilines * p->nvel_xlines * nsm;
        float *vel2 = (float *)malloc(sizeof(float) * nv);
        for (size_t k = 0; k < nv; ++k) vel2[k] = vel[k] * vel[k]; /* Kirchhoff.c:617 */
        int64_t n = migrate_job_fp32(&c, traces, coords, ntr, vel2, (float *)km, (float *)fold);
        free(vel2);
        return n;
    }
    return migrate_job_fp64(&c, traces, coords, ntr, vel, mode == ORC_FP64_FULL,
                            (double *)km, (double *)fold);
}

int64_t orc_migrate_job(const orc_params *p, int ns, float dt, int nsm, float dtm, int job,
                        const float *traces, const int32_t *coords, int64_t ntr,
                        const float *vel, int mode, void *km, void *fold)
{
    return run_job(p, ns, dt, nsm, dtm, job, traces, coords, ntr, vel, mode, km, fold, 0);
}

int64_t orc_migrate_all(const orc_params *p, int ns, float dt, int nsm, float dtm,
                        const float *traces, const int32_t *coords, int64_t ntr,
                        const float *vel, int mode, void *km, void *fold, int nthreads)
{
    orc_grid g;
    orc_grid_init(p, &g);
    int64_t total = 0;
    int bad = 0;
    if (nthreads < 1) nthreads = 1;
#pragma omp parallel for schedule(dynamic, 1) num_threads(nthreads) reduction(+ : total, bad)
    for (int job = 0; job < g.no_job; ++job) {
        int64_t n = run_job(p, ns, dt, nsm, dtm, job, traces, coords, ntr, vel, mode, km,
                            fold, 1);
        if (n < 0) bad += 1; else total += n;
    }
    return bad ? -1 : total;
}


/* ----------------------------------------------------------------- A11 -- */
/* The post-stack filter of Kirchhoff.c:1423-1479 runs FFTW's unnormalised
 * double-complex DFT (FFTW_FORWARD then FFTW_BACKWARD, Kirchhoff.c:1432-1433).
 * FFTW is not available here, so the DFT is evaluated from its definition:
 *   X[k] = sum_t x[t] exp(-2 pi i k t / n),  y[t] = sum_k Y[k] exp(+2 pi i k t / n).
 * Twiddles use the true pi (FFTW's accurate trigonometry), reduced by k*t mod n. */
typedef struct dft_tab { int n; double *c, *s; } dft_tab;

static void dft_init(dft_tab *d, int n)
{
    const double two_pi = 2.0 * acos(-1.0);
    d->n = n;
    d->c = (double *)malloc(sizeof(double) * (size_t)n);
    d->s = (double *)malloc(sizeof(double) * (size_t)n);
    for (int m = 0; m < n; ++m) {
        d->c[m] = cos(two_pi * m / n);
        d->s[m] = sin(two_pi * m / n);
    }
}

static void dft_free(dft_tab *d) { free(d->c); free(d->s); }

/* forward DFT of a real sequence x[n] -> (re, im) */
static void dft_fwd_real(const dft_tab *d, const double *x, double *re, double *im)
{
    const int n = d->n;
    for (int k = 0; k < n; ++k) {
        double sr = 0.0, si = 0.0;
        int m = 0;
        for (int t = 0; t < n; ++t) {
            sr += x[t] * d->c[m];
            si -= x[t] * d->s[m];
            m += k;
            if (m >= n) m -= n;
        }
        re[k] = sr;
        im[k] = si;
    }
}

/* real part of the backward DFT of (re, im) */
static void dft_bwd_re(const dft_tab *d, const double *re, const double *im, double *y)
{
    const int n = d->n;
    for (int t = 0; t < n; ++t) {
        double s = 0.0;
        int m = 0;
        for (int k = 0; k < n; ++k) {
            s += re[k] * d->c[m] - im[k] * d->s[m];
            m += t;
            if (m >= n) m -= n;
        }
        y[t] = s;
    }
}

void orc_poststack_fp32(int ns, float dt, const float *anti_alias, const float *km,
                        const float *fold, int64_t ntraces, float *out)
{
    dft_tab d;
    dft_init(&d, ns);
    double *x = (double *)malloc(sizeof(double) * (size_t)ns);
    double *re = (double *)malloc(sizeof(double) * (size_t)ns);
    double *im = (double *)malloc(sizeof(double) * (size_t)ns);
    double *y = (double *)malloc(sizeof(double) * (size_t)ns);
    for (int64_t tr = 0; tr < ntraces; ++tr) {
        const float *k = km + (size_t)tr * ns;
        const float *f = fold + (size_t)tr * ns;
        /* Kirchhoff.c:1457-1462: float expression stored into a double */
        x[0] = 0.0;
        for (int t = 1; t < ns; ++t)
            x[t] = -dt * t * (k[t] - k[t - 1]) * dt / f[t];
        dft_fwd_real(&d, x, re, im);
        /* Kirchhoff.c:1467-1473: a, b, c, d are floats */
        for (int t = 0; t < ns; ++t) {
            float a = anti_alias[t], b = 0.0f, c = (float)re[t], dd = (float)im[t];
            re[t] = a * c - b * dd;
            im[t] = a * dd + b * c;
        }
        dft_bwd_re(&d, re, im, y);
        /* Kirchhoff.c:1475-1479 */
        float *o = out + (size_t)tr * ns;
        o[0] = 0.0f;
        for (int t = 1; t < ns; ++t) o[t] = (float)(y[t] / ns);
    }
    free(x); free(re); free(im); free(y);
    dft_free(&d);
}

void orc_poststack_fp64(int ns, float dt, const float *anti_alias, const double *km,
                        const double *fold, int64_t ntraces, double *out)
{
    dft_tab d;
    dft_init(&d, ns);
    double *x = (double *)malloc(sizeof(double) * (size_t)ns);
    double *re = (double *)malloc(sizeof(double) * (size_t)ns);
    double *im = (double *)malloc(sizeof(double) * (size_t)ns);
    double *y = (double *)malloc(sizeof(double) * (size_t)ns);
    const double dtd = dt;
    for (int64_t tr = 0; tr < ntraces; ++tr) {
        const double *k = km + (size_t)tr * ns;
        const double *f = fold + (size_t)tr * ns;
        x[0] = 0.0;
        for (int t = 1; t < ns; ++t)
            x[t] = -dtd * t * (k[t] - k[t - 1]) * dtd / f[t];
        dft_fwd_real(&d, x, re, im);
        for (int t = 0; t < ns; ++t) {
            re[t] *= (double)anti_alias[t];
            im[t] *= (double)anti_alias[t];
        }
        dft_bwd_re(&d, re, im, y);
        double *o = out + (size_t)tr * ns;
        o[0] = 0.0;
        for (int t = 1; t < ns; ++t) o[t] = y[t] / ns;
    }
    free(x); free(re); free(im); free(y);
    dft_free(&d);
}

/* ------------------------------------------------------------ A12-A14 -- */
static void put_be32(uint8_t *o, int32_t v)
{
    uint32_t u;
    memcpy(&u, &v, 4);
    o[0] = (uint8_t)(u >> 24); o[1] = (uint8_t)(u >> 16);
    o[2] = (uint8_t)(u >> 8);  o[3] = (uint8_t)u;
}

static void put_be16(uint8_t *o, uint16_t v)
{
    o[0] = (uint8_t)(v >> 8); o[1] = (uint8_t)v;
}

/* The 240 bytes of one output trace header as the reference emits them: the
 * fwrite sequence of Kirchhoff.c:1499-1590 (identical in EOM.c:1437-1497),
 * 4- and 2-byte big-endian items in order.  Only the items the reference
 * sets to something other than zero are named; every other item is the
 * zeroed struct field (Kirchhoff.c:1140-1225).  Item indices: */
enum {
    H_TRACL = 0, H_CDP = 5, H_OFFSET = 11, H_SX = 21, H_SY, H_GX, H_GY,
    H_NS = 38, H_DT = 39, H_CDPX = 71, H_CDPY, H_ILINE, H_CLINE,
    H_MESS0 = 78, H_MESS1 = 85, H_MESS2 = 86, H_MESS3 = 89, H_END = 90, H_COUNT = 91
};
/* width of each item in write order: tracl..cdpt (7 x 4), trid..duse (4 x 2),
 * offset..gwdep (8 x 4), scalel, scalco (2 x 2), sx..gy (4 x 4), counit..mute
 * (13 x 2), ns, dt (2 x 2), gain..otrav (31 x 2), cdpx..shotpt (5 x 4), then
 * messs, messs, mess, 6 x messs, mess, mess, messs, messs, mess, 9999 */
static int header_item_width(int k)
{
    if (k < 7) return 4;
    if (k < 11) return 2;
    if (k < 19) return 4;
    if (k < 21) return 2;
    if (k < 25) return 4;
    if (k < 71) return 2;
    if (k < 76) return 4;
    static const int tail[H_COUNT - 76] = {2, 2, 4, 2, 2, 2, 2, 2, 2, 4, 4, 2, 2, 4, 4};
    return tail[k - 76];
}

static void emit_header(const int64_t item[H_COUNT], uint8_t out[240])
{
    int pos = 0;
    for (int k = 0; k < H_COUNT; ++k) {
        const int w = header_item_width(k);
        if (w == 4) put_be32(out + pos, (int32_t)(uint32_t)item[k]);
        else put_be16(out + pos, (uint16_t)item[k]);
        pos += w;
    }
}

void orc_segy_trace_header(const orc_params *p, const orc_grid *g, int ns, float dt,
                           int job, int xc, int yc, int ih, int tracl, int32_t mess,
                           uint8_t out[240])
{
    const int xslice = job / p->jobs_y, yslice = job % p->jobs_y;
    const int X = xc + xslice * g->nx, Y = yc + yslice * g->ny;
    int64_t item[H_COUNT] = {0};
    /* Kirchhoff.c:1444 */
    int cdp = X + g->Nx * Y + 1;
    /* Kirchhoff.c:1482-1486 */
    int iline = (int)(p->start_iline + X * p->inc_iline);
    int cline = (int)(p->start_xline + Y * p->inc_xline);
    int offset = (int)ceil(p->oh + ih * p->dh);
    int xc1 = (int)((p->start_iline + X * p->inc_iline) * p->iline_distance);
    int yc1 = (int)((p->start_xline + Y * p->inc_xline) * p->xline_distance);
    /* Kirchhoff.c:1492-1493: int + float - float, truncated to int */
    int cdpx = (int)((int)p->ox + g->cosA * xc1 - g->sinA * yc1);
    int cdpy = (int)((int)p->oy + g->sinA * xc1 + g->cosA * yc1);
    item[H_TRACL] = tracl;
    item[H_CDP] = cdp;
    item[H_OFFSET] = offset;
    item[H_SX] = item[H_SY] = item[H_GX] = item[H_GY] = 1; /* htonl(1), Kirchhoff.c:1165-1168 */
    item[H_NS] = (uint16_t)ns;
    item[H_DT] = (uint16_t)(dt * 1000000);
    item[H_CDPX] = cdpx;
    item[H_CDPY] = cdpy;
    item[H_ILINE] = iline;
    item[H_CLINE] = cline;
    item[H_MESS0] = item[H_MESS1] = item[H_MESS2] = item[H_MESS3] = mess;
    item[H_END] = 9999; /* Kirchhoff.c:1589-1590 */
    emit_header(item, out);
}

void orc_segy_job_headers(const orc_params *p, const orc_grid *g, int ns, float dt, int job,
                          uint8_t *out)
{
    /* Kirchhoff.c:1437-1598: it from nx*ny*nh*job + 1, one record per
     * (xc, yc, ih); mess = 0 at 1224 and 9999 after the first record */
    int it = g->nx * g->ny * p->nh * job + 1;
    int32_t mess = 0;
    for (int xc = 0; xc < g->nx; ++xc)
        for (int yc = 0; yc < g->ny; ++yc)
            for (int ih = 0; ih < p->nh; ++ih) {
                orc_segy_trace_header(p, g, ns, dt, job, xc, yc, ih, it, mess, out);
                mess = 9999;
                out += 240;
                ++it;
            }
}

void orc_segy_file_header(const orc_params *p, const orc_grid *g, int ns, float dt,
                          int fix_count, uint8_t out[3600])
{
    static const char msg[] = "written by G. O'Brien\n"; /* Kirchhoff.c:757-758 */
    memset(out, 0, 3600);
    memcpy(out, msg, sizeof msg - 1);
    uint8_t *b = out + 3200;
    /* Kirchhoff.c:781-783 writes an uninitialised short here; the count it
     * meant to write is written on request, zero otherwise */
    if (fix_count) put_be16(b + 12, (uint16_t)(g->Nx * g->Ny * p->nh));
    put_be16(b + 16, (uint16_t)(dt * 1000000));
    put_be16(b + 20, (uint16_t)ns);
    put_be16(b + 22, (uint16_t)ns);
    put_be16(b + 24, 5);
    put_be16(b + 26, 1);
    put_be16(b + 28, 1);
    put_be16(b + 30, 1);
}

int orc_count_size(const uint8_t *head, int64_t length, int *ns, float *dt, int *nt)
{
    /* Kirchhoff.c:650-659: trace-1 header bytes 115-118 */
    if (length < 3600 + 240) return -1;
    uint16_t ns1 = (uint16_t)((head[3714] << 8) | head[3715]);
    uint16_t dt1 = (uint16_t)((head[3716] << 8) | head[3717]);
    *ns = (int)ns1;
    *dt = (float)(dt1 / 1000000.0);
    *nt = (int)((length - 3600) / (240 + ns1 * 4));
    return 0;
}

/* ------------------------------------------------------------------------ */
/* Differentiable offset-class migration and its velocity gradient           */
/* (an extension of the reference operator; no counterpart in Kirchhoff.c,   */
/* whose nearest-neighbour velocity and floor(T) have zero derivative almost */
/* everywhere).  FP64 throughout.                                            */
/* ------------------------------------------------------------------------ */

/* centred cubic B-spline, support (-2, 2) */
static double bspline3(double z)
{
    const double a = fabs(z);
    if (a >= 2.0) return 0.0;
    if (a >= 1.0) return (2.0 - a) * (2.0 - a) * (2.0 - a) / 6.0;
    return 2.0 / 3.0 - a * a + 0.5 * a * a * a;
}

void orc_vgrad_velocity(int ns, int nxp, int nt_nodes, int nx_nodes, const double *V,
                        double *v)
{
    const double dtn = (double)(ns - 1) / (nt_nodes - 3);
    const double dxn = nxp > 1 ? (double)(nxp - 1) / (nx_nodes - 3) : 1.0;
    for (int X = 0; X < nxp; ++X)
        for (int t = 0; t < ns; ++t) {
            double s = 0.0;
            for (int i = 0; i < nt_nodes; ++i) {
                const double bt = bspline3(t / dtn + 1.0 - i);
                if (bt == 0.0) continue;
                for (int j = 0; j < nx_nodes; ++j) {
                    const double bx = bspline3(X / dxn + 1.0 - j);
                    if (bx != 0.0) s += V[(int64_t)i * nx_nodes + j] * bt * bx;
                }
            }
            v[(int64_t)X * ns + t] = s;
        }
}

/* tap k of the filtered trace, zero outside [0, ns) */
static double ftap(const double *F, int ns, int64_t k)
{
    return (k >= 0 && k < ns) ? F[k] : 0.0;
}

double orc_vgrad(int ns, double dt, double aperture, int nh, int nxp, const float *colx,
                 float coly, const double *F, const float *geo, const int32_t *off,
                 int64_t ntr, int nt_nodes, int nx_nodes, const double *V, double *cig,
                 double *dJdV)
{
    double *v = (double *)malloc(sizeof(double) * (size_t)nxp * ns);
    double *I = (double *)calloc((size_t)nh * ns, sizeof(double));
    double *S = (double *)calloc((size_t)nh * ns, sizeof(double));   /* sum of dp/dv */
    double *gv = (double *)calloc((size_t)nxp * ns, sizeof(double)); /* dJ/dv(X,t) */
    orc_vgrad_velocity(ns, nxp, nt_nodes, nx_nodes, V, v);
    double J = 0.0;
    for (int X = 0; X < nxp; ++X) {
        memset(I, 0, sizeof(double) * (size_t)nh * ns);
        memset(S, 0, sizeof(double) * (size_t)nh * ns);
        const float IX = colx[X], IY = coly;
        for (int64_t i = 0; i < ntr; ++i) {
            const int h = off[i];
            if (h < 0 || h >= nh) continue;
            const float *g = geo + 6 * i;
            /* aperture gate of Kirchhoff.c:1353-1354 (float distance) */
            const float cx = IX - g[4], cy = IY - g[5];
            const float dist = (float)sqrt((double)(cx * cx + cy * cy));
            if (!(dist < aperture)) continue;
            const double a = ((double)IX - g[0]) * ((double)IX - g[0]) +
                             ((double)IY - g[1]) * ((double)IY - g[1]);
            const double b = ((double)IX - g[2]) * ((double)IX - g[2]) +
                             ((double)IY - g[3]) * ((double)IY - g[3]);
            const double *Fi = F + i * (int64_t)ns;
            for (int t = 0; t < ns; ++t) {
                const double vv = v[(int64_t)X * ns + t];
                const double th = 0.5 * t * dt, t2 = th * th;
                const double s1 = sqrt(t2 + a / (vv * vv)), s2 = sqrt(t2 + b / (vv * vv));
                double q = (s1 + s2) / dt;
                if (q > ns + 2.0) q = ns + 2.0;
                const double Tf = floor(q), u = q - Tf;
                const int64_t T = (int64_t)Tf;
                const double f0 = ftap(Fi, ns, T - 1), f1 = ftap(Fi, ns, T);
                const double f2 = ftap(Fi, ns, T + 1), f3 = ftap(Fi, ns, T + 2);
                /* Catmull-Rom cubic through f0..f3 at u in [0,1) */
                const double c1 = 0.5 * (f2 - f0);
                const double c2 = f0 - 2.5 * f1 + 2.0 * f2 - 0.5 * f3;
                const double c3 = 1.5 * (f1 - f2) + 0.5 * (f3 - f0);
                I[(int64_t)h * ns + t] += f1 + u * (c1 + u * (c2 + u * c3));
                if (dJdV) {
                    const double dp = c1 + u * (2.0 * c2 + 3.0 * u * c3);
                    /* dq/dv = -(a/s1 + b/s2) / (v^3 dt); s = 0 only with a = 0 */
                    const double ds = (s1 > 0.0 ? a / s1 : 0.0) + (s2 > 0.0 ? b / s2 : 0.0);
                    S[(int64_t)h * ns + t] += dp * (-ds / (vv * vv * vv * dt));
                }
            }
        }
        for (int t = 0; t < ns; ++t) {
            double gsum = 0.0;
            for (int h = 0; h < nh; ++h) {
                const double Ih = I[(int64_t)h * ns + t];
                const double dn = h + 1 < nh ? I[(int64_t)(h + 1) * ns + t] - Ih : 0.0;
                const double dp = h > 0 ? Ih - I[(int64_t)(h - 1) * ns + t] : 0.0;
                if (h + 1 < nh) J += dn * dn;
                gsum += 2.0 * (dp - dn) * S[(int64_t)h * ns + t];
                if (cig) cig[((int64_t)X * nh + h) * ns + t] = Ih;
            }
            gv[(int64_t)X * ns + t] = gsum;
        }
    }
    if (dJdV) {
        const double dtn = (double)(ns - 1) / (nt_nodes - 3);
        const double dxn = nxp > 1 ? (double)(nxp - 1) / (nx_nodes - 3) : 1.0;
        for (int i = 0; i < nt_nodes; ++i)
            for (int j = 0; j < nx_nodes; ++j) {
                double s = 0.0;
                for (int X = 0; X < nxp; ++X) {
                    const double bx = bspline3(X / dxn + 1.0 - j);
                    if (bx == 0.0) continue;
                    for (int t = 0; t < ns; ++t) {
                        const double bt = bspline3(t / dtn + 1.0 - i);
                        if (bt != 0.0) s += gv[(int64_t)X * ns + t] * bt * bx;
                    }
                }
                dJdV[(int64_t)i * nx_nodes + j] = s;
            }
    }
    free(v); free(I); free(S); free(gv);
    return J;
}

/* ------------------------------------------------------------------------ */
/* Equivalent-offset migration: common-scatter-point gathers (EOM.c csp())   */
/* ------------------------------------------------------------------------ */
#define ORC_EOM_PI 3.14159265359 /* EOM.c:124 */

void orc_eom_grid_init(const orc_eom_params *p, orc_eom_grid *g)
{
    /* EOM.c:279-320: Nx = ceil((excsp-oxcsp)/dcspX) (1 when dcspX == 0) */
    memset(g, 0, sizeof *g);
    g->Nx = p->dcsp_x == 0.0 ? 1 : (int)ceil((p->excsp - p->oxcsp) / p->dcsp_x);
    g->Ny = p->dcsp_y == 0.0 ? 1 : (int)ceil((p->eycsp - p->oycsp) / p->dcsp_y);
    if (g->Nx == 0) g->Nx = 1;
    if (g->Ny == 0) g->Ny = 1;
    g->sinA = (float)sin(p->angle_e_to_n * ORC_EOM_PI / 180.0);
    g->cosA = (float)cos(p->angle_e_to_n * ORC_EOM_PI / 180.0);
    g->start_x = (float)(g->cosA * (p->oxcsp - p->ox) + g->sinA * (p->oycsp - p->oy));
    g->start_y = (float)(-g->sinA * (p->oxcsp - p->ox) + g->cosA * (p->oycsp - p->oy));
    g->no_job = p->jobs_x * p->jobs_y;
    g->nx = (int)ceil(1.0 * g->Nx / p->jobs_x);
    g->ny = (int)ceil(1.0 * g->Ny / p->jobs_y);
    if (g->nx == 0) g->nx = 1;
    if (g->ny == 0) g->ny = 1;
}

void orc_eom_filters(int ns, float dt, double cutoff, float *lowpass, float *rootrho)
{
    /* EOM.c:759-845; calloc'd arrays mirrored from ns/2 upwards, so for odd
     * ns the bins ns/2 and ns/2+1 read a not-yet-set zero */
    const float fs = (float)(2.0 / dt / ns);
    const float fmax = (float)cutoff;
    float *a = (float *)calloc((size_t)ns, sizeof(float));
    float *r = (float *)calloc((size_t)ns, sizeof(float));
    for (int i = 0; i <= ns / 2; ++i) {
        const float f = i * fs;
        a[i] = (float)(1.0 / sqrt(1.0 + f * f * f * f / fmax / fmax / fmax / fmax));
        r[i] = (float)sqrt(i * fs / 2.0);
    }
    for (int i = ns / 2; i < ns; ++i) {
        a[i] = a[ns - i];
        r[i] = r[ns - i];
    }
    if (lowpass) memcpy(lowpass, a, sizeof(float) * (size_t)ns);
    if (rootrho) memcpy(rootrho, r, sizeof(float) * (size_t)ns);
    free(a);
    free(r);
}

int64_t orc_eom_job(const orc_eom_params *p, int ns, float dt, int job, const float *traces,
                    const int32_t *coords, int64_t ntr, const float *vel, float *csp, float *fold)
{
    orc_eom_grid g;
    orc_eom_grid_init(p, &g);
    if (job < 0 || job >= g.no_job) return -1;
    const int nx = g.nx, ny = g.ny, nh = p->nh;
    const int xs = job / p->jobs_y, ys = job % p->jobs_y;
    float *v2 = (float *)malloc(sizeof(float) * (size_t)ns);
    float *t2d = (float *)malloc(sizeof(float) * (size_t)ns);
    for (int t = 0; t < ns; ++t) {
        v2[t] = vel[t] * vel[t];                                  /* EOM.c:877 */
        t2d[t] = (float)(sqrt((double)(t * dt)) * t * dt);        /* EOM.c:1196 */
    }
    const float aperture2 = (float)(p->aperture * p->aperture);   /* EOM.c:1176 */
    const size_t n = (size_t)nx * ny * nh * ns;
    for (size_t k = 0; k < n; ++k) { csp[k] = 0.0f; fold[k] = 1.0f; }  /* EOM.c:1202-1211 */
    int64_t evaluated = 0;
    for (int64_t i = 0; i < ntr; ++i) {
        const uint32_t *w = (const uint32_t *)(coords + 4 * i);
        /* EOM.c:1245-1256: scaleXY * htonl(word) - ox, rotated */
        float io = (float)(p->scale_xy * w[0] - p->ox), jo = (float)(p->scale_xy * w[1] - p->oy);
        const float sxt = g.cosA * io + g.sinA * jo;
        const float syt = -g.sinA * io + g.cosA * jo;
        io = (float)(p->scale_xy * w[2] - p->ox);
        jo = (float)(p->scale_xy * w[3] - p->oy);
        const float gxt = g.cosA * io + g.sinA * jo;
        const float gyt = -g.sinA * io + g.cosA * jo;
        const float h2 = (float)(0.25 * ((sxt - gxt) * (sxt - gxt) + (syt - gyt) * (syt - gyt)));
        const float cmpX = (float)(0.5 * (sxt + gxt));
        const float cmpY = (float)(0.5 * (syt + gyt));
        const float *tr = traces + i * (int64_t)ns;
        for (int xc = 0; xc < nx; ++xc)
            for (int yc = 0; yc < ny; ++yc) {
                /* EOM.c:1261-1270 */
                const float cspX = (float)(g.start_x + (xc + xs * nx) * p->dcsp_x);
                const float cspY = (float)(g.start_y + (yc + ys * ny) * p->dcsp_y);
                const float x2 = (cmpX - cspX) * (cmpX - cspX) + (cmpY - cspY) * (cmpY - cspY);
                const float dist = x2 + h2;
                if (!(dist < aperture2)) continue;
                const int talpha = (int)round(2.0 * sqrt((double)(h2 / v2[0])) / dt);  /* :1272 */
                if (talpha < ns) evaluated += ns - talpha;
                float *cs = csp + ((size_t)xc * ny + yc) * nh * ns;
                float *fo = fold + ((size_t)xc * ny + yc) * nh * ns;
                for (int t = talpha; t < ns; ++t) {
                    /* EOM.c:1285-1306 */
                    const float crossterm2 =
                        (float)((4.0 * x2 * h2) / (double)((float)t * dt * (float)t * dt * v2[t]));
                    const float eq = (float)sqrt((double)(x2 + h2 - crossterm2));
                    const double qo = (double)(eq - (float)p->oh) / p->dh;
                    const double fl = floor(qo);
                    /* off < nh-1 && off >= 0; a NaN offset (undefined in the
                     * reference's int conversion) contributes nothing */
                    if (!(fl >= 0.0 && fl < (double)(nh - 1))) continue;
                    const int off = (int)fl;
                    const float tc = (float)(2.25 * eq / sqrt((double)v2[t]));
                    const float scale3 =
                        (float)(0.5 + atan(200.0 * (double)((float)t * dt - tc) / 0.080) / ORC_EOM_PI);
                    const float scale2 = (float)(qo - off);
                    const float scale1 = (float)(1.0 - scale2);
                    const float scaleT = scale3;
                    float *c0 = cs + (size_t)off * ns, *f0 = fo + (size_t)off * ns;
                    c0[t] = c0[t] + scale1 * scaleT * tr[t] / t2d[t];
                    f0[t] = (float)(f0[t] + 1.0);
                    c0[ns + t] = c0[ns + t] + scale2 * scaleT * tr[t] / t2d[t];
                    f0[ns + t] = (float)(f0[ns + t] + 1.0);
                }
            }
    }
    free(v2);
    free(t2d);
    return evaluated;
}

void orc_eom_post(int ns, float dt, const float *lowpass, const float *rootrho, const float *csp,
                  const float *fold, int64_t ntraces, float *out)
{
    /* EOM.c:1386-1413: x = (csp/fold) dt t, FFT, low-pass, (1+i) sqrt-rho,
     * inverse FFT, real part / ns; the spectrum passes through float a,b,c,d */
    dft_tab d;
    dft_init(&d, ns);
    double *x = (double *)malloc(sizeof(double) * (size_t)ns);
    double *re = (double *)malloc(sizeof(double) * (size_t)ns);
    double *im = (double *)malloc(sizeof(double) * (size_t)ns);
    double *y = (double *)malloc(sizeof(double) * (size_t)ns);
    for (int64_t j = 0; j < ntraces; ++j) {
        const float *c = csp + j * (int64_t)ns, *f = fold + j * (int64_t)ns;
        x[0] = 0.0;
        for (int t = 1; t < ns; ++t) x[t] = (c[t] / f[t]) * dt * t;
        dft_fwd_real(&d, x, re, im);
        for (int k = 0; k < ns; ++k) {
            float a = lowpass[k], b = 0.0f, cc = (float)re[k], dd = (float)im[k];
            re[k] = a * cc - b * dd;
            im[k] = a * dd + b * cc;
            a = rootrho[k];
            b = rootrho[k];
            cc = (float)re[k];
            dd = (float)im[k];
            re[k] = a * cc - b * dd;
            im[k] = a * dd + b * cc;
        }
        dft_bwd_re(&d, re, im, y);
        float *o = out + j * (int64_t)ns;
        o[0] = 0.0f;
        for (int t = 1; t < ns; ++t) o[t] = (float)(y[t] / ns);
    }
    free(x); free(re); free(im); free(y);
    dft_free(&d);
}

void orc_eom_trace_header(const orc_eom_params *p, int ns, float dt, int job, int xc, int yc,
                          int ih, int it, int32_t mess, uint8_t out[240])
{
    orc_eom_grid g;
    orc_eom_grid_init(p, &g);
    const int xs = job / p->jobs_y, ys = job % p->jobs_y;
    int64_t item[H_COUNT] = {0};
    /* EOM.c:1420-1435 */
    const int offset = (int)ceil(p->oh + ih * p->dh);
    const float cspX = (float)(g.start_x + (xc + xs * g.nx) * p->dcsp_x);
    const float cspY = (float)(g.start_y + (yc + ys * g.ny) * p->dcsp_y);
    const int xc1 = (int)(g.cosA * cspX - g.sinA * cspY);
    const int yc1 = (int)(g.sinA * cspX + g.cosA * cspY);
    const int csp = g.nx * g.ny * job + 1 + xc * g.ny + yc;   /* EOM.c:1346, 1536 */
    item[H_TRACL] = it;
    item[H_CDP] = csp;
    item[H_OFFSET] = offset;
    item[H_SX] = 13;   /* EOM.c:1120-1123 */
    item[H_SY] = 14;
    item[H_GX] = 15;
    item[H_GY] = 16;
    item[H_NS] = (uint16_t)ns;
    item[H_DT] = (uint16_t)(dt * 1000000);
    /* htonl(ox + xc1): a double converted to the uint32_t parameter (taken
     * through int64 so a negative sum wraps as on x86-64) */
    item[H_CDPX] = (uint32_t)(int64_t)(p->ox + xc1);
    item[H_CDPY] = (uint32_t)(int64_t)(p->oy + yc1);
    item[H_ILINE] = csp;
    item[H_CLINE] = 1;   /* htonl(1.0) */
    item[H_MESS0] = item[H_MESS1] = item[H_MESS2] = item[H_MESS3] = mess;
    item[H_END] = 9999;  /* EOM.c:1527-1528 */
    emit_header(item, out);
}

int64_t orc_eom_job_records(const orc_eom_params *p, int ns, float dt, int job,
                            const float *post, uint8_t *out)
{
    orc_eom_grid g;
    orc_eom_grid_init(p, &g);
    if (job < 0 || job >= g.no_job) return -1;
    const int64_t limit = (int64_t)g.Nx * g.Ny * p->nh + 1;   /* EOM.c:1417 */
    int64_t it = (int64_t)g.nx * g.ny * p->nh * job + 1;       /* EOM.c:1347 */
    int32_t mess = 0;                                          /* EOM.c:1179 */
    int64_t written = 0;
    const float *x = post;
    for (int xc = 0; xc < g.nx; ++xc)
        for (int yc = 0; yc < g.ny; ++yc)
            for (int ih = 0; ih < p->nh; ++ih, ++it, x += ns) {
                if (!(it < limit)) continue;
                orc_eom_trace_header(p, ns, dt, job, xc, yc, ih, (int)it, mess, out);
                mess = 9999;
                out += 240;
                /* data[0] = 0, data[t] = ntohf(out[t]) (EOM.c:1409-1412) */
                put_be32(out, 0);
                for (int t = 1; t < ns; ++t) {
                    int32_t w;
                    memcpy(&w, x + t, 4);
                    put_be32(out + 4 * t, w);
                }
                out += 4 * (int64_t)ns;
                ++written;
            }
    return written;
}

void orc_eom_file_header(int ns, float dt, uint8_t out[3600])
{
    /* EOM.c:479-480: the second sprintf overwrites the first; the rest of
     * the 3200-byte buffer is left as zero here */
    static const char msg[] = "written by G. O'Brien\n";
    memset(out, 0, 3600);
    memcpy(out, msg, sizeof msg - 1);
    uint8_t *b = out + 3200;
    put_be16(b + 16, (uint16_t)(dt * 1000000));   /* EOM.c:507-509 */
    put_be16(b + 20, (uint16_t)ns);
    put_be16(b + 22, (uint16_t)ns);
    put_be16(b + 24, 5);                          /* IEEE, EOM.c:518-520 */
    put_be16(b + 26, 1);
    put_be16(b + 28, 1);
    put_be16(b + 30, 1);
}
