/*
 * ktm_setup.c -- host setup of the migration: the part of Kirchhoff.c's
 * main() and kirchhoff() that depends only on the configuration, turned into
 * lookup tables for the kernels.
 *
 * Compiled with -ffp-contract=off: each table entry is produced with the
 * promotions and rounding points of the reference expression it replaces, so
 * that the kernels see bit-identical coordinates, weights and indices.
 */
#include "ktm_internal.h"

#include <math.h>
#include <stdlib.h>
#include <string.h>

/* Kirchhoff.c:97: the program's own value of pi, used for every angle. */
static const double REF_PI = 3.14159265359;

void ktm_config_default(ktm_config *c)
{
    /* the macro block of Kirchhoff.c:17-65 */
    memset(c, 0, sizeof *c);
    c->start_iline = 1500; c->end_iline = 1600;
    c->start_xline = 2500; c->end_xline = 2600;
    c->inc_iline = 1.0; c->inc_xline = 1.0;
    c->dh = 100.0; c->nh = 50; c->oh = 200;
    c->ox = 444479.29; c->oy = 478003.98; c->angle_e_to_n = -10.3690;
    c->iline_distance = 25.0; c->xline_distance = 12.5;
    c->aperture = 3000.0; c->anti_alias_cutoff = 50.0;
    c->nvel_ilines = 472; c->nvel_xlines = 643;
    c->dvel_ilines = 4; c->dvel_xlines = 8;
    c->ovel_ilines = 852; c->ovel_xlines = 752;
    c->ntblock = 50000; c->jobs_x = 36; c->jobs_y = 1;
    c->scale_xy = 0.0;
}

static int cells_along(int start, int end, double inc)
{
    /* Kirchhoff.c:266-277: ceil of the span, with the 0 -> 1 special cases */
    if (inc == 0.0) return 1;
    int n = (int)ceil((end - start) / inc);
    return n == 0 ? 1 : n;
}

/* create_aliasfilter(), Kirchhoff.c:380-441: returns trangeMax and writes the
 * spectral low-pass.  The mirror loop of 419-423 starts at ns/2, so for odd
 * ns its first two steps copy a never-written (zeroed) bin into bins ns/2 and
 * ns/2+1; the array is cleared first to reproduce that. */
static int alias_filter(const ktm_config *c, int ns, float dt, float *aa)
{
    const float tan45 = (float)tan(45.0f * REF_PI / 180.0);
    const float spacing = (float)fmin(c->iline_distance, c->xline_distance);
    const float f_alias = (float)(1500.0f / 4.0 / spacing / tan45);
    const int trange_max = (int)floor(1.0 / 2.0 / f_alias / dt);
    if (aa) {
        const float bin = (float)(2.0 / dt / ns);
        const float fc = (float)((float)c->anti_alias_cutoff * 3.0);
        const int half = ns / 2;
        memset(aa, 0, sizeof(float) * (size_t)ns);
        for (int i = 0; i <= half; ++i) {
            const float f = i * bin;
            const float r4 = f * f * f * f / fc / fc / fc / fc;
            aa[i] = (float)(1.0 / sqrt(1.0 + r4));
        }
        for (int i = half; i < ns; ++i) aa[i] = aa[ns - i];
    }
    return trange_max;
}

int ktm_setup_init(const ktm_config *c, int ns, float dt, int nsm, float dtm,
                   ktm_setup *s, float *anti_alias)
{
    if (!c || !s || ns <= 0 || nsm <= 0 || !(dt > 0.0f) || !(dtm > 0.0f)) return KTM_ERR_ARG;
    if (c->jobs_x <= 0 || c->jobs_y <= 0 || c->nh <= 0 || !(c->dh != 0.0)) return KTM_ERR_ARG;
    if (!(c->aperture >= 1e-3) || !(c->aperture < 1e30)) return KTM_ERR_ARG;
    memset(s, 0, sizeof *s);
    s->ns = ns; s->nsm = nsm; s->dt = dt; s->dtm = dtm;
    s->Nx = cells_along(c->start_iline, c->end_iline, c->inc_iline);
    s->Ny = cells_along(c->start_xline, c->end_xline, c->inc_xline);
    if (s->Nx < 0 || s->Ny < 0) return KTM_ERR_ARG;   /* end line before start line */
    s->no_job = c->jobs_x * c->jobs_y;
    s->nx = (int)ceil(1.0 * s->Nx / c->jobs_x);
    s->ny = (int)ceil(1.0 * s->Ny / c->jobs_y);
    if (s->nx == 0) s->nx = 1;
    if (s->ny == 0) s->ny = 1;
    s->nxp = s->nx * c->jobs_x;
    s->nyp = s->ny * c->jobs_y;
    s->nh = c->nh;
    s->sinA = (float)sin(c->angle_e_to_n * REF_PI / 180.0);
    s->cosA = (float)cos(c->angle_e_to_n * REF_PI / 180.0);
    s->trange_max = alias_filter(c, ns, dt, anti_alias);
    /* the fixed trange=2 taps (Kirchhoff.c:1391-1392) index dataP[T+tm+-2] */
    if (s->trange_max < 2) return KTM_ERR_UNSUPPORTED;
    return KTM_OK;
}

int ktm_vel_index(const ktm_config *c, int col, int axis)
{
    /* Kirchhoff.c:1357-1372; the origin term is an integer division */
    const int start = axis ? c->start_xline : c->start_iline;
    const int ovel = axis ? c->ovel_xlines : c->ovel_ilines;
    const int dvel = axis ? c->dvel_xlines : c->dvel_ilines;
    const int nvel = axis ? c->nvel_xlines : c->nvel_ilines;
    const double inc = axis ? c->inc_xline : c->inc_iline;
    if (dvel == 0 || nvel <= 0) return -1;
    const int origin = c->image_origin_zero ? 0 : (start - ovel) / dvel;
    int k = (int)round(origin + col * (inc / dvel));
    if (k > nvel - 1) k = nvel - 1;
    if (k < 0) k = c->vel_lower_clamp ? 0 : -1;
    return k;
}

float ktm_gate_threshold(double A)
{
    /* The reference gates are (double)(float)sqrt((double)x) < A for a float x
     * (Kirchhoff.c:1334-1338, 1352-1354).  The double rounding of sqrt to
     * float is innocuous (53 >= 2*24+2), so the left side is RN_f(sqrt(x)).
     * With B the largest float below A and M the midpoint of B and its
     * successor, RN_f(sqrt(x)) <= B  <=>  sqrt(x) < M  <=>  x < M^2: a tie
     * would need x == M^2, impossible since M has 25 significant bits and M^2
     * needs at least 49.  M^2 is exact in double; the float threshold is the
     * smallest float above it. */
    float B = (float)A;
    if ((double)B >= A) B = nextafterf(B, -INFINITY);
    const double M = 0.5 * ((double)B + (double)nextafterf(B, INFINITY));
    const double M2 = M * M;
    float thr = (float)M2;
    if ((double)thr <= M2) thr = nextafterf(thr, INFINITY);
    return thr;
}

/* image coordinate of column `col` (possibly fractional: the slab centre) */
static float image_coord(const ktm_config *c, double col, int axis)
{
    /* Kirchhoff.c:1332-1333,1346-1347 (origin dropped in the GHTest1 form) */
    const double start = c->image_origin_zero ? 0.0
                         : (double)(axis ? c->start_xline : c->start_iline);
    const double inc = axis ? c->inc_xline : c->inc_iline;
    const double d = axis ? c->xline_distance : c->iline_distance;
    return (float)((start + col * inc) * d);
}

void ktm_tables_free(ktm_tables *tb)
{
    free(tb->anti_alias); free(tb->t2); free(tb->t1); free(tb->colx); free(tb->coly);
    free(tb->velx); free(tb->vely); free(tb->slab_cx); free(tb->slab_cy);
    free(tb->slab_thr); free(tb->slab_x2thr); free(tb->w); free(tb->hconv);
    memset(tb, 0, sizeof *tb);
}

int ktm_tables_build(const ktm_config *c, int ns, float dt, int nsm, float dtm,
                     ktm_tables *tb)
{
    memset(tb, 0, sizeof *tb);
    int st = ktm_setup_init(c, ns, dt, nsm, dtm, &tb->s, NULL);
    if (st != KTM_OK) return st;
    const ktm_setup *s = &tb->s;
    tb->anti_alias = (float *)malloc(sizeof(float) * (size_t)ns);
    tb->t2 = (float *)malloc(sizeof(float) * (size_t)ns);
    tb->t1 = (int32_t *)malloc(sizeof(int32_t) * (size_t)ns);
    tb->colx = (float *)malloc(sizeof(float) * (size_t)s->nxp);
    tb->coly = (float *)malloc(sizeof(float) * (size_t)s->nyp);
    tb->velx = (int32_t *)malloc(sizeof(int32_t) * (size_t)s->nxp);
    tb->vely = (int32_t *)malloc(sizeof(int32_t) * (size_t)s->nyp);
    tb->slab_cx = (float *)malloc(sizeof(float) * (size_t)s->no_job);
    tb->slab_cy = (float *)malloc(sizeof(float) * (size_t)s->no_job);
    tb->slab_thr = (double *)malloc(sizeof(double) * (size_t)s->no_job);
    tb->slab_x2thr = (float *)malloc(sizeof(float) * (size_t)s->no_job);
    tb->w = (float *)malloc(sizeof(float) * (size_t)(ns + 2));
    tb->hconv = (double *)malloc(sizeof(double) * (size_t)ns);
    if (!tb->anti_alias || !tb->t2 || !tb->t1 || !tb->colx || !tb->coly || !tb->velx ||
        !tb->vely || !tb->slab_cx || !tb->slab_cy || !tb->slab_thr || !tb->slab_x2thr || !tb->w || !tb->hconv) {
        ktm_tables_free(tb);
        return KTM_ERR_NOMEM;
    }
    alias_filter(c, ns, dt, tb->anti_alias);

    /* traveltime tables, Kirchhoff.c:1246-1247 and 1376-1378 */
    const float dt_over_dtm = dt / dtm;
    for (int t = 0; t < ns; ++t) {
        tb->t2[t] = (float)(0.25 * t * t * dt * dt);
        int t1 = (int)round(t * dt_over_dtm);
        tb->t1[t] = t1 > nsm - 1 ? nsm - 1 : t1;
    }
    /* image columns */
    int bad = 0;
    for (int x = 0; x < s->nxp; ++x) {
        tb->colx[x] = image_coord(c, (double)x, 0);
        tb->velx[x] = ktm_vel_index(c, x, 0);
        bad |= tb->velx[x] < 0;
    }
    for (int y = 0; y < s->nyp; ++y) {
        tb->coly[y] = image_coord(c, (double)y, 1);
        tb->vely[y] = ktm_vel_index(c, y, 1);
        bad |= tb->vely[y] < 0;
    }
    /* slab prefilter, Kirchhoff.c:1332-1338 (job -> slice as in 287-295) */
    const float add_on = (float)(0.5 * sqrt(s->nx * c->inc_iline * c->iline_distance * s->nx *
                                                c->inc_iline * c->iline_distance +
                                            s->ny * c->inc_xline * c->xline_distance * s->ny *
                                                c->inc_xline * c->xline_distance));
    for (int j = 0; j < s->no_job; ++j) {
        const int xs = j / c->jobs_y, ys = j % c->jobs_y;
        tb->slab_cx[j] = image_coord(c, s->nx / 2.0 + xs * s->nx, 0);
        tb->slab_cy[j] = image_coord(c, s->ny / 2.0 + ys * s->ny, 1);
        tb->slab_thr[j] = c->aperture + add_on;
        tb->slab_x2thr[j] = ktm_gate_threshold(tb->slab_thr[j]);
    }
    tb->ap_x2thr = ktm_gate_threshold(c->aperture);
    /* CLOSED-mode weights: x[t] = tr[t] / (4 t2[t]) for 1 <= t < ns,
     * x[0] = tr[0] (Kirchhoff.c:1314), tail x[t] = tr[ns-1] / (dt^2 t^2)
     * for t >= ns (Kirchhoff.c:1319-1320). */
    tb->w[0] = 1.0f;
    for (int t = 1; t < ns; ++t) tb->w[t] = (float)(1.0 / (4.0 * (double)tb->t2[t]));
    for (int t = ns; t < ns + 2; ++t)
        tb->w[t] = (float)(1.0 / ((double)dt * dt * (double)t * t));
    /* post-stack kernel: h[m] = (1/ns) sum_k aa[k] cos(2 pi k m / ns); the
     * spectrum is even (aa[k] == aa[ns-k], odd-ns zero bins included), so the
     * filter of Kirchhoff.c:1465-1478 is this real circular convolution */
    const long double two_pi = 2.0L * acosl(-1.0L);
    for (int m = 0; m < ns; ++m) {
        long double acc = 0.0L;
        long long km = 0;
        for (int k = 0; k < ns; ++k) {
            acc += (long double)tb->anti_alias[k] * cosl(two_pi * km / ns);
            km += m;
            if (km >= ns) km -= ns;
        }
        tb->hconv[m] = (double)(acc / ns);
    }
    if (bad) {
        ktm_tables_free(tb);
        return KTM_ERR_VELOCITY;
    }
    return KTM_OK;
}
