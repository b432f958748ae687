/*
 * ktm_eom.hip -- equivalent-offset migration (EOM.c, csp(job_id)) on the
 * same engine: common-scatter-point gathers.
 *
 * For every trace and every CSP (common scatter point) of the job slab within
 * the aperture (EOM.c:1261-1270), each output sample t >= talpha is binned by
 * its equivalent offset
 *   he = sqrt(x2 + h2 - 4 x2 h2 / ((t dt)^2 v^2))                 EOM.c:1285-1286
 * into the two offset bins around he / dh with linear weights, a time taper
 * 0.5 + atan(200 (t dt - 2.25 he / v) / 0.08) / pi and 1 / (sqrt(t dt) t dt)
 * (EOM.c:1287-1306); fold counts the adds.  After all traces, each gather
 * trace is scaled by t dt / fold and filtered by the low-pass and the
 * (1+i) sqrt(rho) spectra (EOM.c:1386-1413).
 *
 * k_eom_geometry: per trace h2, CMP and talpha with the reference's float and
 *   double promotions (EOM.c:1245-1272).
 * k_eom_accum: one workgroup = one CSP x one segment of 256 samples, one lane
 *   per sample.  The lane owns its sample of every offset bin of its CSP for
 *   the whole launch, so it adds into the gather without atomics and in trace
 *   order (the reference's order of additions per bin); the two bins of the
 *   latest trace stay in registers until another trace needs their slot.
 *   Traces stream in chunks of 256: the aperture gate of each is evaluated
 *   lane-parallel, the passing ones are compacted in order through LDS.
 * k_postfft<PostEom> (ktm_postfft.h; k_postmx<PostEom> of ktm_postmx.h on
 *   the matrix cores beyond 4096 samples): the post filter as a real
 *   even-kernel circular convolution in FP64 (the (1+i) factor of
 *   the sqrt-rho spectrum leaves the real part of a Hermitian product
 *   unchanged), g = IDFT(lowpass * rootrho) / ns.
 * k_eom_records: the bytes of image<job>.su (EOM.c:1346-1537), one thread
 *   per 32-bit word.
 * The gathers are either the whole padded volume [nxp][nyp][nh][ns] or, for
 * the ktm_eom_job_* entry points, one job's slab [nx][ny][nh][ns] (the
 * reference's csp_data/fold of csp(), EOM.c:1190-1211).
 */
#include <hip/hip_runtime.h>

#include <math.h>
#include <stdint.h>
#include <stdlib.h>
#include <string.h>

#include "ktm_internal.h"
#include "ktm_devmath.h"
#include "ktm_postmx.h"
#include "ktm_postfft.h"

#define EOM_NT 256
#define EOM_RUN 16     /* consecutive tiles per XCD run */
#define EOM_WAVE 64
#define EOM_PI 3.14159265359 /* EOM.c:124 */

#define EOM_TRY(expr)                                  \
    do {                                               \
        hipError_t e_ = (expr);                        \
        if (e_ != hipSuccess) return KTM_ERR_DEVICE;   \
    } while (0)

/* ------------------------------------------------------------------------ */
/* host setup                                                                */
/* ------------------------------------------------------------------------ */
extern "C" void ktm_eom_config_default(ktm_eom_config *c)
{
    memset(c, 0, sizeof *c);
    c->ox = 789868.0;           /* EOM.c:66-68 */
    c->oy = 265844.0;
    c->angle_e_to_n = 0.3;
    c->oxcsp = 789868.0;        /* EOM.c:73-76 */
    c->oycsp = 265844.0;
    c->excsp = 819945.0;
    c->eycsp = 265999.0;
    c->scale_xy = 0.001;        /* EOM.c:77 */
    c->dcsp_x = 12.5;           /* EOM.c:79-81 */
    c->dcsp_y = 0.0;
    c->dh = 6.25;
    c->aperture = 6000.0;       /* EOM.c:85-86 */
    c->low_pass_cutoff = 50.0;
    c->nh = 1000;               /* EOM.c:82 */
    c->oh = 0;                  /* EOM.c:99 */
    c->jobs_x = 36;             /* EOM.c:91-92 */
    c->jobs_y = 1;
    c->ntblock = 20000;         /* EOM.c:89 */
    c->velocity_start = 3000.0; /* EOM.c:61-62 */
    c->velocity_end = 3000.0;
}

extern "C" int ktm_eom_setup_init(const ktm_eom_config *c, int ns, float dt, ktm_eom_setup *s,
                                  float *lowpass, float *rootrho)
{
    if (!c || !s || ns < 2 || !(dt > 0.0f)) return KTM_ERR_ARG;
    if (c->jobs_x <= 0 || c->jobs_y <= 0 || c->nh < 2 || !(c->dh != 0.0)) return KTM_ERR_ARG;
    if (!(c->aperture > 0.0) || !(c->aperture < 1e30)) return KTM_ERR_ARG;
    memset(s, 0, sizeof *s);
    s->ns = ns;
    s->dt = dt;
    /* EOM.c:279-320 */
    s->Nx = c->dcsp_x == 0.0 ? 1 : (int)ceil((c->excsp - c->oxcsp) / c->dcsp_x);
    s->Ny = c->dcsp_y == 0.0 ? 1 : (int)ceil((c->eycsp - c->oycsp) / c->dcsp_y);
    if (s->Nx < 0 || s->Ny < 0) return KTM_ERR_ARG;   /* end before origin */
    if (s->Nx == 0) s->Nx = 1;
    if (s->Ny == 0) s->Ny = 1;
    s->sinA = (float)sin(c->angle_e_to_n * EOM_PI / 180.0);
    s->cosA = (float)cos(c->angle_e_to_n * EOM_PI / 180.0);
    s->start_x = (float)(s->cosA * (c->oxcsp - c->ox) + s->sinA * (c->oycsp - c->oy));
    s->start_y = (float)(-s->sinA * (c->oxcsp - c->ox) + s->cosA * (c->oycsp - c->oy));
    s->no_job = c->jobs_x * c->jobs_y;
    s->nx = (int)ceil(1.0 * s->Nx / c->jobs_x);
    s->ny = (int)ceil(1.0 * s->Ny / c->jobs_y);
    if (s->nx == 0) s->nx = 1;
    if (s->ny == 0) s->ny = 1;
    s->nxp = s->nx * c->jobs_x;
    s->nyp = s->ny * c->jobs_y;
    s->nh = c->nh;
    /* EOM.c:759-845: calloc'd spectra mirrored from ns/2 upwards (for odd ns
     * the bins ns/2 and ns/2+1 read a not-yet-set zero) */
    if (lowpass || rootrho) {
        float *a = (float *)calloc((size_t)ns, sizeof(float));
        float *r = (float *)calloc((size_t)ns, sizeof(float));
        if (!a || !r) { free(a); free(r); return KTM_ERR_NOMEM; }
        const float fs = (float)(2.0 / dt / ns);
        const float fmax = (float)c->low_pass_cutoff;
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
    return KTM_OK;
}

struct ktm_eom_plan {
    ktm_eom_config cfg;
    ktm_eom_setup s;
    int device;
    float aperture2;
    float *d_cspx, *d_cspy;   /* [nxp], [nyp] CSP coordinates, EOM.c:1261-1262 */
    float *d_v2, *d_t2d;      /* [ns] v^2 (EOM.c:877), sqrt(t dt) t dt (EOM.c:1196) */
    double *d_sv;             /* [ns] sqrt(v^2) in double (EOM.c:1290) */
    double *d_g2;             /* [2 ns] post-filter kernel, g2[m] = g[m mod ns] */
    double2 *d_pf_g, *d_pf_tw; /* [N] FFT form of the post-filter kernel and twiddles */
    int pf_n;                  /* points of the FFT form, 0: the direct product */
    float4 *d_geo;            /* [cap] per-trace h2, cmpX, cmpY, talpha */
    int64_t geo_cap;
    float v2_0;               /* v^2 at t = 0 (talpha, EOM.c:1272) */
};

extern "C" int ktm_eom_plan_destroy(ktm_eom_plan *p)
{
    if (!p) return KTM_OK;
    (void)hipSetDevice(p->device);
    void *bufs[] = {p->d_cspx, p->d_cspy, p->d_v2, p->d_t2d, p->d_sv, p->d_g2, p->d_geo,
                    p->d_pf_g, p->d_pf_tw};
    for (void *b : bufs)
        if (b) (void)hipFree(b);
    free(p);
    return KTM_OK;
}

extern "C" int ktm_eom_plan_create(ktm_eom_plan **out, const ktm_eom_config *c, int ns, float dt,
                                   const float *vel)
{
    if (!out || !c || !vel) return KTM_ERR_ARG;
    *out = NULL;
    ktm_eom_plan *p = (ktm_eom_plan *)calloc(1, sizeof *p);
    if (!p) return KTM_ERR_NOMEM;
    float *lp = (float *)malloc(sizeof(float) * (size_t)ns);
    float *rr = (float *)malloc(sizeof(float) * (size_t)ns);
    int st = (lp && rr) ? ktm_eom_setup_init(c, ns, dt, &p->s, lp, rr) : KTM_ERR_NOMEM;
    const ktm_eom_setup *s = &p->s;
    float *cx = NULL, *cy = NULL, *v2 = NULL, *t2d = NULL;
    double *sv = NULL, *g = NULL;
    if (st == KTM_OK) {
        p->cfg = *c;
        p->aperture2 = (float)(c->aperture * c->aperture);   /* EOM.c:1176 */
        p->v2_0 = vel[0] * vel[0];
        cx = (float *)malloc(sizeof(float) * (size_t)s->nxp);
        cy = (float *)malloc(sizeof(float) * (size_t)s->nyp);
        v2 = (float *)malloc(sizeof(float) * (size_t)ns);
        t2d = (float *)malloc(sizeof(float) * (size_t)ns);
        sv = (double *)malloc(sizeof(double) * (size_t)ns);
        g = (double *)malloc(sizeof(double) * 2 * (size_t)ns);
        if (!cx || !cy || !v2 || !t2d || !sv || !g) st = KTM_ERR_NOMEM;
    }
    if (st == KTM_OK) {
        for (int X = 0; X < s->nxp; ++X) cx[X] = (float)(s->start_x + X * c->dcsp_x);
        for (int Y = 0; Y < s->nyp; ++Y) cy[Y] = (float)(s->start_y + Y * c->dcsp_y);
        for (int t = 0; t < ns; ++t) {
            v2[t] = vel[t] * vel[t];
            if (!(v2[t] > 0.0f) || !(v2[t] < 3.0e38f)) st = KTM_ERR_VELOCITY;
            t2d[t] = (float)(sqrt((double)(t * dt)) * t * dt);
            sv[t] = sqrt((double)v2[t]);
        }
        /* g[m] = (1/ns) sum_k lowpass[k] rootrho[k] cos(2 pi k m / ns): the
         * spectrum is even (both factors are mirrored), so the filtered trace
         * is the circular convolution with g */
        const long double two_pi = 2.0L * acosl(-1.0L);
        for (int m = 0; m < ns; ++m) {
            long double acc = 0.0L;
            long long km = 0;
            for (int k = 0; k < ns; ++k) {
                acc += (long double)lp[k] * (long double)rr[k] * cosl(two_pi * km / ns);
                km += m;
                if (km >= ns) km -= ns;
            }
            g[m] = g[m + ns] = (double)(acc / ns);
        }
    }
    if (st == KTM_OK && hipGetDevice(&p->device) != hipSuccess) st = KTM_ERR_DEVICE;
    if (st == KTM_OK &&
        (hipMalloc(&p->d_cspx, sizeof(float) * s->nxp) != hipSuccess ||
         hipMalloc(&p->d_cspy, sizeof(float) * s->nyp) != hipSuccess ||
         hipMalloc(&p->d_v2, sizeof(float) * ns) != hipSuccess ||
         hipMalloc(&p->d_t2d, sizeof(float) * ns) != hipSuccess ||
         hipMalloc(&p->d_sv, sizeof(double) * ns) != hipSuccess ||
         hipMalloc(&p->d_g2, sizeof(double) * 2 * ns) != hipSuccess))
        st = KTM_ERR_NOMEM;
    if (st == KTM_OK &&
        (hipMemcpy(p->d_cspx, cx, sizeof(float) * s->nxp, hipMemcpyHostToDevice) != hipSuccess ||
         hipMemcpy(p->d_cspy, cy, sizeof(float) * s->nyp, hipMemcpyHostToDevice) != hipSuccess ||
         hipMemcpy(p->d_v2, v2, sizeof(float) * ns, hipMemcpyHostToDevice) != hipSuccess ||
         hipMemcpy(p->d_t2d, t2d, sizeof(float) * ns, hipMemcpyHostToDevice) != hipSuccess ||
         hipMemcpy(p->d_sv, sv, sizeof(double) * ns, hipMemcpyHostToDevice) != hipSuccess ||
         hipMemcpy(p->d_g2, g, sizeof(double) * 2 * ns, hipMemcpyHostToDevice) != hipSuccess))
        st = KTM_ERR_DEVICE;
    if (st == KTM_OK && (p->pf_n = postfft_size(ns)) != 0) {
        const size_t n = (size_t)p->pf_n;
        double *t = (double *)malloc(sizeof(double) * 4 * n);
        if (!t) st = KTM_ERR_NOMEM;
        if (st == KTM_OK) {
            postfft_host_tables(g, ns, p->pf_n, t, t + 2 * n);
            if (hipMalloc(&p->d_pf_g, sizeof(double2) * n) != hipSuccess ||
                hipMalloc(&p->d_pf_tw, sizeof(double2) * n) != hipSuccess)
                st = KTM_ERR_NOMEM;
            else if (hipMemcpy(p->d_pf_g, t, sizeof(double2) * n, hipMemcpyHostToDevice) != hipSuccess ||
                     hipMemcpy(p->d_pf_tw, t + 2 * n, sizeof(double2) * n, hipMemcpyHostToDevice) !=
                         hipSuccess)
                st = KTM_ERR_DEVICE;
        }
        free(t);
    }
    free(lp); free(rr); free(cx); free(cy); free(v2); free(t2d); free(sv); free(g);
    if (st != KTM_OK) {
        ktm_eom_plan_destroy(p);
        return st;
    }
    *out = p;
    return KTM_OK;
}

extern "C" int ktm_eom_plan_setup(const ktm_eom_plan *p, ktm_eom_setup *out)
{
    if (!p || !out) return KTM_ERR_ARG;
    *out = p->s;
    return KTM_OK;
}

extern "C" int64_t ktm_eom_image_size(const ktm_eom_plan *p)
{
    return p ? (int64_t)p->s.nxp * p->s.nyp * p->s.nh * p->s.ns : -1;
}

/* ------------------------------------------------------------------------ */
/* kernels                                                                   */
/* ------------------------------------------------------------------------ */
struct EomGeoArgs {
    double scale_xy, ox, oy;
    float sinA, cosA, v2_0, dt;
};

/* EOM.c:1245-1272 per trace: scaleXY * htonl(word) - ox, rotation, h2, CMP,
 * talpha = round(2 sqrt(h2 / v^2[0]) / dt) */
__global__ void k_eom_geometry(const EomGeoArgs a, const int32_t *__restrict__ coords, int64_t ntr,
                               float4 *__restrict__ geo)
{
    const int64_t i = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
    if (i >= ntr) return;
    const uint32_t *w = reinterpret_cast<const uint32_t *>(coords + 4 * i);
    float io = (float)(a.scale_xy * (double)w[0] - a.ox), jo = (float)(a.scale_xy * (double)w[1] - a.oy);
    const float sxt = a.cosA * io + a.sinA * jo;
    const float syt = -a.sinA * io + a.cosA * jo;
    io = (float)(a.scale_xy * (double)w[2] - a.ox);
    jo = (float)(a.scale_xy * (double)w[3] - a.oy);
    const float gxt = a.cosA * io + a.sinA * jo;
    const float gyt = -a.sinA * io + a.cosA * jo;
    const float h2 = (float)(0.25 * (double)((sxt - gxt) * (sxt - gxt) + (syt - gyt) * (syt - gyt)));
    const float cmpX = (float)(0.5 * (double)(sxt + gxt));
    const float cmpY = (float)(0.5 * (double)(syt + gyt));
    const double ta = round(2.0 * sqrt_cr((double)(h2 / a.v2_0)) / (double)a.dt);
    geo[i] = make_float4(h2, cmpX, cmpY, (float)ta);   /* talpha: an integer, exact in float */
}

struct EomArgs {
    const float *traces;          /* [ntr][ns] */
    const float4 *geo;            /* [ntr] h2, cmpX, cmpY, talpha */
    int64_t ntr;
    const float *cspx, *cspy, *v2, *t2d;
    const double *sv;
    float *csp, *fold;            /* [nxp][nyp][nh][ns], or the job's [nx][ny][nh][ns] */
    unsigned long long *counter;
    float aperture2, dt;
    double dh;
    int oh, ns, nh, nseg;
    int rx0, ry0, nrx, nry;       /* CSP range of the launch */
    int cx0, cy0, nyl;            /* gather array: CSP (cx0, cy0) first, nyl per X */
};

__global__ void __launch_bounds__(EOM_NT) k_eom_accum(const EomArgs a)
{
    /* per passing trace: 4 x2 h2 (exact in double) and RN(x2 + h2), the
     * trace-only terms of EOM.c:1285-1286 */
    __shared__ double sXH4[EOM_NT];
    __shared__ float sXpH[EOM_NT];
    __shared__ int sTa[EOM_NT], sIdx[EOM_NT], sCnt[EOM_NT / EOM_WAVE];
    const int tid = threadIdx.x, lane = tid & (EOM_WAVE - 1), wv = tid / EOM_WAVE;
    /* XCD-aware order: consecutive workgroup ids go to different XCDs; each
     * XCD takes runs of EOM_RUN consecutive (CSP, segment) tiles (neighbouring
     * CSPs share traces in its L2), and the runs are dealt round-robin so
     * every XCD gets CSPs from the whole survey (the trace count per CSP
     * falls towards the survey's edges) */
    const int ntiles = a.nrx * a.nry * a.nseg;
    const int i = blockIdx.x >> 3;
    const int tile = (((i / EOM_RUN) << 3) + (int)(blockIdx.x & 7)) * EOM_RUN + i % EOM_RUN;
    if (tile >= ntiles) return;
    const int seg = tile % a.nseg, col = tile / a.nseg;
    const int X = a.rx0 + col / a.nry, Y = a.ry0 + col % a.nry;
    const float cspX = a.cspx[X], cspY = a.cspy[Y];
    const int t = seg * EOM_NT + tid;
    const bool tin = t < a.ns;
    const int tc = tin ? t : a.ns - 1;
    /* per-sample constants of this lane */
    const float tdt = (float)t * a.dt;                            /* t*dt (float) */
    const float den0 = tdt * (float)t * a.dt * a.v2[tc];          /* t*dt*t*dt*vel_t[t] */
    const double svt = a.sv[tc];
    /* reciprocals for the correctly rounded divisions below (div_cr): the
     * quotients are those of the reference's double divisions */
    const double rden0 = 1.0 / (double)den0;   /* +inf at t = 0: NaN offset, as 0-divides */
    const double rsvt = 1.0 / svt, rdh = 1.0 / a.dh;
    constexpr double R008 = 1.0 / 0.080, RPI = 1.0 / EOM_PI;
    const float t2dt = a.t2d[tc];
    const float rt2d = 1.0f / t2dt;   /* t = 0 never bins: its offset is NaN */
    const int64_t col_base = ((int64_t)(X - a.cx0) * a.nyl + (Y - a.cy0)) * a.nh;
    uint64_t nev = 0;
    /* the lane's sample of two offset bins, held in registers between traces */
    int tagE = -1, tagO = -1;
    float cE = 0.0f, fE = 0.0f, cO = 0.0f, fO = 0.0f;
    for (int64_t c0 = 0; c0 < a.ntr; c0 += EOM_NT) {
        /* ---- aperture gate of EOM.c:1268-1270 for 256 traces, in order ---- */
        const int64_t i = c0 + tid;
        bool pass = false;
        float x2 = 0.0f, h2 = 0.0f;
        int ta = 0;
        if (i < a.ntr) {
            const float4 g = a.geo[i];
            const float dx = g.y - cspX, dy = g.z - cspY;
            x2 = dx * dx + dy * dy;
            h2 = g.x;
            ta = (int)g.w;
            pass = x2 + h2 < a.aperture2;
        }
        const uint64_t m = __ballot(pass);
        const int pre = __popcll(m & ((1ull << lane) - 1ull));
        if (lane == 0) sCnt[wv] = __popcll(m);
        __syncthreads();
        int base = 0, total = 0;
#pragma unroll
        for (int q = 0; q < EOM_NT / EOM_WAVE; ++q) {
            base += q < wv ? sCnt[q] : 0;
            total += sCnt[q];
        }
        if (pass) {
            sXH4[base + pre] = 4.0 * (double)x2 * (double)h2;
            sXpH[base + pre] = x2 + h2;
            sTa[base + pre] = ta;
            sIdx[base + pre] = tid;
        }
        __syncthreads();
        for (int j = 0; j < total; ++j) {
            const int taj = sTa[j];
            const bool live = tin && t >= taj;
            nev += __popcll(__ballot(live));
            if (!live) continue;
            /* EOM.c:1285-1306 with the reference's promotions */
            const float cross = (float)div_cr(sXH4[j], (double)den0, rden0);
            const float arg = sXpH[j] - cross;
            if (!(arg >= 0.0f)) continue;                 /* NaN offset: no contribution */
            /* (float)sqrt((double)arg) of the reference: the double square
             * root of a float, rounded to float, is the correctly rounded
             * float square root (double rounding is innocuous for sqrt when
             * 53 >= 2 x 24 + 2), which __builtin_sqrtf is without fast-math */
            const float eq = __builtin_sqrtf(arg);
            const double qo = div_cr((double)(eq - (float)a.oh), a.dh, rdh);
            const double fl = floor(qo);
            if (!(fl >= 0.0 && fl < (double)(a.nh - 1))) continue;
            const int off = (int)fl;
            const float tcv = (float)div_cr(2.25 * (double)eq, svt, rsvt);
            const double z = div_cr(200.0 * (double)(tdt - tcv), 0.080, R008);
            const float s3 = (float)(0.5 + div_cr(atan(z), EOM_PI, RPI));
            const float s2 = (float)(qo - (double)off);
            const float s1 = (float)(1.0 - (double)s2);
            const float x = a.traces[(c0 + sIdx[j]) * a.ns + t];
            const float v1 = div_cr(s1 * s3 * x, t2dt, rt2d), v2 = div_cr(s2 * s3 * x, t2dt, rt2d);
            /* bins off and off+1 have opposite parity: the even one goes to
             * slot E, the odd one to slot O; a slot that holds another bin
             * is written back and refilled, so every bin lives in exactly
             * one place and receives its adds in trace order */
            const bool ev = (off & 1) == 0;
            const int be = ev ? off : off + 1, bo = ev ? off + 1 : off;
            if (tagE != be) {
                if (tagE >= 0) {
                    const int64_t w = (col_base + tagE) * a.ns + t;
                    a.csp[w] = cE;
                    a.fold[w] = fE;
                }
                const int64_t r = (col_base + be) * a.ns + t;
                cE = a.csp[r];
                fE = a.fold[r];
                tagE = be;
            }
            if (tagO != bo) {
                if (tagO >= 0) {
                    const int64_t w = (col_base + tagO) * a.ns + t;
                    a.csp[w] = cO;
                    a.fold[w] = fO;
                }
                const int64_t r = (col_base + bo) * a.ns + t;
                cO = a.csp[r];
                fO = a.fold[r];
                tagO = bo;
            }
            cE = cE + (ev ? v1 : v2);
            fE = fE + 1.0f;
            cO = cO + (ev ? v2 : v1);
            fO = fO + 1.0f;
        }
        __syncthreads();
    }
    if (tagE >= 0) {
        const int64_t w = (col_base + tagE) * a.ns + t;
        a.csp[w] = cE;
        a.fold[w] = fE;
    }
    if (tagO >= 0) {
        const int64_t w = (col_base + tagO) * a.ns + t;
        a.csp[w] = cO;
        a.fold[w] = fO;
    }
    if (lane == 0 && a.counter) atomicAdd(a.counter, (unsigned long long)nev);
}

__global__ void k_eom_reset(float *csp, float *fold, int64_t n)
{
    for (int64_t i = (int64_t)blockIdx.x * blockDim.x + threadIdx.x; i < n;
         i += (int64_t)gridDim.x * blockDim.x) {
        csp[i] = 0.0f;    /* EOM.c:1202-1211 */
        fold[i] = 1.0f;
    }
}

/* x of the post filter, EOM.c:1391: the float (csp / fold) dt t, stored into
 * a double; out = g (*) x in FP64 (k_postfft or k_postmx), out[0] = 0 */
struct PostEom {
    static __device__ __forceinline__ double x(const float *c, const float *f, int u, float dt)
    {
        return (double)((c[u] / f[u]) * dt * (float)u);
    }
};

/* ------------------------------------------------------------------------ */
/* entry points                                                              */
/* ------------------------------------------------------------------------ */
extern "C" int ktm_eom_reset(const ktm_eom_plan *p, float *csp, float *fold, void *stream)
{
    if (!p || !csp || !fold) return KTM_ERR_ARG;
    hipLaunchKernelGGL(k_eom_reset, dim3(2048), dim3(256), 0, (hipStream_t)stream, csp, fold,
                       ktm_eom_image_size(p));
    EOM_TRY(hipGetLastError());
    return KTM_OK;
}

/* Bin a block of traces into the gathers of every slab (job < 0) or one;
 * local: the gathers hold the job's slab only, [nx][ny][nh][ns]. */
static int eom_block(const ktm_eom_plan *cp, const float *traces, const int32_t *coords,
                     int64_t ntr, int job, bool local, float *csp, float *fold, uint64_t *counter,
                     void *stream)
{
    if (!cp || ntr < 0) return KTM_ERR_ARG;
    if (!ntr) return KTM_OK;
    if (!traces || !coords || !csp || !fold) return KTM_ERR_ARG;
    ktm_eom_plan *p = const_cast<ktm_eom_plan *>(cp);
    const ktm_eom_setup *s = &p->s;
    if (job >= s->no_job || (local && job < 0)) return KTM_ERR_ARG;
    hipStream_t st = (hipStream_t)stream;
    EOM_TRY(hipSetDevice(p->device));
    if (ntr > p->geo_cap) {
        if (p->d_geo) EOM_TRY(hipFree(p->d_geo));
        p->d_geo = nullptr;
        p->geo_cap = 0;
        if (hipMalloc((void **)&p->d_geo, sizeof(float4) * (size_t)ntr) != hipSuccess) return KTM_ERR_NOMEM;
        p->geo_cap = ntr;
    }
    EomGeoArgs ga;
    ga.scale_xy = p->cfg.scale_xy; ga.ox = p->cfg.ox; ga.oy = p->cfg.oy;
    ga.sinA = s->sinA; ga.cosA = s->cosA; ga.v2_0 = p->v2_0; ga.dt = s->dt;
    hipLaunchKernelGGL(k_eom_geometry, dim3((unsigned)((ntr + 255) / 256)), dim3(256), 0, st, ga,
                       coords, ntr, p->d_geo);
    EomArgs a;
    a.traces = traces; a.geo = p->d_geo; a.ntr = ntr;
    a.cspx = p->d_cspx; a.cspy = p->d_cspy; a.v2 = p->d_v2; a.t2d = p->d_t2d; a.sv = p->d_sv;
    a.csp = csp; a.fold = fold; a.counter = (unsigned long long *)counter;
    a.aperture2 = p->aperture2; a.dt = s->dt; a.dh = p->cfg.dh; a.oh = p->cfg.oh;
    a.ns = s->ns; a.nh = s->nh;
    a.nseg = (s->ns + EOM_NT - 1) / EOM_NT;
    if (job < 0) {
        a.rx0 = 0; a.ry0 = 0; a.nrx = s->nxp; a.nry = s->nyp;
    } else {   /* the job's slab, EOM.c:1261-1262 */
        a.rx0 = (job / p->cfg.jobs_y) * s->nx; a.ry0 = (job % p->cfg.jobs_y) * s->ny;
        a.nrx = s->nx; a.nry = s->ny;
    }
    if (local) {
        a.cx0 = a.rx0; a.cy0 = a.ry0; a.nyl = s->ny;
    } else {
        a.cx0 = 0; a.cy0 = 0; a.nyl = s->nyp;
    }
    const int64_t ntiles = (int64_t)a.nrx * a.nry * a.nseg;
    if (ntiles > 0x7fffff00LL) return KTM_ERR_UNSUPPORTED;
    const int64_t span = 8 * EOM_RUN;   /* tiles per round of runs over the 8 XCDs */
    const int64_t nblk = (ntiles + span - 1) / span * span;
    hipLaunchKernelGGL(k_eom_accum, dim3((unsigned)nblk), dim3(EOM_NT), 0, st, a);
    EOM_TRY(hipGetLastError());
    return KTM_OK;
}

extern "C" int ktm_eom_block(const ktm_eom_plan *p, const float *traces, const int32_t *coords,
                             int64_t ntr, int job, float *csp, float *fold, uint64_t *counter,
                             void *stream)
{
    return eom_block(p, traces, coords, ntr, job, false, csp, fold, counter, stream);
}

extern "C" int ktm_eom_job_block(const ktm_eom_plan *p, const float *traces,
                                 const int32_t *coords, int64_t ntr, int job, float *csp,
                                 float *fold, uint64_t *counter, void *stream)
{
    return eom_block(p, traces, coords, ntr, job, true, csp, fold, counter, stream);
}


static int eom_post(const ktm_eom_plan *p, const float *csp, const float *fold, int64_t ntraces,
                    float *out, void *stream)
{
    if (!p || !csp || !fold || !out) return KTM_ERR_ARG;
    const int ns = p->s.ns;
    if (ns > 8192) return KTM_ERR_UNSUPPORTED;
    EOM_TRY(hipSetDevice(p->device));
    const hipError_t e =
        p->pf_n ? postfft_launch<PostEom>(csp, fold, ntraces, ns, p->s.dt, p->d_pf_g, p->d_pf_tw,
                                             out, (hipStream_t)stream)
                   : postmx_launch<PostEom>(csp, fold, ntraces, ns, p->s.dt, p->d_g2, out,
                                            (hipStream_t)stream);
    return e == hipSuccess ? KTM_OK : e == hipErrorInvalidValue ? KTM_ERR_UNSUPPORTED : KTM_ERR_DEVICE;
}

extern "C" int ktm_eom_post(const ktm_eom_plan *p, const float *csp, const float *fold, float *out,
                            void *stream)
{
    if (!p) return KTM_ERR_ARG;
    return eom_post(p, csp, fold, (int64_t)p->s.nxp * p->s.nyp * p->s.nh, out, stream);
}

extern "C" int64_t ktm_eom_job_size(const ktm_eom_plan *p)
{
    return p ? (int64_t)p->s.nx * p->s.ny * p->s.nh * p->s.ns : -1;
}

extern "C" int ktm_eom_job_reset(const ktm_eom_plan *p, float *csp, float *fold, void *stream)
{
    if (!p || !csp || !fold) return KTM_ERR_ARG;
    hipLaunchKernelGGL(k_eom_reset, dim3(2048), dim3(256), 0, (hipStream_t)stream, csp, fold,
                       ktm_eom_job_size(p));
    EOM_TRY(hipGetLastError());
    return KTM_OK;
}

extern "C" int ktm_eom_job_post(const ktm_eom_plan *p, const float *csp, const float *fold,
                                float *out, void *stream)
{
    if (!p) return KTM_ERR_ARG;
    return eom_post(p, csp, fold, (int64_t)p->s.nx * p->s.ny * p->s.nh, out, stream);
}

/* ------------------------------------------------------------------------ */
/* image<job>.su records, EOM.c:1346-1537                                    */
/* ------------------------------------------------------------------------ */
struct EomRecArgs {
    const float *cspx, *cspy;     /* plan tables, global CSP index */
    int64_t nrec;                 /* records written (a prefix of the job's) */
    int64_t it0;                  /* running number of the job's first record */
    int nx, ny, nh, ns, X0, Y0, csp0;
    double ox, oy, dh;
    int oh;
    float cosA, sinA;
    uint32_t dt_us;
};

/* one thread per 32-bit word of the records: header word or sample */
__global__ void k_eom_records(const EomRecArgs a, const float *__restrict__ post,
                              uint32_t *__restrict__ rec)
{
    const int64_t e = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
    const int64_t W = 60 + (int64_t)a.ns;
    if (e >= a.nrec * W) return;
    const int64_t r = e / W;
    const int w = (int)(e - r * W);
    uint32_t v = 0u;
    if (w >= 60) {
        /* data[0] = 0, data[t] = ntohf(out[t]) (EOM.c:1409-1412) */
        const int t = w - 60;
        v = t == 0 ? 0u : __builtin_bswap32(__float_as_uint(post[r * a.ns + t]));
    } else {
        const int ih = (int)(r % a.nh);
        const int64_t c = r / a.nh;
        const int xc = (int)(c / a.ny), yc = (int)(c % a.ny);
        switch (w) {
        case 0: v = __builtin_bswap32((uint32_t)(a.it0 + r)); break;                   /* tracl = it */
        case 5: case 47:                                                                /* cdp, iline = csp */
            v = __builtin_bswap32((uint32_t)(a.csp0 + xc * a.ny + yc)); break;
        case 9: v = __builtin_bswap32((uint32_t)(int)ceil(a.oh + ih * a.dh)); break;   /* offset */
        case 18: case 19: case 20: case 21:                                            /* sx..gy = 13..16 */
            v = __builtin_bswap32((uint32_t)(13 + w - 18)); break;
        case 28: v = __builtin_bswap32((uint32_t)(a.ns & 0xffff)); break;              /* ns, byte 114 */
        case 29: v = __builtin_bswap32(a.dt_us << 16); break;                          /* dt, byte 116 */
        case 45: case 46: {                                   /* cdpx, cdpy = htonl(o + (int)rotated) */
            const float cX = a.cspx[a.X0 + xc], cY = a.cspy[a.Y0 + yc];
            const int xc1 = (int)(a.cosA * cX - a.sinA * cY);
            const int yc1 = (int)(a.sinA * cX + a.cosA * cY);
            const double o = w == 45 ? a.ox + (double)xc1 : a.oy + (double)yc1;
            v = __builtin_bswap32((uint32_t)(int64_t)o);
            break;
        }
        case 48: v = __builtin_bswap32(1u); break;                                     /* cline */
        case 51: case 55: case 56: case 58:   /* mess: 0 in the first record written, then 9999 */
            v = r > 0 ? __builtin_bswap32(9999u) : 0u; break;
        case 59: v = __builtin_bswap32(9999u); break;                                  /* EOM.c:1527-1528 */
        default: break;
        }
    }
    rec[e] = v;
}

extern "C" int ktm_eom_job_records(const ktm_eom_plan *p, const float *post, int job,
                                   void *records, int64_t *nrec, void *stream)
{
    if (!p || !post || !records || !nrec || ((uintptr_t)records & 3u)) return KTM_ERR_ARG;
    const ktm_eom_setup *s = &p->s;
    if (job < 0 || job >= s->no_job) return KTM_ERR_ARG;
    /* it runs from nx*ny*nh*job + 1 and a record is written while
     * it < Nx*Ny*nh + 1 (EOM.c:1347, 1417): a prefix of the job's records */
    const int64_t total = (int64_t)s->nx * s->ny * s->nh;
    const int64_t it0 = total * job + 1;
    int64_t n = (int64_t)s->Nx * s->Ny * s->nh + 1 - it0;
    n = n < 0 ? 0 : n > total ? total : n;
    *nrec = n;
    if (!n) return KTM_OK;
    EomRecArgs a;
    a.cspx = p->d_cspx; a.cspy = p->d_cspy;
    a.nrec = n; a.it0 = it0;
    a.nx = s->nx; a.ny = s->ny; a.nh = s->nh; a.ns = s->ns;
    a.X0 = (job / p->cfg.jobs_y) * s->nx; a.Y0 = (job % p->cfg.jobs_y) * s->ny;
    a.csp0 = s->nx * s->ny * job + 1;                       /* EOM.c:1346 */
    a.ox = p->cfg.ox; a.oy = p->cfg.oy; a.dh = p->cfg.dh; a.oh = p->cfg.oh;
    a.cosA = s->cosA; a.sinA = s->sinA;
    a.dt_us = (uint32_t)(uint16_t)(s->dt * 1000000);
    const int64_t words = n * (60 + (int64_t)s->ns);
    const int nt = 256;
    if ((words + nt - 1) / nt > 0x7fffffffLL) return KTM_ERR_UNSUPPORTED;
    EOM_TRY(hipSetDevice(p->device));
    hipLaunchKernelGGL(k_eom_records, dim3((unsigned)((words + nt - 1) / nt)), dim3(nt), 0,
                       (hipStream_t)stream, a, post, (uint32_t *)records);
    EOM_TRY(hipGetLastError());
    return KTM_OK;
}
