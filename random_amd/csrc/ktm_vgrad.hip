/*
 * ktm_vgrad.hip -- velocity gradient of a gather-flatness objective.
 *
 * A differentiable variant of the migration of Kirchhoff.c:1332-1404 for
 * velocity analysis.  The reference picks the velocity of the nearest grid
 * node (1357-1358,1376) and the sample floor(dsr/dt) (1384); both have zero
 * derivative almost everywhere.  Here:
 *   v(t, X)   = sum_ij V[i][j] B(t/dtn + 1 - i) B(X/dxn + 1 - j)
 *               uniform cubic B-spline, B centred with support (-2, 2),
 *               dtn = (ns-1)/(nt_nodes-3) samples, dxn = (nxp-1)/(nx_nodes-3)
 *               columns, so every image point has its four nodes per axis;
 *   q         = (sqrt(t2 + a/v^2) + sqrt(t2 + b/v^2)) / dt, t2 = (t dt / 2)^2,
 *               the double-square-root time of Kirchhoff.c:1381-1383 in
 *               samples (not floored);
 *   I[X][h][t] = sum over the traces of offset class h that pass the aperture
 *               gate of Kirchhoff.c:1354 of F(q), F the CLOSED filtered trace
 *               interpolated by the Catmull-Rom cubic (C1 in q), zero outside
 *               [0, ns); q is clamped to ns + 2 where every tap is zero;
 *   J         = sum_{X,t} sum_{h < nh-1} (I[X][h+1][t] - I[X][h][t])^2.
 *
 * k_vgrad: one workgroup owns one image column X and one time chunk of
 * `chunk` (<= 256) samples, one lane per sample.  The offset classes stream
 * in order: pass 1 sums class h+1 (Kahan) while the lane holds I[h-1..h]
 * of its sample in registers, adds its share of J (FP64), and with dJ/dI[h]
 * known pass 2 walks class h's traces again and recomputes each traveltime
 * to accumulate
 *   dJ/dv(X,t) = sum_h dJ/dI[h][t] sum_c F'(q_c) dq_c/dv.
 * Nothing per contribution is stored: the time chunk and a window of three
 * classes bound the working set to registers.  Traces are binned on the host
 * by (offset class, CMP inline coordinate), so a column's traces of one class
 * are a contiguous range; every wave walks the ranges itself (gates
 * lane-parallel, ballot, then the passing traces in order) and needs no
 * barrier.  At bind, k_vg_tables lays the geometry and the traces out in
 * that order, each trace zero-padded (one sample before, five after): the
 * four taps of a contribution are one 16-byte load, with no clamp and no
 * trace-index lookup in the loop.
 * k_vg_project_t / k_vg_project_x: dJ/dV = B_t^T (dJ/dv) B_x in FP64.
 */
#include <hip/hip_runtime.h>

#include <math.h>
#include <stdint.h>
#include <stdlib.h>
#include <string.h>

#include "ktm_internal.h"

#define VG_NT 256
#define VG_WAVE 64
/* passing traces evaluated per step, their tap loads issued together: 4
 * (44 VGPRs, 8 waves per SIMD) measured 8 % faster than 1 or 2 and 2 %
 * faster than 6 (60 VGPRs) or 8 (76 VGPRs, 6 waves) */
#define VG_G 4

struct ktm_vgrad {
    ktm_config cfg;
    ktm_tables tb;
    int device;
    int ns, nxp, nh, nt_nodes, nx_nodes, chunk, nchunks;
    float dt;
    double dtn, dxn;
    float *d_colx;
    float *d_v;            /* [nxp][ns] velocity on the image grid */
    float *d_gv;           /* [nxp][ns] dJ/dv */
    double *d_part;        /* [nxp * nchunks] shares of J */
    double *d_G1;          /* [nxp][nt_nodes] */
    double *h_part;
    int *d_flag;
    int32_t *d_order;      /* [nbin] trace indices sorted by (class, cmpx) */
    int32_t *d_ranges;     /* [nxp][nh][2] */
    int64_t order_cap;
    ktm_trace_geom *d_sgeom;  /* [nbin] geometry in binned order */
    float *d_taps;         /* [nbin][ns + 6] zero-padded traces in binned order */
    int64_t taps_cap;      /* rows of d_taps and d_sgeom */
    int64_t ntr;
    int64_t pairs;         /* (trace, column) pairs passing the aperture gate */
    int bound;
};

#define VG_TRY(expr)                                   \
    do {                                               \
        hipError_t e_ = (expr);                        \
        if (e_ != hipSuccess) return KTM_ERR_DEVICE;   \
    } while (0)

/* ------------------------------------------------------------------------ */
/* spline                                                                    */
/* ------------------------------------------------------------------------ */
__device__ __forceinline__ double bspline3(double z)
{
    const double a = fabs(z);
    if (a >= 2.0) return 0.0;
    if (a >= 1.0) return (2.0 - a) * (2.0 - a) * (2.0 - a) * (1.0 / 6.0);
    return 2.0 / 3.0 - a * a + 0.5 * a * a * a;
}

/* v(t, X) for every image point; flag = 1 if any v is not a positive finite
 * number. */
__global__ void k_vg_velocity(const float *__restrict__ V, int nt_nodes, int nx_nodes,
                              double dtn, double dxn, int ns, int nxp, float *__restrict__ v,
                              int *__restrict__ flag)
{
    const int64_t e = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
    if (e >= (int64_t)nxp * ns) return;
    const int X = (int)(e / ns), t = (int)(e % ns);
    const double zt = t / dtn + 1.0, zx = X / dxn + 1.0;
    const int it = (int)floor(zt), ix = (int)floor(zx);
    double s = 0.0;
    for (int i = it - 1; i <= it + 2; ++i) {
        if (i < 0 || i >= nt_nodes) continue;
        const double bt = bspline3(zt - i);
        for (int j = ix - 1; j <= ix + 2; ++j) {
            if (j < 0 || j >= nx_nodes) continue;
            s += (double)V[(int64_t)i * nx_nodes + j] * bt * bspline3(zx - j);
        }
    }
    const float vf = (float)s;
    v[e] = vf;
    if (!(vf > 0.0f) || !(vf < 3.0e38f)) *flag = 1;
}

/* G1[X][i] = sum_t gv[X][t] B(t/dtn + 1 - i) */
__global__ void k_vg_project_t(const float *__restrict__ gv, int ns, int nxp, int nt_nodes,
                               double dtn, double *__restrict__ G1)
{
    const int64_t e = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
    if (e >= (int64_t)nxp * nt_nodes) return;
    const int X = (int)(e / nt_nodes), i = (int)(e % nt_nodes);
    /* B(t/dtn + 1 - i) != 0 for (i - 3) dtn < t < (i + 1) dtn */
    const int t0 = max(0, (int)floor((i - 3) * dtn));
    const int t1 = min(ns - 1, (int)ceil((i + 1) * dtn));
    double s = 0.0;
    for (int t = t0; t <= t1; ++t) s += (double)gv[(int64_t)X * ns + t] * bspline3(t / dtn + 1.0 - i);
    G1[e] = s;
}

/* dJdV[i][j] = sum_X G1[X][i] B(X/dxn + 1 - j) */
__global__ void k_vg_project_x(const double *__restrict__ G1, int nxp, int nt_nodes,
                               int nx_nodes, double dxn, double *__restrict__ dV)
{
    const int e = blockIdx.x * blockDim.x + threadIdx.x;
    if (e >= nt_nodes * nx_nodes) return;
    const int i = e / nx_nodes, j = e % nx_nodes;
    const int x0 = max(0, (int)floor((j - 3) * dxn));
    const int x1 = min(nxp - 1, (int)ceil((j + 1) * dxn));
    double s = 0.0;
    for (int X = x0; X <= x1; ++X) s += G1[(int64_t)X * nt_nodes + i] * bspline3(X / dxn + 1.0 - j);
    dV[e] = s;
}

/* ------------------------------------------------------------------------ */
/* migration into offset-class gathers, objective and dJ/dv                  */
/* ------------------------------------------------------------------------ */
struct VgArgs {
    const float *taps;             /* [nbin][trow] zero-padded CLOSED traces, binned order */
    int64_t trow;                  /* ns + 6 */
    const ktm_trace_geom *sgeom;   /* [nbin] geometry in binned order */
    const int32_t *ranges;
    const float *v, *colx;
    float coly, ap_x2thr, dt, rdt;
    int ns, nh, nxp, chunk, nchunks, run;
    float *cig;                    /* optional [nxp][nh][ns] */
    float *gv;                     /* optional [nxp][ns] */
    double *part;                  /* [nxp * nchunks] */
};

/* four consecutive floats at a 4-byte-aligned address (one dwordx4 load) */
typedef float vg_f4 __attribute__((ext_vector_type(4), aligned(4)));

/* Catmull-Rom cubic through the taps f = F[T-1..T+2] at q = T + u, u in
 * [0, 1) (taps outside [0, ns) are zero); returns the value, or with DERIV
 * its q-derivative */
template <bool DERIV>
__device__ __forceinline__ float cr_eval(const vg_f4 f, float u)
{
    const float c1 = 0.5f * (f.z - f.x);
    const float c2 = __builtin_fmaf(-0.5f, f.w, __builtin_fmaf(2.0f, f.z, __builtin_fmaf(-2.5f, f.y, f.x)));
    const float c3 = __builtin_fmaf(1.5f, f.y - f.z, 0.5f * (f.w - f.x));
    if (DERIV) return __builtin_fmaf(u, __builtin_fmaf(3.0f * u, c3, 2.0f * c2), c1);
    return __builtin_fmaf(u, __builtin_fmaf(u, __builtin_fmaf(u, c3, c2), c1), f.y);
}

/* The binned tables: sgeom[k] = geom[order[k]] and taps[k][p] = F[p - 1] of
 * trace order[k] for 1 <= p <= ns, zero for p = 0 and ns < p < trow
 * (trow = ns + 6), so the taps F[T-1..T+2] of every interval T in
 * [0, ns + 2] are taps[k][T..T+3], read without a clamp.  Grid: x over p,
 * y strides over the binned traces. */
__global__ void k_vg_tables(const float *__restrict__ prep, int64_t row,
                            const ktm_trace_geom *__restrict__ geom,
                            const int32_t *__restrict__ order, int64_t nbin, int ns, int64_t trow,
                            float *__restrict__ taps, ktm_trace_geom *__restrict__ sgeom)
{
    const int p = blockIdx.x * blockDim.x + threadIdx.x;
    if (p >= trow) return;
    for (int64_t k = blockIdx.y; k < nbin; k += gridDim.y) {
        const int32_t ti = order[k];
        taps[k * trow + p] = (p >= 1 && p <= ns) ? prep[(int64_t)ti * row + p - 1] : 0.0f;
        if (p == 0) sgeom[k] = geom[ti];
    }
}

/* One pass over the traces of class h for this lane's sample t.  The
 * traveltime is formed as q = t + m/dt with the moveout
 *   m = a r / (s1 + th) + b r / (s2 + th),  s1 = sqrt(th^2 + a r),  th = t dt / 2,
 * (equal to (s1 + s2)/dt - t), so FP32 rounds the moveout and not the whole
 * traveltime: ulp(m) instead of ulp(q) (1e-4 samples at q ~ 2000).
 * FWD: returns the Kahan sum of F(q); otherwise sum F'(q) (a/s1 + b/s2).
 * lo..hi are positions in the binned order, so they index sgeom and taps
 * directly.  The interval T = t + floor(min(m/dt, ns + 2 - t)) stays in
 * [0, ns + 2] for every lane (m >= 0; inactive lanes have r = 0). */
template <bool FWD>
__device__ __forceinline__ float vg_class_pass(const VgArgs &a, int lo, int hi, float IX,
                                               float r, float th, float t2, int t, int lane,
                                               float2 *sAB)
{
    const float dmax = (float)(a.ns + 2 - t);   /* q clamped to ns + 2: a zero interval */
    float acc = 0.0f, cmp = 0.0f;
    for (int b0 = lo; b0 < hi; b0 += VG_WAVE) {
        const int k = b0 + lane;
        float ga = 0.0f, gb = 0.0f;
        bool pass = false;
        if (k < hi) {
            const ktm_trace_geom g = a.sgeom[k];
            const float ax = IX - g.sxt, ay = a.coly - g.syt;
            const float bx = IX - g.gxt, by = a.coly - g.gyt;
            const float cx = IX - g.cmpx, cy = a.coly - g.cmpy;
            ga = ax * ax + ay * ay;
            gb = bx * bx + by * by;
            pass = cx * cx + cy * cy < a.ap_x2thr;
        }
        /* the squared distances of the wave's 64 traces in its LDS row: a
         * passing trace's pair is one broadcast read below, not two
         * v_readlane per contribution (the wave's own LDS accesses are
         * ordered, so the row is rewritten only after the reads above it) */
        sAB[lane] = make_float2(ga, gb);
        uint64_t m = __ballot(pass);
        while (m) {
            /* up to VG_G passing traces at a time, in order; slots past the
             * last passing trace repeat the first and are not added */
            int jc[VG_G], nb = 0;
#pragma unroll
            for (int c = 0; c < VG_G; ++c) {
                if (m) {
                    jc[c] = __builtin_ctzll(m);
                    m &= m - 1ull;
                    nb = c + 1;
                } else {
                    jc[c] = jc[0];
                }
            }
            vg_f4 cf[VG_G];
            float uu[VG_G], dqv[VG_G];
#pragma unroll
            for (int c = 0; c < VG_G; ++c) {
                const int j = jc[c];
                const float2 ab = sAB[j];
                const float A = ab.x, B = ab.y;
                const char *Fr = (const char *)(a.taps + (int64_t)(b0 + j) * a.trow);
                const float ar = A * r, br = B * r;
                /* t2 carries +1e-30 (see k_vgrad): u >= 1e-30 keeps s + th
                 * >= 1e-15 > 0 without a guard per contribution */
                const float u1 = ar + t2, u2 = br + t2;
                const float i1 = __builtin_amdgcn_rsqf(u1), i2 = __builtin_amdgcn_rsqf(u2);
                const float w1 = __builtin_amdgcn_rcpf(__builtin_fmaf(u1, i1, th));
                const float w2 = __builtin_amdgcn_rcpf(__builtin_fmaf(u2, i2, th));
                const float d = fminf(__builtin_fmaf(ar, w1, br * w2) * a.rdt, dmax);
                const float Df = floorf(d);
                uu[c] = d - Df;
                dqv[c] = __builtin_fmaf(A, i1, B * i2);
                /* taps T-1..T+2 at a 32-bit byte offset from the trace's
                 * (scalar) row base */
                cf[c] = *(const vg_f4 *)(Fr + (uint32_t)(t + (int)Df) * 4u);
            }
#pragma unroll
            for (int c = 0; c < VG_G; ++c) {
                if (c < nb) {
                    if (FWD) {
                        const float y = cr_eval<false>(cf[c], uu[c]) - cmp;
                        const float s = acc + y;
                        cmp = (s - acc) - y;
                        acc = s;
                    } else {
                        acc = __builtin_fmaf(cr_eval<true>(cf[c], uu[c]), dqv[c], acc);
                    }
                }
            }
        }
    }
    return acc;
}

__global__ void __launch_bounds__(VG_NT) k_vgrad(const VgArgs a)
{
    __shared__ double sRed[VG_NT / VG_WAVE];
    __shared__ float2 sABw[VG_NT / VG_WAVE][VG_WAVE];   /* per wave: the gate lanes' distances */
    const int tid = threadIdx.x, lane = tid & (VG_WAVE - 1), wv = tid / VG_WAVE;
    /* XCD-aware order: consecutive workgroup ids go to different XCDs; each
     * XCD takes runs of a.run consecutive (column, chunk) tiles (neighbouring
     * columns share traces in its L2), dealt round-robin so that every XCD
     * gets columns from the whole line (fewer traces reach its ends) */
    const int ntiles = a.nxp * a.nchunks;
    const int wi = blockIdx.x >> 3;
    const int tile = (((wi / a.run) << 3) + (int)(blockIdx.x & 7)) * a.run + wi % a.run;
    if (tile >= ntiles) return;
    const int X = tile / a.nchunks, ch = tile % a.nchunks;
    const int t = ch * a.chunk + tid;
    const bool act = tid < a.chunk && t < a.ns;
    const float IX = a.colx[X];
    float r = 0.0f, th = 0.0f, t2 = 0.0f, kfac = 0.0f;
    if (act) {
        const float v = a.v[(int64_t)X * a.ns + t];
        r = 1.0f / (v * v);
        th = 0.5f * (float)t * a.dt;
        t2 = th * th;
        kfac = -(r / v) / a.dt;     /* dq/dv = kfac (a/s1 + b/s2) */
    }
    /* + 1e-30 replaces a guard u = max(a r + t2, 1e-30) per contribution
     * (u = 0, at t = 0 with a source or receiver on the column, would give
     * s + th = 0): for t >= 1 (t2 >= 4e-6) t2 is unchanged, and at t = 0
     * u = a r + 1e-30 equals the guarded value for a r = 0 and for
     * a r >= 3.4e-23 (1e-30 below half its ulp) -- every case but a source
     * or receiver within ~6e-8 m of the column and not on it, where u moves
     * by at most 1e-30 */
    t2 = t2 + 1.0e-30f;
    const int32_t *rg = a.ranges + (int64_t)X * a.nh * 2;

    /* The offset classes stream in order: the lane holds I[h-1..h+1] of its
     * sample, the objective share adds (I[h+1] - I[h])^2, and as soon as
     * dJ/dI[h] = 2 ((I[h] - I[h-1]) - (I[h+1] - I[h])) is known the second
     * pass recomputes class h's traveltimes for dJ/dv.  No per-class array,
     * so no LDS bounds the occupancy. */
    double jl = 0.0;
    float dprev = 0.0f, g = 0.0f;
    float Ih = vg_class_pass<true>(a, rg[0], rg[1], IX, r, th, t2, t, lane, sABw[wv]);
    if (!act) Ih = 0.0f;
    for (int h = 0; h < a.nh; ++h) {
        float In = 0.0f, dn = 0.0f;
        if (h + 1 < a.nh) {
            In = vg_class_pass<true>(a, rg[2 * h + 2], rg[2 * h + 3], IX, r, th, t2, t, lane, sABw[wv]);
            if (!act) In = 0.0f;
            dn = In - Ih;
        }
        jl += (double)dn * (double)dn;
        if (a.cig && act) a.cig[((int64_t)X * a.nh + h) * a.ns + t] = Ih;
        if (a.gv) {
            const float rh = 2.0f * (dprev - dn);
            const float S = vg_class_pass<false>(a, rg[2 * h], rg[2 * h + 1], IX, r, th, t2, t, lane, sABw[wv]);
            g = __builtin_fmaf(rh, S, g);
        }
        dprev = dn;
        Ih = In;
    }
    for (int o = VG_WAVE / 2; o > 0; o >>= 1) jl += __shfl_down(jl, o);
    if (lane == 0) sRed[wv] = jl;
    __syncthreads();
    if (tid == 0) {
        double s = 0.0;
        for (int w = 0; w < VG_NT / VG_WAVE; ++w) s += sRed[w];
        a.part[tile] = s;
    }
    if (a.gv && act) a.gv[(int64_t)X * a.ns + t] = kfac * g;
}

/* ------------------------------------------------------------------------ */
/* host side                                                                 */
/* ------------------------------------------------------------------------ */
extern "C" int ktm_vgrad_destroy(ktm_vgrad *p)
{
    if (!p) return KTM_OK;
    (void)hipSetDevice(p->device);
    void *bufs[] = {p->d_colx, p->d_v, p->d_gv, p->d_part, p->d_G1, p->d_flag,
                    p->d_order, p->d_ranges, p->d_sgeom, p->d_taps};
    for (void *b : bufs)
        if (b) (void)hipFree(b);
    free(p->h_part);
    ktm_tables_free(&p->tb);
    free(p);
    return KTM_OK;
}

extern "C" int ktm_vgrad_create(ktm_vgrad **out, const ktm_config *cfg, int ns, float dt,
                                int nt_nodes, int nx_nodes, int chunk)
{
    if (!out || !cfg) return KTM_ERR_ARG;
    *out = NULL;
    if (ns < 4 || nt_nodes < 4 || nx_nodes < 4 || chunk < 1 || chunk > VG_NT) return KTM_ERR_ARG;
    ktm_vgrad *p = (ktm_vgrad *)calloc(1, sizeof *p);
    if (!p) return KTM_ERR_NOMEM;
    int st = ktm_tables_build(cfg, ns, dt, ns, dt, &p->tb);
    if (st != KTM_OK) { free(p); return st; }
    const ktm_setup *s = &p->tb.s;
    if (s->nyp != 1 || s->nxp < 2) { ktm_vgrad_destroy(p); return KTM_ERR_UNSUPPORTED; }
    p->cfg = *cfg;
    p->ns = ns; p->dt = dt; p->nxp = s->nxp; p->nh = s->nh;
    p->nt_nodes = nt_nodes; p->nx_nodes = nx_nodes; p->chunk = chunk;
    p->nchunks = (ns + chunk - 1) / chunk;
    p->dtn = (double)(ns - 1) / (nt_nodes - 3);
    p->dxn = (double)(p->nxp - 1) / (nx_nodes - 3);
    if (hipGetDevice(&p->device) != hipSuccess) { ktm_vgrad_destroy(p); return KTM_ERR_DEVICE; }
    const size_t img = (size_t)p->nxp * ns;
    const size_t nparts = (size_t)p->nxp * p->nchunks;
    p->h_part = (double *)malloc(sizeof(double) * nparts);
    if (!p->h_part ||
        hipMalloc(&p->d_colx, sizeof(float) * p->nxp) != hipSuccess ||
        hipMalloc(&p->d_v, sizeof(float) * img) != hipSuccess ||
        hipMalloc(&p->d_gv, sizeof(float) * img) != hipSuccess ||
        hipMalloc(&p->d_part, sizeof(double) * nparts) != hipSuccess ||
        hipMalloc(&p->d_G1, sizeof(double) * (size_t)p->nxp * nt_nodes) != hipSuccess ||
        hipMalloc(&p->d_flag, sizeof(int)) != hipSuccess ||
        hipMalloc(&p->d_ranges, sizeof(int32_t) * 2 * (size_t)p->nxp * p->nh) != hipSuccess) {
        ktm_vgrad_destroy(p);
        return KTM_ERR_NOMEM;
    }
    if (hipMemcpy(p->d_colx, p->tb.colx, sizeof(float) * p->nxp, hipMemcpyHostToDevice) != hipSuccess) {
        ktm_vgrad_destroy(p);
        return KTM_ERR_DEVICE;
    }
    *out = p;
    return KTM_OK;
}

typedef struct { int32_t off; float cmpx; int32_t idx; } vg_key;

static int vg_key_cmp(const void *x, const void *y)
{
    const vg_key *a = (const vg_key *)x, *b = (const vg_key *)y;
    if (a->off != b->off) return a->off < b->off ? -1 : 1;
    if (a->cmpx != b->cmpx) return a->cmpx < b->cmpx ? -1 : 1;
    return a->idx < b->idx ? -1 : (a->idx > b->idx);
}

/* first key in [lo, hi) with cmpx >= x (lower) or > x (upper) */
static int64_t vg_bound(const vg_key *k, int64_t lo, int64_t hi, float x, int upper)
{
    while (lo < hi) {
        const int64_t mid = lo + (hi - lo) / 2;
        const int go_right = upper ? !(k[mid].cmpx > x) : (k[mid].cmpx < x);
        if (go_right) lo = mid + 1; else hi = mid;
    }
    return lo;
}

extern "C" int ktm_vgrad_bind(ktm_vgrad *p, const float *prepared, int64_t row,
                              const ktm_trace_geom *geom, int64_t ntr, void *stream)
{
    if (!p || !prepared || !geom || ntr <= 0 || ntr > INT32_MAX || row < p->ns + 1)
        return KTM_ERR_ARG;
    hipStream_t st = (hipStream_t)stream;
    VG_TRY(hipSetDevice(p->device));
    p->bound = 0;
    ktm_trace_geom *hg = (ktm_trace_geom *)malloc(sizeof(ktm_trace_geom) * (size_t)ntr);
    vg_key *keys = (vg_key *)malloc(sizeof(vg_key) * (size_t)ntr);
    int32_t *ranges = (int32_t *)malloc(sizeof(int32_t) * 2 * (size_t)p->nxp * p->nh);
    int32_t *order = (int32_t *)malloc(sizeof(int32_t) * (size_t)ntr);
    int64_t *cstart = (int64_t *)calloc((size_t)p->nh + 1, sizeof(int64_t));
    int rc = KTM_OK;
    if (!hg || !keys || !ranges || !order || !cstart) { rc = KTM_ERR_NOMEM; goto done; }
    if (hipMemcpyAsync(hg, geom, sizeof(ktm_trace_geom) * (size_t)ntr, hipMemcpyDeviceToHost, st) !=
            hipSuccess ||
        hipStreamSynchronize(st) != hipSuccess) {
        rc = KTM_ERR_DEVICE;
        goto done;
    }
    {
        int64_t n = 0;
        for (int64_t i = 0; i < ntr; ++i) {
            if (hg[i].off < 0 || hg[i].off >= p->nh) continue;   /* Kirchhoff.c:1311 */
            keys[n].off = hg[i].off;
            keys[n].cmpx = hg[i].cmpx;
            keys[n].idx = (int32_t)i;
            ++n;
        }
        qsort(keys, (size_t)n, sizeof(vg_key), vg_key_cmp);
        int64_t pairs = 0;
        for (int64_t k = 0; k < n; ++k) order[k] = keys[k].idx;
        for (int64_t k = 0, h = 0; h <= p->nh; ++h) {
            while (k < n && keys[k].off < h) ++k;
            cstart[h] = k;
        }
        cstart[p->nh] = n;
        /* a superset of the traces whose CMP passes the aperture gate of a
         * column: |cmpx - IX| <= distance < aperture; the exact gate runs in
         * the kernel */
        const float reach = (float)(p->cfg.aperture * (1.0 + 1e-6) + 1.0);
        for (int X = 0; X < p->nxp; ++X) {
            const float IX = p->tb.colx[X];
            for (int h = 0; h < p->nh; ++h) {
                const int64_t lo = vg_bound(keys, cstart[h], cstart[h + 1], IX - reach, 0);
                const int64_t hi = vg_bound(keys, lo, cstart[h + 1], IX + reach, 1);
                ranges[((int64_t)X * p->nh + h) * 2] = (int32_t)lo;
                ranges[((int64_t)X * p->nh + h) * 2 + 1] = (int32_t)hi;
                /* the kernel's gate, same float arithmetic (no contraction) */
                for (int64_t k = lo; k < hi; ++k) {
                    const ktm_trace_geom *g = hg + keys[k].idx;
                    const float cx = IX - g->cmpx, cy = p->tb.coly[0] - g->cmpy;
                    pairs += cx * cx + cy * cy < p->tb.ap_x2thr;
                }
            }
        }
        if (n > p->order_cap) {
            (void)hipFree(p->d_order);
            p->d_order = NULL;
            p->order_cap = 0;
            if (hipMalloc(&p->d_order, sizeof(int32_t) * (size_t)(n > 0 ? n : 1)) != hipSuccess) {
                rc = KTM_ERR_NOMEM;
                goto done;
            }
            p->order_cap = n;
        }
        const int64_t trow = (int64_t)p->ns + 6;
        if (n > p->taps_cap) {
            (void)hipFree(p->d_sgeom);
            (void)hipFree(p->d_taps);
            p->d_sgeom = NULL;
            p->d_taps = NULL;
            p->taps_cap = 0;
            if (hipMalloc(&p->d_sgeom, sizeof(ktm_trace_geom) * (size_t)n) != hipSuccess ||
                hipMalloc(&p->d_taps, sizeof(float) * (size_t)n * (size_t)trow) != hipSuccess) {
                rc = KTM_ERR_NOMEM;
                goto done;
            }
            p->taps_cap = n;
        }
        if ((n > 0 && hipMemcpy(p->d_order, order, sizeof(int32_t) * (size_t)n,
                                hipMemcpyHostToDevice) != hipSuccess) ||
            hipMemcpy(p->d_ranges, ranges, sizeof(int32_t) * 2 * (size_t)p->nxp * p->nh,
                      hipMemcpyHostToDevice) != hipSuccess) {
            rc = KTM_ERR_DEVICE;
            goto done;
        }
        if (n > 0) {
            const dim3 grid((unsigned)((trow + 255) / 256), (unsigned)(n < 65535 ? n : 65535));
            hipLaunchKernelGGL(k_vg_tables, grid, dim3(256), 0, st, prepared, row, geom,
                               p->d_order, n, p->ns, trow, p->d_taps, p->d_sgeom);
        }
        if (hipGetLastError() != hipSuccess || hipStreamSynchronize(st) != hipSuccess) {
            rc = KTM_ERR_DEVICE;
            goto done;
        }
        p->ntr = ntr;
        p->pairs = pairs;
        p->bound = 1;
    }
done:
    free(hg); free(keys); free(ranges); free(order); free(cstart);
    return rc;
}

extern "C" int ktm_vgrad_velocity(ktm_vgrad *p, const float *V, float *v, void *stream)
{
    if (!p || !V || !v) return KTM_ERR_ARG;
    hipStream_t st = (hipStream_t)stream;
    VG_TRY(hipSetDevice(p->device));
    const int64_t n = (int64_t)p->nxp * p->ns;
    int zero = 0, flag = 0;
    VG_TRY(hipMemcpyAsync(p->d_flag, &zero, sizeof(int), hipMemcpyHostToDevice, st));
    hipLaunchKernelGGL(k_vg_velocity, dim3((unsigned)((n + 255) / 256)), dim3(256), 0, st, V,
                       p->nt_nodes, p->nx_nodes, p->dtn, p->dxn, p->ns, p->nxp, v, p->d_flag);
    VG_TRY(hipGetLastError());
    VG_TRY(hipMemcpyAsync(&flag, p->d_flag, sizeof(int), hipMemcpyDeviceToHost, st));
    VG_TRY(hipStreamSynchronize(st));
    return flag ? KTM_ERR_VELOCITY : KTM_OK;
}

extern "C" int ktm_vgrad_eval(ktm_vgrad *p, const float *V, double *J, double *dJdV, float *cig,
                              void *stream)
{
    if (!p || !V || !J) return KTM_ERR_ARG;
    if (!p->bound) return KTM_ERR_ARG;
    hipStream_t st = (hipStream_t)stream;
    int rc = ktm_vgrad_velocity(p, V, p->d_v, stream);
    if (rc != KTM_OK) return rc;
    VgArgs a;
    a.taps = p->d_taps; a.trow = (int64_t)p->ns + 6; a.sgeom = p->d_sgeom;
    a.ranges = p->d_ranges;
    a.v = p->d_v; a.colx = p->d_colx;
    a.coly = p->tb.coly[0]; a.ap_x2thr = p->tb.ap_x2thr;
    a.dt = p->dt; a.rdt = 1.0f / p->dt;
    a.ns = p->ns; a.nh = p->nh; a.nxp = p->nxp; a.chunk = p->chunk; a.nchunks = p->nchunks;
    a.cig = cig; a.gv = dJdV ? p->d_gv : NULL; a.part = p->d_part;
    const int ntiles = p->nxp * p->nchunks;
    /* runs of 16 columns, but at least 8 runs per XCD */
    a.run = ntiles / 64 < 16 * p->nchunks ? ntiles / 64 : 16 * p->nchunks;
    if (a.run < 1) a.run = 1;
    const int64_t span = 8 * (int64_t)a.run;
    const int64_t nblk = ((int64_t)ntiles + span - 1) / span * span;
    hipLaunchKernelGGL(k_vgrad, dim3((unsigned)nblk), dim3(VG_NT), 0, st, a);
    VG_TRY(hipGetLastError());
    if (dJdV) {
        const int64_t n1 = (int64_t)p->nxp * p->nt_nodes;
        hipLaunchKernelGGL(k_vg_project_t, dim3((unsigned)((n1 + 255) / 256)), dim3(256), 0, st,
                           p->d_gv, p->ns, p->nxp, p->nt_nodes, p->dtn, p->d_G1);
        const int n2 = p->nt_nodes * p->nx_nodes;
        hipLaunchKernelGGL(k_vg_project_x, dim3((unsigned)((n2 + 255) / 256)), dim3(256), 0, st,
                           p->d_G1, p->nxp, p->nt_nodes, p->nx_nodes, p->dxn, dJdV);
        VG_TRY(hipGetLastError());
    }
    const size_t nparts = (size_t)ntiles;
    VG_TRY(hipMemcpyAsync(p->h_part, p->d_part, sizeof(double) * nparts, hipMemcpyDeviceToHost, st));
    VG_TRY(hipStreamSynchronize(st));
    double s = 0.0;
    for (size_t k = 0; k < nparts; ++k) s += p->h_part[k];
    *J = s;
    return KTM_OK;
}

extern "C" int64_t ktm_vgrad_pairs(const ktm_vgrad *p)
{
    return p && p->bound ? p->pairs : -1;
}
