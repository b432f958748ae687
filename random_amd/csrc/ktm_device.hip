/*
 * ktm_device.hip -- gfx950 kernels and the device half of the C ABI of the
 * Kirchhoff prestack time-migration engine.
 *
 * Kernels (every one replaces a loop of Kirchhoff.c's kirchhoff()):
 *   k_geometry      per-trace rotation / CMP / offset class   Kirchhoff.c:1296-1310
 *                   (k_geometry_segy: from the raw trace headers)
 *   k_segy_decode   raw SEG-Y records -> FP32 traces + header words,
 *                   Kirchhoff.c:1281-1290 (the prep kernels below decode
 *                   records themselves when ingesting a raw block)
 *   k_prep_ref      FP32 double integral + trange=2 taps       Kirchhoff.c:1312-1328,1391-1392
 *   k_prep_closed   same operator as a 3-tap filter (FP32), from FP32 or
 *                   BF16/FP16-stored traces              (exact rewrite of 1312-1392)
 *   k_quantize      BF16/FP16 storage with per-trace scale and sixth-order
 *                   error feedback
 *   k_migrate       the image loop, Kirchhoff.c:1332-1404: one workgroup owns a
 *                   tile of image columns (one wave per column, all ns samples
 *                   in registers), streams the traces of a block through LDS in
 *                   trace order and accumulates F(floor(dsr/dt)) per sample
 *   k_postfft<PostKtm>  derivative/fold + low-pass, Kirchhoff.c:1457-1479, as
 *                   an FP64 FFT convolution (ktm_postfft.h; beyond 4096
 *                   samples the FP64 matrix product of ktm_postmx.h); its
 *                   transpose k_postfft<PostImg> + k_poststack_adjoint_tail
 *   k_image_*       image<job>.su trace records (headers + big-endian
 *                   samples), Kirchhoff.c:1437-1598
 *   k_demigrate     exact transpose of k_migrate (CLOSED prep)
 *   k_synth_*       seeded synthetic surveys
 *
 * Build flags: -ffp-contract=off (no implicit FMA: every product and sum
 * that the reference rounds separately is rounded separately here; FMAs are
 * written explicitly where intended) and the default correctly-rounded FP32
 * divide/sqrt; packed-FP32 ops are disabled.
 */
#include <hip/hip_runtime.h>
#include <rocprim/device/device_radix_sort.hpp>

#include <mutex>

#include <stdint.h>
#include <stdlib.h>
#include <string.h>

#include "ktm_internal.h"
#include "ktm_devmath.h"
#include "ktm_postmx.h"
#include "ktm_postfft.h"

#define KTM_WAVE 64

/* ------------------------------------------------------------------------ */
/* plan                                                                      */
/* ------------------------------------------------------------------------ */
struct ktm_plan {
    ktm_config cfg;
    ktm_tables tb;
    int device;
    int64_t row;            /* prepared trace row (elements), multiple of 8, >= ns+1 */
    float *d_t2;
    int32_t *d_t1;
    float *d_colx, *d_coly;
    int32_t *d_velx, *d_vely;
    float *d_slab_cx, *d_slab_cy;
    float *d_slab_x2thr;
    float *d_vel2;          /* [nvel_il][nvel_xl][nsm], squared like Kirchhoff.c:617 */
    float *d_w;             /* [ns+2] CLOSED weights */
    double *d_hconv2;       /* [2 ns] post-stack kernel, h2[m] = h[m mod ns] */
    double2 *d_pf_g, *d_pf_tw; /* [N] FFT form of the post-stack kernel and twiddles */
    int pf_n;                /* points of the FFT form, 0: the direct product */
    float *d_stage;         /* FP32 filtered traces of a BF16/FP16 block */
    size_t stage_cap;       /* elements of d_stage */
    float *d_pix_v2, *d_pix_rs; /* per-pixel v^2 and RN(1/(v^2 dt^2)), built for demigration */
    /* demigration trace order: sort keys and indices (in, out) and the
     * radix sort's temporary storage, grown on demand */
    uint64_t *d_dm_keys;
    int32_t *d_dm_idx;
    void *d_dm_tmp;
    size_t dm_cap, dm_tmp_bytes;
    uint32_t *d_dm_amax;    /* bits of max |km| of the last demigration */
    float *d_scratch;       /* this device's copy of `scratch` (ktm_migrate_block_static) */
    /* compensated migration of small images: partial images of parts 1..
     * parts_max-1 of a trace block, [part][km, fold, comp][image or region] */
    float *d_parts;
    int parts_max;
};

static int scratch_address(int dev, float **out);
static int parts_workspace(ktm_plan *p);

#define HIP_TRY(expr)                                  \
    do {                                               \
        hipError_t e_ = (expr);                        \
        if (e_ != hipSuccess) return KTM_ERR_DEVICE;   \
    } while (0)

/* ------------------------------------------------------------------------ */
/* per-trace geometry, Kirchhoff.c:1296-1310 (and Kirchhoff_GHTest1.c:       */
/* 1302-1315 when scale_xy > 0)                                              */
/* ------------------------------------------------------------------------ */
struct GeoArgs {
    double ox, oy, scale, oh, dh;
    float sinA, cosA;
};

__device__ __forceinline__ void rotate_point(const GeoArgs &g, uint32_t hx, uint32_t hy,
                                             float &xt, float &yt)
{
    if (g.scale == 0.0) {
        /* htonl() gives uint32_t; minus the double origin; into an int */
        const int i = (int)((double)hx - g.ox);
        const int j = (int)((double)hy - g.oy);
        xt = g.cosA * (float)i + g.sinA * (float)j;
        yt = -g.sinA * (float)i + g.cosA * (float)j;
    } else {
        const int i = (int)((double)hx - g.ox / g.scale);
        const int j = (int)((double)hy - g.oy / g.scale);
        xt = (float)((double)(g.cosA * (float)i) * g.scale + (double)(g.sinA * (float)j) * g.scale);
        yt = (float)((double)(-g.sinA * (float)i) * g.scale + (double)(g.cosA * (float)j) * g.scale);
    }
}

__device__ __forceinline__ ktm_trace_geom trace_geometry(const GeoArgs &g, const int32_t *c)
{
    ktm_trace_geom r;
    rotate_point(g, (uint32_t)c[0], (uint32_t)c[1], r.sxt, r.syt);
    rotate_point(g, (uint32_t)c[2], (uint32_t)c[3], r.gxt, r.gyt);
    const float ex = r.sxt - r.gxt, ey = r.syt - r.gyt;
    const float h2 = (float)(0.25 * (double)(ex * ex + ey * ey));
    r.cmpx = (float)(0.5 * (double)(r.sxt + r.gxt));
    r.cmpy = (float)(0.5 * (double)(r.syt + r.gyt));
    r.off = (int)floor((2.0 * sqrt((double)h2) - g.oh) / g.dh);
    r.scale = 1.0f;
    return r;
}

__global__ void k_geometry(GeoArgs g, const int32_t *__restrict__ coords, int64_t ntr,
                           ktm_trace_geom *__restrict__ geom)
{
    const int64_t i = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
    if (i >= ntr) return;
    geom[i] = trace_geometry(g, coords + 4 * i);
}

/* The same from raw SEG-Y trace records: header words 18-21 (sx, sy, gx,
 * gy; bytes 73-88, big-endian, Kirchhoff.c:1297,1302) of each record of
 * rec_words 4-byte words; the header words also go to coords if given. */
__global__ void k_geometry_segy(GeoArgs g, const uint32_t *__restrict__ raw, int64_t rec_words,
                                int64_t ntr, ktm_trace_geom *__restrict__ geom,
                                int32_t *__restrict__ coords)
{
    const int64_t i = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
    if (i >= ntr) return;
    int32_t c[4];
#pragma unroll
    for (int k = 0; k < 4; ++k) c[k] = (int32_t)__builtin_bswap32(raw[i * rec_words + 18 + k]);
    if (coords) {
#pragma unroll
        for (int k = 0; k < 4; ++k) coords[4 * i + k] = c[k];
    }
    geom[i] = trace_geometry(g, c);
}

/* ------------------------------------------------------------------------ */
/* trace preparation                                                         */
/* ------------------------------------------------------------------------ */

/* Where the prep kernels read trace samples: FP32 traces [ntr][ns], or raw
 * SEG-Y trace records as stored in the file (240-byte header + ns big-endian
 * words, Kirchhoff.c:1281-1290) with IBM (format 1) or IEEE (format 5)
 * samples.  Decoding inside the prep kernels makes the ingest of a block
 * one pass: the byte swap and IBM decode of Kirchhoff.c:1284-1290 happen
 * where the sample is consumed. */
enum { SRC_F32 = 0, SRC_SEGY_IBM = 1, SRC_SEGY_IEEE = 2, SRC_Q_BF16 = 3, SRC_Q_F16 = 4 };

struct TraceSrc {
    const void *p;
    int64_t rec_words;      /* 4-byte words per trace: ns, or 60 + ns for a record */
};

/* IBM hex float -> float as Kirchhoff.c:1287-1289 evaluates it:
 * (sign * fraction / 2^24) * (float)16^(exponent - 64), the power of 16
 * rounded to float on its own (inf above 2^127, 0 below 2^-149; a zero
 * fraction times inf is NaN).  The power is built from its bits: e4 is a
 * multiple of 4, so no value falls between two floats. */
__device__ __forceinline__ float ibm_word(uint32_t w)
{
    const float frac = (float)(w & 0x00ffffffu) * (1.0f / 16777216.0f);
    const int e4 = 4 * (int)((w >> 24) & 0x7fu) - 256;
    /* all three forms computed, then selected (no divergent branch) */
    const uint32_t norm = (uint32_t)(e4 + 127) << 23;                 /* -126 <= e4 <= 127 */
    const int sh = e4 + 149;                                          /* -149 <= e4 < -126 */
    const uint32_t sub = sh >= 0 ? (1u << (sh & 31)) : 0u;
    const uint32_t pb = e4 > 127 ? 0x7f800000u : (e4 >= -126 ? norm : sub);
    const float sf = (w >> 31) ? -frac : frac;
    return sf * __uint_as_float(pb);
}

/* Raw 32-bit word of sample t of trace i (no decode yet), and its decode:
 * split so that a tile's loads can all be in flight before any decode. */
template <int SRC>
__device__ __forceinline__ uint32_t load_word(const TraceSrc &s, int64_t i, int t)
{
    if (SRC == SRC_F32) return reinterpret_cast<const uint32_t *>(s.p)[i * s.rec_words + t];
    return reinterpret_cast<const uint32_t *>(s.p)[i * s.rec_words + 60 + t];
}

template <int SRC>
__device__ __forceinline__ float decode_word(uint32_t u)
{
    if (SRC == SRC_F32) return __uint_as_float(u);
    const uint32_t w = __builtin_bswap32(u);
    return SRC == SRC_SEGY_IBM ? ibm_word(w) : __uint_as_float(w);
}

template <int SRC>
__device__ __forceinline__ float load_sample(const TraceSrc &s, int64_t i, int t)
{
    return decode_word<SRC>(load_word<SRC>(s, i, t));
}

/* REFERENCE mode: one lane per trace runs the two FP32 running sums of
 * Kirchhoff.c:1314-1324 in a single pass, keeping the last five values of
 * the double integral p[] in registers, and emits
 *   F[T] = (float)(-2.0*p[T] + p[T-2] + p[T+2])         (Kirchhoff.c:1392)
 * as soon as p[T+2] exists (F[t-2] at step t).  p[-1] = p[1] and p[-2] =
 * p[2] are the mirror pad of Kirchhoff.c:1327-1328.  A workgroup is one
 * wave of 64 traces; samples pass through LDS in tiles of 64 so that every
 * global load and store is one contiguous 256-byte run of a trace row (a
 * lane walking its own row would touch 64 rows per instruction). */
#define PREP_TS 32
template <int SRC>
__global__ void __launch_bounds__(KTM_WAVE)
k_prep_ref(const TraceSrc src, int64_t ntr, int ns, int64_t row, float dt,
           const float *__restrict__ t2, float *__restrict__ out)
{
    /* s_inc[j][k]: the increment of integral 1 at t = t0 + k of trace j,
     * (double)tr[t] / (4 t2[t]) (Kirchhoff.c:1314); at t = 0 the sample
     * itself.  It does not depend on the running sum, so it is computed in
     * the load phase, lane-parallel, and the sequential recurrence is left
     * with one double add and two conversions per step.  Loads and stores
     * move two traces x 32 samples per instruction (128-byte runs). */
    __shared__ double s_inc[KTM_WAVE][PREP_TS + 1];
    __shared__ float s_out[KTM_WAVE][PREP_TS + 1];
    const int lane = threadIdx.x;
    const int sub = lane & (PREP_TS - 1), half = lane / PREP_TS;
    const int64_t tr0 = (int64_t)blockIdx.x * KTM_WAVE;
    const int ntl = (int)min((int64_t)KTM_WAVE, ntr - tr0);
    const int64_t i = tr0 + (lane < ntl ? lane : 0);
    const float last = load_sample<SRC>(src, i, ns - 1);
    float q = 0.0f, p = 0.0f;
    float pm1 = 0.0f, pm2 = 0.0f, pm3 = 0.0f, pm4 = 0.0f;
    /* unconditional loads at clamped (in-bounds) positions, so that all 32
     * are in flight before the first wait; out-of-range values are replaced
     * after the decode.  The next tile's words are loaded as soon as this
     * tile's are decoded, so their latency overlaps the recurrence. */
    uint32_t wv[KTM_WAVE / 2];
    float t2v;
    {
        const int tc = sub < ns ? sub : ns - 1;
#pragma unroll
        for (int m = 0; m < KTM_WAVE / 2; ++m)
            wv[m] = load_word<SRC>(src, tr0 + min(2 * m + half, ntl - 1), tc);
        t2v = t2[tc];
    }
    for (int t0 = 0; t0 < ns + 2; t0 += PREP_TS) {
        const int t = t0 + sub;
        const double den = t < ns && t > 0 ? 4.0 * (double)t2v : 1.0;
        float xv[KTM_WAVE / 2];
#pragma unroll
        for (int m = 0; m < KTM_WAVE / 2; ++m)
            xv[m] = t < ns ? decode_word<SRC>(wv[m]) : 0.0f;
#pragma unroll
        for (int m = 0; m < KTM_WAVE / 2; ++m)
            s_inc[2 * m + half][sub] = t == 0 ? (double)xv[m] : (double)xv[m] / den;
        if (t0 + PREP_TS < ns + 2) {
            const int tn = t0 + PREP_TS + sub;
            const int tc = tn < ns ? tn : ns - 1;
#pragma unroll
            for (int m = 0; m < KTM_WAVE / 2; ++m)
                wv[m] = load_word<SRC>(src, tr0 + min(2 * m + half, ntl - 1), tc);
            t2v = t2[tc];
        }
        __syncthreads();
        const int s_end = min(PREP_TS, ns + 2 - t0);
        if (t0 >= 4 && t0 + PREP_TS <= ns) {
            /* interior tile: every step is the t >= 4, t < ns case; unrolled
             * so that the LDS reads and the F arithmetic overlap the
             * recurrence (one double add and two conversions per step) */
            double inc[PREP_TS];
#pragma unroll
            for (int k = 0; k < PREP_TS; ++k) inc[k] = s_inc[lane][k];
#pragma unroll
            for (int k = 0; k < PREP_TS; ++k) {
                q = (float)((double)q + inc[k]);
                pm4 = pm3; pm3 = pm2; pm2 = pm1; pm1 = p;
                p = p + q;
                s_out[lane][k] = (float)(-2.0 * (double)pm2 + (double)pm4 + (double)p);
            }
        } else for (int k = 0; k < s_end; ++k) {
            const int tk = t0 + k;
            const double inc = s_inc[lane][k];
            if (tk == 0) {
                q = (float)inc;   /* integral 1 at t */
                p = q;            /* integral 2 at t */
                continue;
            }
            if (tk < ns)
                q = (float)((double)q + inc);
            else
                q = q + last / dt / dt / (float)tk / (float)tk;
            pm4 = pm3; pm3 = pm2; pm2 = pm1; pm1 = p;
            p = p + q;
            /* window: pm4..p = p[t-4]..p[t]; F[t-2] goes to s_out[.][k] */
            if (tk == 2)
                s_out[lane][k] = (float)(-2.0 * (double)pm2 + (double)p + (double)p);
            else if (tk == 3)
                s_out[lane][k] = (float)(-2.0 * (double)pm2 + (double)pm2 + (double)p);
            else if (tk >= 4)
                s_out[lane][k] = (float)(-2.0 * (double)pm2 + (double)pm4 + (double)p);
        }
        __syncthreads();
        const int T = t0 - 2 + sub;           /* F index held in s_out[.][sub] */
        if (sub < s_end && T >= 0 && T < ns) {
#pragma unroll 8
            for (int m = 0; m < KTM_WAVE / 2; ++m) {
                const int j = 2 * m + half;
                if (j < ntl) out[(tr0 + j) * row + T] = s_out[j][sub];
            }
        }
        __syncthreads();
    }
    for (int j = 0; j < ntl; ++j)
        for (int64_t T = ns + lane; T < row; T += KTM_WAVE) out[(tr0 + j) * row + T] = 0.0f;
}

/* CLOSED mode: F as the 3-tap filter of x[t] = trace[min(t,ns-1)] * w[t].
 * Derivation from Kirchhoff.c:1314-1328 (exact arithmetic): with
 * Q[k] = sum_{j<=k} x[j] and p[k] = sum_{j<=k} Q[j],
 *   T >= 2 : p[T+2] - 2p[T] + p[T-2] = x[T] + 2x[T+1] + x[T+2]
 *   T == 1 : p[3] - p[1]           = 2x[0] + 2x[1] + 2x[2] + x[3]
 *   T == 0 : 2(p[2] - p[0])        = 4x[0] + 4x[1] + 2x[2]
 * Shared by the F32 prep kernel and the BF16/FP16 staging in k_migrate so the
 * two storage modes differ only by quantisation. */
__device__ __forceinline__ float bf16_to_f32_(uint16_t h) { return __uint_as_float((uint32_t)h << 16); }
__device__ __forceinline__ float f16_to_f32_(uint16_t u)
{
    _Float16 h;
    __builtin_memcpy(&h, &u, 2);
    return (float)h;
}

__device__ __forceinline__ float closed_tap(int T, float x0, float x1, float x2, float x3)
{
    /* for T >= 2, x0..x2 are x[T], x[T+1], x[T+2]; for T < 2 they are x[0..3] */
    if (T >= 2) return (x0 + 2.0f * x1) + x2;
    if (T == 1) return ((2.0f * x0 + 2.0f * x1) + 2.0f * x2) + x3;
    return (4.0f * x0 + 4.0f * x1) + 2.0f * x2;
}

/* src: any 32-bit trace source of load_sample, or BF16/FP16 traces stored
 * by k_quantize (SRC_Q_*, [ntr][row] halves, rec_words = row) with the
 * per-trace scale in geom; the stored value times the scale stands for the
 * trace sample. */
template <int SRC>
__global__ void k_prep_closed(const TraceSrc src, const ktm_trace_geom *__restrict__ geom,
                              int64_t ntr, int ns, int64_t row, const float *__restrict__ w,
                              float *__restrict__ out)
{
    constexpr bool Q16 = SRC == SRC_Q_BF16 || SRC == SRC_Q_F16;
    const int64_t i = blockIdx.y;
    const int T = blockIdx.x * blockDim.x + threadIdx.x;
    if (i >= ntr || T >= row) return;
    float v = 0.0f;
    if (T < ns) {
        float x[4];
        const int b = T >= 2 ? T : 0;
        const float sc = Q16 ? geom[i].scale : 1.0f;
#pragma unroll
        for (int k = 0; k < 4; ++k) {
            const int t = b + k;
            const int tt = t < ns ? t : ns - 1;
            float tr;
            if (Q16) {
                const uint16_t hq = reinterpret_cast<const uint16_t *>(src.p)[i * src.rec_words + tt];
                tr = sc * (SRC == SRC_Q_BF16 ? bf16_to_f32_(hq) : f16_to_f32_(hq));
            } else {
                tr = load_sample<SRC>(src, i, tt);
            }
            x[k] = t <= ns + 1 ? tr * w[t] : 0.0f;
        }
        v = closed_tap(T, x[0], x[1], x[2], x[3]);
    }
    out[i * row + T] = v;
}

/* BF16 (round to nearest even) and FP16 storage helpers */
__device__ __forceinline__ uint16_t f32_to_bf16(float f)
{
    uint32_t u = __float_as_uint(f);
    if ((u & 0x7fffffffu) > 0x7f800000u) return (uint16_t)((u >> 16) | 0x40u); /* NaN */
    u += 0x7fffu + ((u >> 16) & 1u);
    return (uint16_t)(u >> 16);
}
__device__ __forceinline__ float bf16_to_f32(uint16_t h) { return __uint_as_float((uint32_t)h << 16); }
__device__ __forceinline__ uint16_t f32_to_f16(float f)
{
    _Float16 h = (_Float16)f;
    uint16_t u;
    __builtin_memcpy(&u, &h, 2);
    return u;
}
__device__ __forceinline__ float f16_to_f32(uint16_t u)
{
    _Float16 h;
    __builtin_memcpy(&h, &u, 2);
    return (float)h;
}

/* scale = max |tr| (1 for an all-zero trace); the trace divided by scale is
 * rounded to BF16 or FP16 with sixth-order error feedback: sample t is
 * quantised as v = tr/scale + sum_k QA_k e[t-k] (e the rounding errors of
 * the previous samples), so the noise reaching the image is the rounding
 * noise shaped by N(z) = 1 - sum_k QA_k z^-k and then filtered by the [1,2,1]
 * anti-alias taps of the migration, H(z) = (1 + z^-1)^2.  QA minimise the
 * power of the shaped, filtered noise (|H|^2 has autocorrelation 6, 4, 1)
 * under N(1) = 0, which keeps the zero at DC of first-order feedback
 * (QA = 1): power 1.88 against 4 for first-order and 6 for none, the
 * low-frequency part 1.4 x first-order's where the unconstrained 4-tap
 * optimum has 4.4 x.  Measured on 74 random configurations: BF16 image
 * error (relative L2 against FP64) median 9.2e-4 -> 6.4e-4, max 1.6e-3 ->
 * 1.0e-3.  The coefficients are exact binary fractions and the feedback is
 * an FIR filter of bounded rounding errors, so it cannot run away.  A
 * workgroup is one wave of 64 traces: the max-abs sweep reads two traces x
 * 32 samples per load (max is exact in any order); the sequential
 * error-feedback sweep stages 32-sample tiles through LDS like k_prep_ref,
 * one lane per trace. */
#define QA1 1.625f
#define QA2 (-1.625f)
#define QA3 1.5f
#define QA4 (-1.0f)
#define QA5 0.625f
#define QA6 (-0.125f)
template <int DT, int SRC>
__global__ void __launch_bounds__(KTM_WAVE)
k_quantize(const TraceSrc src, int64_t ntr, int ns, int64_t row, uint16_t *__restrict__ q,
           ktm_trace_geom *__restrict__ geom)
{
    __shared__ float s_in[KTM_WAVE][PREP_TS + 1];
    __shared__ uint16_t s_q[KTM_WAVE][PREP_TS + 2];
    __shared__ float s_scale[KTM_WAVE];
    const int lane = threadIdx.x;
    const int sub = lane & (PREP_TS - 1), half = lane / PREP_TS;
    const int64_t tr0 = (int64_t)blockIdx.x * KTM_WAVE;
    const int ntl = (int)min((int64_t)KTM_WAVE, ntr - tr0);
    for (int m = 0; m < KTM_WAVE / 2; ++m) {
        const int j = min(2 * m + half, ntl - 1);
        float mx = 0.0f;
#pragma unroll 8
        for (int t = sub; t < ns; t += PREP_TS) mx = fmaxf(mx, fabsf(load_sample<SRC>(src, tr0 + j, t)));
        for (int o = PREP_TS / 2; o > 0; o >>= 1) mx = fmaxf(mx, __shfl_xor(mx, o));
        if (sub == 0) s_scale[2 * m + half] = mx;
    }
    __syncthreads();
    const float m_own = s_scale[lane];
    const float scale = m_own > 0.0f ? m_own : 1.0f;
    const float inv = 1.0f / scale;
    float e1 = 0.0f, e2 = 0.0f, e3 = 0.0f, e4 = 0.0f, e5 = 0.0f, e6 = 0.0f;   /* e[t-1] .. */
    for (int t0 = 0; t0 < ns; t0 += PREP_TS) {
        const int t = t0 + sub;
        const int tc = t < ns ? t : ns - 1;
        uint32_t wv[KTM_WAVE / 2];
#pragma unroll
        for (int m = 0; m < KTM_WAVE / 2; ++m)
            wv[m] = load_word<SRC>(src, tr0 + min(2 * m + half, ntl - 1), tc);
#pragma unroll
        for (int m = 0; m < KTM_WAVE / 2; ++m) s_in[2 * m + half][sub] = decode_word<SRC>(wv[m]);
        __syncthreads();
        const int s_end = min(PREP_TS, ns - t0);
        if (s_end == PREP_TS) {
            float xk[PREP_TS];
#pragma unroll
            for (int k = 0; k < PREP_TS; ++k) xk[k] = s_in[lane][k];
#pragma unroll
            for (int k = 0; k < PREP_TS; ++k) {
                const float fb = ((((QA1 * e1 + QA2 * e2) + QA3 * e3) + QA4 * e4) + QA5 * e5) + QA6 * e6;
                const float v = xk[k] * inv + fb;
                const uint16_t h = DT == KTM_TRACE_BF16 ? f32_to_bf16(v) : f32_to_f16(v);
                s_q[lane][k] = h;
                e6 = e5; e5 = e4; e4 = e3; e3 = e2; e2 = e1;
                e1 = v - (DT == KTM_TRACE_BF16 ? bf16_to_f32(h) : f16_to_f32(h));
            }
        } else for (int k = 0; k < s_end; ++k) {
            const float fb = ((((QA1 * e1 + QA2 * e2) + QA3 * e3) + QA4 * e4) + QA5 * e5) + QA6 * e6;
            const float v = s_in[lane][k] * inv + fb;
            const uint16_t h = DT == KTM_TRACE_BF16 ? f32_to_bf16(v) : f32_to_f16(v);
            s_q[lane][k] = h;
            e6 = e5; e5 = e4; e4 = e3; e3 = e2; e2 = e1;
            e1 = v - (DT == KTM_TRACE_BF16 ? bf16_to_f32(h) : f16_to_f32(h));
        }
        __syncthreads();
        if (sub < s_end) {
#pragma unroll 8
            for (int m = 0; m < KTM_WAVE / 2; ++m) {
                const int j = 2 * m + half;
                if (j < ntl) q[(tr0 + j) * row + t] = s_q[j][sub];
            }
        }
        __syncthreads();
    }
    for (int j = 0; j < ntl; ++j)
        for (int64_t T = ns + lane; T < row; T += KTM_WAVE) q[(tr0 + j) * row + T] = 0;
    if (lane < ntl) geom[tr0 + lane].scale = scale;
}

/* ------------------------------------------------------------------------ */
/* migration                                                                 */
/* ------------------------------------------------------------------------ */
struct MigArgs {
    const float *prep;               /* [ntr][row] filtered traces F(T) */
    const ktm_trace_geom *geom;
    int64_t ntr, row;
    float *km, *fold, *comp;
    const float *vel2;
    const int32_t *t1, *velx, *vely;
    const float *t2, *colx, *coly, *slab_cx, *slab_cy, *slab_x2thr;
    unsigned long long *counter;
    float ap_x2thr;                  /* aperture gate on the squared CMP distance */
    float aperture;                  /* for the conservative tile test only */
    float dt, rdt;                   /* dt and RN(1/dt) */
    int ns, nh, nxp, nyp, nx, ny, jobs_y, nvel_xl, nsm;
    int tiles_x, ntiles, tx_cols, ty_cols, run;
    int rx0, ry0, rx1, ry1;          /* migrated region [rx0,rx1) x [ry0,ry1) */
    /* Kahan variants: the block's 512-trace chunks dealt round-robin to
     * nparts parts (chunk c to part c mod nparts), part = blockIdx.z, so that
     * every part spans the whole survey and its workgroups carry like loads;
     * part 0 adds into km/fold/comp, part p > 0 writes its sums into
     * parts + (p-1) * 3 * part_n, region layout [X-rx0][Y-ry0][h][t] */
    int64_t nparts, part_n;
    float *parts;
};

/* Reference semantics of one traveltime index (Kirchhoff.c:1381-1384):
 * FP32 true divisions and sums, the two square roots in double and summed in
 * double, rounded to float, FP32 division by dt, floor.  Returns ns when the
 * sample falls past the trace (the T<ns test of Kirchhoff.c:1387).  r and
 * rdt are RN(1/v2) and RN(1/dt). */
__device__ __forceinline__ int exact_T(float dIS2, float dIR2, float t2, float v2, float r,
                                       float dt, float rdt, int ns)
{
    const float s1 = t2 + div_cr(dIS2, v2, r);
    const float s2 = t2 + div_cr(dIR2, v2, r);
    const float dsr = (float)(sqrt_cr((double)s1) + sqrt_cr((double)s2));
    const float qq = div_cr(dsr, dt, rdt);
    return qq < (float)ns ? (int)qq : ns;
}

/* Fast path of the traveltime index.  With r = RN(1/v2) and rdt = RN(1/dt),
 *   q = RN( RN( sqrt_hw(fma(dIS2, r, t2)) + sqrt_hw(fma(dIR2, r, t2)) ) * rdt ).
 * Against the reference value Q = RN(dsr/dt) of exact_T (u = 2^-24):
 *   fma vs RN(t2 + RN(dIS2/v2))            relative 4u on the radicand
 *   -> 2u after the square root, plus the hardware sqrt (<= 1 ulp = 2u)
 *   sum rounding vs the double sum rounded to float         2u
 *   RN(1/dt) and the product rounding vs RN(dsr/dt)         3u
 * so |q - Q| <= 9u Q (first order).  KTM_FAST_EPS = 10u = 5.96e-7 rounded up;
 * the 1 ulp figure of v_sqrt_f32 is checked exhaustively on [1, 4) by the GPU
 * tests, the whole bound on random and near-integer inputs. */
#define KTM_FAST_EPS 6.0e-7f

__device__ __forceinline__ float fast_q(float dIS2, float dIR2, float r, float t2, float rdt)
{
    const float a1 = __builtin_fmaf(dIS2, r, t2);
    const float a2 = __builtin_fmaf(dIR2, r, t2);
    const float d = __builtin_amdgcn_sqrtf(a1) + __builtin_amdgcn_sqrtf(a2);
    return d * rdt;
}

/* The migration kernel's form, in samples from the start: with
 * rs = RN(1/(v2 dt^2)) (product and quotient in double, one rounding) and
 * tq = 0.25 t^2 (exact for t < 4096, else one rounding),
 *   q = RN( sqrt_hw(fma(dIS2, rs, tq)) + sqrt_hw(fma(dIR2, rs, tq)) ).
 * a1 = RN(fma(dIS2, rs, tq)) is within 2u of (t2e + dIS2/v2)/dt^2 (t2e =
 * t^2 dt^2 / 4), the reference's s1 = RN(t2 + RN(dIS2/v2)) within 2u of
 * t2e + dIS2/v2 (its t2 is t2e rounded once): 4u on the radicand, 2u after
 * the square root, plus 2u for v_sqrt_f32; the sum's rounding u; the
 * reference's RN(dsr) and RN(dsr/dt) 2u: |q - Q| <= 7u Q to first order;
 * its tolerance KTM_FAST_EPS_S = 8u leaves 1u for the second-order terms.
 * No multiply by 1/dt is left in the hot loop. */
#define KTM_FAST_EPS_S 4.8e-7f
__device__ __forceinline__ float fast_qs(float dIS2, float dIR2, float rs, float tq)
{
    return __builtin_amdgcn_sqrtf(__builtin_fmaf(dIS2, rs, tq)) +
           __builtin_amdgcn_sqrtf(__builtin_fmaf(dIR2, rs, tq));
}

/* per-sample constants of fast_qs */
__device__ __forceinline__ float fast_rs(float v2, float dt)
{
    return (float)(1.0 / ((double)v2 * (double)dt * (double)dt));
}
__device__ __forceinline__ float fast_tq(int t) { return (float)(0.25 * (double)t * (double)t); }

/* q clamped to nsh = ns + 0.5 (demigration).  q >= 0 (a sum of square
 * roots).  Beyond ns + 0.5 the reference's Q is > ns as well (|q - Q| <=
 * 1e-6 q), so the sample is discarded either way; the clamped value
 * truncates to ns and is half a sample from an integer, hence never
 * ambiguous.  One clamp serves the index and the ambiguity test.  (The
 * migration kernel clamps at ns instead, see there.) */
__device__ __forceinline__ float fast_clamp(float q, float nsh) { return fminf(q, nsh); }

/* An integer within eps*(q+1) of the clamped q: floor(q) may differ from
 * the reference's floor(Q).  eps is the bound of the fast form in use. */
__device__ __forceinline__ bool fast_ambiguous(float qc, float eps = KTM_FAST_EPS)
{
    const float dev = fabsf(qc - __builtin_rintf(qc));
    return dev <= __builtin_fmaf(qc, eps, eps);
}

/* floor of the clamped q, in [0, ns] */
__device__ __forceinline__ int fast_T(float qc) { return (int)qc; }

/* staged batches whose taps the Kahan variants sum in plain FP32 before one
 * compensated step */
#define KTM_KAHAN_FOLD 4

/* q clamped to [0, hi] in one instruction (v_med3_f32; q >= 0 is a sum of
 * square roots, so this is min(q, hi) without the canonicalisation of
 * fminf's operands, which the compiler otherwise repeats per sample) */
__device__ __forceinline__ float clamp_ns(float q, float hi)
{
    return __builtin_amdgcn_fmed3f(q, 0.0f, hi);
}

/* Slots whose traveltime chains the exact path evaluates side by side
 * under one test for the rare path (K divisible by it, else 2) */
#ifndef KTM_SLOT_GROUP
#define KTM_SLOT_GROUP 2
#endif

/* Queue of the speculative exact path: items per lane and staged batch (a
 * lane's further ambiguous samples of the batch take the inline exact path) */
#define KTM_SPEC_Q 1

/* The reference traveltime index of a sample with every rule of the exact
 * path: the fast form, and where an integer lies within its error bound the
 * reference arithmetic, ns when the sample is past the trace for certain. */
__device__ __forceinline__ int reference_T(float dIS2, float dIR2, float rs, float tq, float t2,
                                           float v2, float dt, float rdt, int ns)
{
    constexpr float EPS = KTM_FAST_EPS_S;
    const float nsf = (float)ns;
    const float q = fast_qs(dIS2, dIR2, rs, tq);
    const float qc = clamp_ns(q, nsf);
    if (!fast_ambiguous(qc, EPS)) return fast_T(qc);
    if (q - nsf >= __builtin_fmaf(q, EPS, EPS)) return ns;
    return exact_T(dIS2, dIR2, t2, v2, 1.0f / v2, dt, rdt, ns);
}

/* One workgroup = 8 waves = NCOL image columns x W time segments.  Each lane
 * owns K samples t = seg*64K + 64k + lane of its column: the image values
 * (km, fold count, Kahan compensation) and the fast path's rs = RN(1/(v^2
 * dt^2)) and tq = t^2/4 stay in registers for the whole launch.  Traces
 * stream through LDS in trace order: a chunk of 512 is filtered to those
 * whose CMP can reach the tile (ordered compaction), then staged 8 at a
 * time; per staged trace the aperture and slab gates of every column are
 * evaluated lane-parallel (one lane per trace) and the passing traces are
 * walked in order.
 *
 * Traveltime index, two variants:
 * - exact (the FP32 reference path, and the Kahan modes unless FASTIDX): the
 *   reference's index, resolved speculatively.  The fast form's floor(q) is
 *   used at once; a sample whose q lies within the form's error bound of an
 *   integer (and is not past the trace for certain) is added with that index
 *   and queued in LDS with the running sum before the add (one per lane and
 *   batch; a lane's further ambiguous samples take the inline exact path).
 *   At the end of the staged batch one wave-wide pass computes the reference
 *   index of every lane's queued sample together (the exact arithmetic used
 *   to run for the whole wave per ambiguous lane and slot); where it differs
 *   from the fast one (a few percent of them) the sample's running sum is
 *   rebuilt from the saved value through the batch's later traces in order,
 *   so every sample receives its additions in the reference's order and the
 *   result equals the inline exact path's bit for bit.  v^2 of the column's
 *   samples stays in LDS for the reference arithmetic.
 * - FASTIDX: T = floor(min(q, ns)) of the fast form alone (CLOSED and 16-bit
 *   modes, whose contract is an error bound against FP64): no ambiguity
 *   test; fold is summed from a row V[T] = (T < ns) staged once in LDS (an
 *   address add and a float add per sample instead of a compare and a carry
 *   add, both quarter-rate on gfx950).
 * q is clamped at ns, so in the exact variants a sample at or past the end of
 * the trace lands on an integer and takes the rare branch, where the samples
 * that do not count towards fold are counted: the common path has no compare
 * for fold.  Every variant fits 128 VGPRs (4 waves per SIMD, two workgroups
 * per CU, one staging while the other computes; the Kahan exact variant
 * spills a few 64-bit addresses that are reloaded once per chunk of 512
 * traces). */
template <int K, int W, int B, bool KAHAN, bool FASTIDX>
__global__ void __launch_bounds__(8 * KTM_WAVE) __attribute__((amdgpu_waves_per_eu(4, 4)))
k_migrate(const MigArgs a)
{
    constexpr int NCOL = 8 / W;             /* image columns per workgroup */
    constexpr int NW = NCOL * W;            /* waves per workgroup (8) */
    constexpr int NT = NW * KTM_WAVE;
    constexpr bool SPEC = !FASTIDX;
    static_assert(B <= 16 && K <= 16, "queue items pack the slot and the trace in 4 bits each");
    extern __shared__ float4 lds4[];
    const int ldl = (int)a.row;             /* LDS row = prepared row (floats) */
    float *sF = reinterpret_cast<float *>(lds4);            /* [B][ldl] */
    float *sG = sF + B * ldl;                               /* [B][8]   */
    int *sSel = reinterpret_cast<int *>(sG + B * 8);        /* [NT]     */
    int *sCnt = sSel + NT;                                  /* [NW]     */
    /* exact index: v^2 per sample [NCOL][W*K*64], then the speculation
     * queue [NW][KTM_SPEC_Q][64] of uint2 (running sum before the add,
     * slot | trace << 4 | index << 8 | mismatch << 31: the fast index when
     * queued, the reference's once resolved); fast index: V[T], [ldl] */
    float *sV2 = reinterpret_cast<float *>(sCnt + NW);
    uint2 *sQ = reinterpret_cast<uint2 *>(sV2 + NCOL * (W * K * KTM_WAVE));

    const int tid = threadIdx.x, lane = tid & (KTM_WAVE - 1), wv = tid / KTM_WAVE;
    /* XCD-aware tile order: consecutive workgroup ids go to different XCDs.
     * Each XCD takes runs of a.run consecutive tiles (a row of tiles in 3-D:
     * neighbouring columns share traces in its L2) and the runs are dealt
     * round-robin, so every XCD gets rows from the whole image; contiguous
     * bands per XCD left the edge bands (fewer traces within aperture) at
     * 0.73 of the mean work and the middle ones at 1.14. */
    const int wi = blockIdx.x >> 3;
    const int tile = (((wi / a.run) << 3) + (int)(blockIdx.x & 7)) * a.run + wi % a.run;
    if (tile >= a.ntiles) return;
    const int h = blockIdx.y;
    const int part = KAHAN ? (int)blockIdx.z : 0;
    const int tx = tile % a.tiles_x, ty = tile / a.tiles_x;
    const int cw = wv / W, seg = wv % W;    /* column within tile, time segment */
    const int X = a.rx0 + tx * a.tx_cols + cw % a.tx_cols;
    const int Y = a.ry0 + ty * a.ty_cols + cw / a.tx_cols;
    const bool col_ok = X < a.rx1 && Y < a.ry1;
    const int Xc = X < a.rx1 ? X : a.rx1 - 1, Yc = Y < a.ry1 ? Y : a.ry1 - 1;
    const float IX = a.colx[Xc], IY = a.coly[Yc];
    const int job = (Xc / a.nx) * a.jobs_y + (Yc / a.ny);
    const float scx = a.slab_cx[job], scy = a.slab_cy[job];
    const float sthr = a.slab_x2thr[job];
    const int ns = a.ns;
    const float nsf = (float)ns;

    /* conservative tile test: the CMP must be within aperture of some
     * column of the tile (exact per-column gates follow) */
    const int x0 = a.rx0 + tx * a.tx_cols, x1 = min(x0 + a.tx_cols, a.rx1) - 1;
    const int y0 = a.ry0 + ty * a.ty_cols, y1 = min(y0 + a.ty_cols, a.ry1) - 1;
    const float ccx = 0.5f * (a.colx[x0] + a.colx[x1]), ccy = 0.5f * (a.coly[y0] + a.coly[y1]);
    const float hx = 0.5f * fabsf(a.colx[x1] - a.colx[x0]);
    const float hy = 0.5f * fabsf(a.coly[y1] - a.coly[y0]);
    const float reach = a.aperture * 1.0001f + sqrtf(hx * hx + hy * hy) + 1.0f;
    const float reach2 = reach * reach;

    /* exact variants: v^2 of this column's samples in LDS for the exact path */
    float *sV2c = sV2 + cw * (W * K * KTM_WAVE);
    if (FASTIDX) {
        for (int c = tid; c < ldl; c += NT) sV2[c] = c < ns ? 1.0f : 0.0f;
    }
    uint2 *sQw = sQ + wv * (KTM_SPEC_Q * KTM_WAVE);
    const int64_t pix = (((int64_t)Xc * a.nyp + Yc) * a.nh + h) * ns;
    const int64_t vbase = ((int64_t)a.velx[Xc] * a.nvel_xl + a.vely[Yc]) * a.nsm;
    float rs[K], tq[K], acc[K], cmp[K];
    int cnt[FASTIDX ? 1 : K];               /* passing traces with T >= ns, per slot */
    float vcnt[FASTIDX ? K : 1];            /* passing traces with T < ns, per slot */
#pragma unroll
    for (int k = 0; k < K; ++k) {
        const int tl = seg * K * KTM_WAVE + k * KTM_WAVE + lane;
        if (!FASTIDX) cnt[FASTIDX ? 0 : k] = 0;
        if (FASTIDX) vcnt[FASTIDX ? k : 0] = 0.0f;
        cmp[k] = 0.0f;
        if (tl < ns) {
            const float v2 = a.vel2[vbase + a.t1[tl]];
            if (!FASTIDX) sV2c[tl] = v2;
            rs[k] = fast_rs(v2, a.dt);
            tq[k] = fast_tq(tl);
            acc[k] = col_ok && part == 0 ? a.km[pix + tl] : 0.0f;
            if (KAHAN && col_ok && part == 0) cmp[k] = a.comp[pix + tl];
        } else {
            if (!FASTIDX) sV2c[tl] = 1.0f;
            /* padding lane: q = 0.5, never flagged by the ambiguity test; it
             * reads F[0] (and V[0]) into values that are not stored */
            rs[k] = 0.0f;
            tq[k] = 0.0625f;
            acc[k] = 0.0f;
        }
    }

    /* Kahan variants: the taps are summed onto y = -c in plain FP32 over
     * KTM_KAHAN_FOLD staged batches (at most KTM_KAHAN_FOLD * B terms, error
     * <= (terms - 1) u times their magnitudes) and added to the image with one
     * compensated step, which keeps the compensated error bound of order
     * u * sum |f| at a fraction of the per-tap cost.  Plain variants add
     * every tap into the image in trace order, as the reference does. */
    float y[KAHAN ? K : 1];
#pragma unroll
    for (int k = 0; k < (KAHAN ? K : 1); ++k) y[k] = -cmp[k];
    int nfold = 0;
    auto fold_y = [&]() __attribute__((always_inline)) {
#pragma unroll
        for (int k = 0; k < (KAHAN ? K : 1); ++k) {
            const float yk = y[k];
            const float s = acc[k] + yk;
            y[k] = yk - (s - acc[k]);           /* -c, c = (s - acc) - y */
            acc[k] = s;
        }
    };
    /* the squared source and receiver distances of staged trace j from this
     * column, as the gate evaluation forms them */
    auto staged_dist = [&](int j, float &dIS2, float &dIR2) __attribute__((always_inline)) {
        const float4 g0 = reinterpret_cast<const float4 *>(sG + j * 8)[0];
        const float ax = IX - g0.x, ay = IY - g0.y, bx = IX - g0.z, by = IY - g0.w;
        dIS2 = ax * ax + ay * ay;
        dIR2 = bx * bx + by * by;
    };

    uint32_t npass = 0;
    const int n4 = ldl / 4;
    const int64_t cbeg = KAHAN ? (int64_t)part * NT : 0;
    const int64_t cstep = KAHAN ? a.nparts * NT : NT;
    const int64_t cend = a.ntr;
    for (int64_t c0 = cbeg; c0 < cend; c0 += cstep) {
        /* ---- select the traces of this chunk that can reach the tile, in order ---- */
        const int64_t i = c0 + tid;
        bool pred = false;
        if (i < cend) {
            const ktm_trace_geom g = a.geom[i];
            const float dx = g.cmpx - ccx, dy = g.cmpy - ccy;
            pred = g.off == h && dx * dx + dy * dy <= reach2;
        }
        const uint64_t m = __ballot(pred);
        const int pre = __popcll(m & ((1ull << lane) - 1ull));
        if (lane == 0) sCnt[wv] = __popcll(m);
        __syncthreads();
        int base = 0, total = 0;
#pragma unroll
        for (int q = 0; q < NW; ++q) {
            const int v = sCnt[q];
            base += q < wv ? v : 0;
            total += v;
        }
        if (pred) sSel[base + pre] = tid;
        __syncthreads();

        for (int b0 = 0; b0 < total; b0 += B) {
            const int nb = min(B, total - b0);
            /* ---- stage nb traces (filtered samples + geometry) into LDS ---- */
            for (int j = 0; j < nb; ++j) {
                const int64_t ti = c0 + sSel[b0 + j];
                const float4 *src = reinterpret_cast<const float4 *>(a.prep + ti * a.row);
                for (int c = tid; c < n4; c += NT)
                    reinterpret_cast<float4 *>(sF + j * ldl)[c] = src[c];
            }
            if (tid < nb * 2) {
                const int j = tid >> 1, half = tid & 1;
                const int64_t ti = c0 + sSel[b0 + j];
                reinterpret_cast<float4 *>(sG + j * 8)[half] =
                    reinterpret_cast<const float4 *>(a.geom + ti)[half];
            }
            __syncthreads();

            if (col_ok) {
                /* gates and distances of Kirchhoff.c:1334-1354 for every staged
                 * trace at once (lane j = trace j), on squared distances with
                 * the exact thresholds of ktm_gate_threshold */
                float g_is = 0.0f, g_ir = 0.0f;
                bool gp = false;
                if (lane < nb) {
                    staged_dist(lane, g_is, g_ir);
                    const float4 g1 = reinterpret_cast<const float4 *>(sG + lane * 8)[1];
                    const float cx = IX - g1.x, cy = IY - g1.y;
                    const float ex = scx - g1.x, ey = scy - g1.y;
                    gp = ex * ex + ey * ey < sthr && cx * cx + cy * cy < a.ap_x2thr;
                }
                const uint64_t pm_batch = __ballot(gp);
                uint64_t pm = pm_batch;
                npass += __popcll(pm);
                int qn = 0;                         /* queued samples of this lane (plain) */
                /* Hot loop, Kirchhoff.c:1374-1399: the tap F[T] of slot k for
                 * one trace (and, fast index, V[T] for fold). */
                auto tap = [&](int k, int j, float dIS2, float dIR2, const float *Fr) __attribute__((always_inline)) {
                    const float q = fast_qs(dIS2, dIR2, rs[k], tq[k]);
                    const float qc = clamp_ns(q, nsf);
                    int T = fast_T(qc);
                    if (FASTIDX) {
                        /* V sits (sV2 - Fr) floats from F: one address add */
                        const float *Fp = Fr + T;
                        vcnt[FASTIDX ? k : 0] += Fp[sV2 - Fr];
                        return *Fp;
                    }
                    constexpr float EPS = KTM_FAST_EPS_S;
                    if (fast_ambiguous(qc, EPS)) {
                        if (q - nsf >= __builtin_fmaf(q, EPS, EPS)) {
                            T = ns;   /* Q >= ns as well: past the trace */
                            cnt[FASTIDX ? 0 : k] += 1;
                        } else if (SPEC && qn < KTM_SPEC_Q) {
                            /* speculate: floor(q) now, the reference index at
                             * the end of the batch */
                            sQw[qn * KTM_WAVE + lane] = make_uint2(
                                __float_as_uint(KAHAN ? y[KAHAN ? k : 0] : acc[k]),
                                (uint32_t)(k | (j << 4) | (T << 8)));
                            ++qn;
                        } else {
                            const int tl = seg * K * KTM_WAVE + k * KTM_WAVE + lane;
                            const float v2 = sV2c[tl];
                            T = exact_T(dIS2, dIR2, a.t2[tl], v2, 1.0f / v2, a.dt, a.rdt, ns);
                            cnt[FASTIDX ? 0 : k] += T >= ns ? 1 : 0;
                        }
                    }
                    return Fr[T];
                };
                /* the rare path of tap() for slot k: ambiguous or past the trace */
                auto rare = [&](int k, int j, float q, float dIS2, float dIR2, int &T) __attribute__((always_inline)) {
                    constexpr float EPS = KTM_FAST_EPS_S;
                    if (q - nsf >= __builtin_fmaf(q, EPS, EPS)) {
                        T = ns;   /* Q >= ns as well: past the trace */
                        cnt[FASTIDX ? 0 : k] += 1;
                    } else if (SPEC && qn < KTM_SPEC_Q) {
                        sQw[qn * KTM_WAVE + lane] = make_uint2(
                            __float_as_uint(KAHAN ? y[KAHAN ? k : 0] : acc[k]),
                            (uint32_t)(k | (j << 4) | (T << 8)));
                        ++qn;
                    } else {
                        const int tl = seg * K * KTM_WAVE + k * KTM_WAVE + lane;
                        const float v2 = sV2c[tl];
                        T = exact_T(dIS2, dIR2, a.t2[tl], v2, 1.0f / v2, a.dt, a.rdt, ns);
                        cnt[FASTIDX ? 0 : k] += T >= ns ? 1 : 0;
                    }
                };
                while (pm) {
                    const int j = __builtin_ctzll(pm);   /* traces in order */
                    pm &= pm - 1ull;
                    const float dIS2 = __int_as_float(__builtin_amdgcn_readlane(__float_as_int(g_is), j));
                    const float dIR2 = __int_as_float(__builtin_amdgcn_readlane(__float_as_int(g_ir), j));
                    const float *Fr = sF + j * ldl;
                    constexpr int G = K % KTM_SLOT_GROUP == 0 ? KTM_SLOT_GROUP : 2;
                    if (FASTIDX || (K % G)) {
#pragma unroll
                        for (int k = 0; k < K; ++k) {
                            const float f = tap(k, j, dIS2, dIR2, Fr);
                            if (KAHAN) y[KAHAN ? k : 0] = y[KAHAN ? k : 0] + f;
                            else acc[k] = acc[k] + f;
                        }
                    } else {
                        /* exact index, G slots per branch: their chains run
                         * side by side, one test guards their rare paths */
#pragma unroll
                        for (int k = 0; k < K; k += G) {
                            constexpr float EPS = KTM_FAST_EPS_S;
                            float qg[G];
                            int Tg[G];
                            bool ag[G], any = false;
#pragma unroll
                            for (int i = 0; i < G; ++i) {
                                qg[i] = fast_qs(dIS2, dIR2, rs[k + i], tq[k + i]);
                                const float qc = clamp_ns(qg[i], nsf);
                                Tg[i] = fast_T(qc);
                                ag[i] = fast_ambiguous(qc, EPS);
                                any = any || ag[i];
                            }
                            if (any) {
#pragma unroll
                                for (int i = 0; i < G; ++i)
                                    if (ag[i]) rare(k + i, j, qg[i], dIS2, dIR2, Tg[i]);
                            }
#pragma unroll
                            for (int i = 0; i < G; ++i) {
                                const float f = Fr[Tg[i]];
                                if (KAHAN) y[KAHAN ? k + i : 0] = y[KAHAN ? k + i : 0] + f;
                                else acc[k + i] = acc[k + i] + f;
                            }
                        }
                    }
                }
                if (SPEC && __ballot(qn > 0)) {
                    /* ---- resolve the queued samples: one pass per queue
                     * depth, every lane's item of that depth together ---- */
                    bool redo = false;
                    for (int r = 0; r < KTM_SPEC_Q; ++r) {
                        if (!__ballot(qn > r)) break;
                        if (qn > r) {
                            uint2 e = sQw[r * KTM_WAVE + lane];
                            const int k = (int)(e.y & 15u), j = (int)((e.y >> 4) & 15u);
                            const int tl = seg * K * KTM_WAVE + k * KTM_WAVE + lane;
                            const int Tf = (int)((e.y >> 8) & 0x7fffu);
                            float dIS2, dIR2;
                            staged_dist(j, dIS2, dIR2);
                            const float v2 = sV2c[tl];
                            const int Te = exact_T(dIS2, dIR2, a.t2[tl], v2, 1.0f / v2, a.dt,
                                                   a.rdt, ns);
                            if (Te >= ns) {
#pragma unroll
                                for (int kk = 0; kk < K; ++kk)
                                    if (kk == k) cnt[FASTIDX ? 0 : kk] += 1;
                            }
                            e.y = (e.y & 0xffu) | ((uint32_t)Te << 8) |
                                  (Te != Tf ? 0x80000000u : 0u);
                            sQw[r * KTM_WAVE + lane] = e;
                            redo = redo || Te != Tf;
                        }
                    }
                    /* ---- rebuild the sums whose queued index was wrong: from
                     * the sum before the earliest wrong one, through the
                     * batch's later passing traces in order (rare) ---- */
                    if (__ballot(redo) && redo) {
                        uint32_t done = 0;              /* slots rebuilt */
                        for (int r = 0; r < qn; ++r) {
                            const uint2 e = sQw[r * KTM_WAVE + lane];
                            const int k = (int)(e.y & 15u), j = (int)((e.y >> 4) & 15u);
                            if (!(e.y & 0x80000000u) || (done >> k & 1u)) continue;
                            done |= 1u << k;
                            const int tl = seg * K * KTM_WAVE + k * KTM_WAVE + lane;
                            const float v2 = sV2c[tl];
                            float rsk = 0.0f;
#pragma unroll
                            for (int kk = 0; kk < K; ++kk)
                                if (kk == k) rsk = rs[kk];
                            const float tqk = fast_tq(tl), t2k = a.t2[tl];
                            float s = __uint_as_float(e.x) +
                                      sF[j * ldl + (int)((e.y >> 8) & 0x7fffu)];
                            uint64_t later = pm_batch & ~((2ull << j) - 1ull);
                            while (later) {
                                const int jj = __builtin_ctzll(later);
                                later &= later - 1ull;
                                float dIS2, dIR2;
                                staged_dist(jj, dIS2, dIR2);
                                const int T = reference_T(dIS2, dIR2, rsk, tqk, t2k, v2, a.dt,
                                                          a.rdt, ns);
                                s = s + sF[jj * ldl + T];
                            }
#pragma unroll
                            for (int kk = 0; kk < K; ++kk) {
                                if (kk != k) continue;
                                if (KAHAN) y[KAHAN ? kk : 0] = s;
                                else acc[kk] = s;
                            }
                        }
                    }
                }
                if (KAHAN && ++nfold == KTM_KAHAN_FOLD) {
                    nfold = 0;
                    fold_y();
                }
            }
            __syncthreads();
        }
    }
    if (KAHAN && nfold) fold_y();

    if (KAHAN && part > 0 && col_ok) {
        /* a later part of the block: its own sums, merged by k_merge_parts */
        float *pk = a.parts + (int64_t)(part - 1) * 3 * a.part_n;
        const int64_t rpix =
            (((int64_t)(X - a.rx0) * (a.ry1 - a.ry0) + (Y - a.ry0)) * a.nh + h) * ns;
#pragma unroll
        for (int k = 0; k < K; ++k) {
            const int t = seg * K * KTM_WAVE + k * KTM_WAVE + lane;
            if (t < ns) {
                pk[rpix + t] = acc[k];
                pk[a.part_n + rpix + t] = FASTIDX ? vcnt[FASTIDX ? k : 0]
                                                  : (float)((int)npass - cnt[FASTIDX ? 0 : k]);
                pk[2 * a.part_n + rpix + t] = -y[KAHAN ? k : 0];
            }
        }
        if (lane == 0 && seg == 0 && a.counter)
            atomicAdd(a.counter, (unsigned long long)npass * (unsigned long long)ns);
    } else if (col_ok) {
#pragma unroll
        for (int k = 0; k < K; ++k) {
            const int t = seg * K * KTM_WAVE + k * KTM_WAVE + lane;
            if (t < ns) {
                a.km[pix + t] = acc[k];
                /* fold adds the passing traces whose sample was inside the
                 * trace (Kirchhoff.c:1387-1396) */
                a.fold[pix + t] = a.fold[pix + t] +
                                  (FASTIDX ? vcnt[FASTIDX ? k : 0]
                                           : (float)((int)npass - cnt[FASTIDX ? 0 : k]));
                if (KAHAN) a.comp[pix + t] = -y[KAHAN ? k : 0];
            }
        }
        if (lane == 0 && seg == 0 && a.counter)
            atomicAdd(a.counter, (unsigned long long)npass * (unsigned long long)ns);
    }
}

/* The compensated sums of parts 1..nparts-1 of a trace block added to the
 * image (part 0 added into it directly): per sample, in FP64, the image's
 * km - comp and each part's km - comp in part order, rounded once to the new
 * km, with comp = km - the FP64 total (so km - comp keeps the total to
 * FP64); fold adds the parts' counts, which are exact.  One thread per
 * sample of the region. */
__global__ void k_merge_parts(const MigArgs a, int nparts)
{
    const int64_t ht = (int64_t)a.nh * a.ns;
    const int rny = a.ry1 - a.ry0;
    for (int64_t i = (int64_t)blockIdx.x * blockDim.x + threadIdx.x; i < a.part_n;
         i += (int64_t)gridDim.x * blockDim.x) {
        const int64_t c = i / ht, r = i - c * ht;
        const int64_t pix = ((int64_t)(a.rx0 + c / rny) * a.nyp + (a.ry0 + c % rny)) * ht + r;
        double tot = (double)a.km[pix] - (double)a.comp[pix];
        float f = a.fold[pix];
        for (int p = 1; p < nparts; ++p) {
            const float *pk = a.parts + (int64_t)(p - 1) * 3 * a.part_n;
            tot += (double)pk[i] - (double)pk[2 * a.part_n + i];
            f += pk[a.part_n + i];
        }
        const float k = (float)tot;
        a.km[pix] = k;
        a.fold[pix] = f;
        a.comp[pix] = (float)((double)k - tot);
    }
}

/* Contributions of the traces outside the offset classes [0, nh): the
 * reference evaluates their loop body for every column in aperture and
 * drops the sum (the off test of Kirchhoff.c:1387); k_migrate never stages
 * them, so their (trace, column) pairs are counted here, with the gates of
 * k_migrate, for the contribution count.  One thread per trace. */
__global__ void k_count_outside(const MigArgs a)
{
    const int64_t i = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
    if (i >= a.ntr) return;
    const ktm_trace_geom g = a.geom[i];
    if (g.off >= 0 && g.off < a.nh) return;
    uint64_t n = 0;
    for (int X = a.rx0; X < a.rx1; ++X) {
        const float cx = a.colx[X] - g.cmpx;
        if (!(cx * cx < a.ap_x2thr)) continue;   /* RN(cx^2 + cy^2) >= cx^2 */
        for (int Y = a.ry0; Y < a.ry1; ++Y) {
            const float cy = a.coly[Y] - g.cmpy;
            const int job = (X / a.nx) * a.jobs_y + (Y / a.ny);
            const float ex = a.slab_cx[job] - g.cmpx, ey = a.slab_cy[job] - g.cmpy;
            n += (ex * ex + ey * ey < a.slab_x2thr[job] && cx * cx + cy * cy < a.ap_x2thr) ? 1u : 0u;
        }
    }
    if (n) atomicAdd(a.counter, (unsigned long long)(n * (uint64_t)a.ns));
}

/* ------------------------------------------------------------------------ */
/* image reset and post-stack                                                */
/* ------------------------------------------------------------------------ */
__global__ void k_image_reset(float *km, float *fold, float *comp, int64_t n)
{
    for (int64_t i = (int64_t)blockIdx.x * blockDim.x + threadIdx.x; i < n;
         i += (int64_t)gridDim.x * blockDim.x) {
        km[i] = 0.0f;
        fold[i] = 1.0f;
        if (comp) comp[i] = 0.0f;
    }
}

/* The post-stack low-pass of Kirchhoff.c:1457-1479 multiplies the DFT by the
 * even real spectrum anti_alias, i.e. it is the circular convolution
 * out[t] = sum_u h[(t-u) mod ns] d[u] with h the (real) inverse DFT of
 * anti_alias divided by ns, evaluated in FP64; h2[m] = h[m mod ns] for
 * m < 2 ns.  PS_NT: threads per workgroup of the transpose. */
#define PS_NT 256

/* x of the post-stack filter, Kirchhoff.c:1460: a float expression stored
 * into a double */
struct PostKtm {
    static __device__ __forceinline__ double x(const float *k, const float *f, int u, float dt)
    {
        return (double)(-dt * (float)u * (k[u] - k[u - 1]) * dt / f[u]);
    }
};

/* ------------------------------------------------------------------------ */
/* seeded synthetic surveys                                                  */
/* ------------------------------------------------------------------------ */
/* One workgroup per trace: the DSR diffraction time of every scatterer,
 *   tau_j = sqrt((t0/2)^2 + |s-x_j|^2/v_j^2) + sqrt((t0/2)^2 + |g-x_j|^2/v_j^2),
 * v_j = v0 + vt*t0_j + vx*x_j, in LDS; then each sample sums the Ricker
 * wavelets (1 - 2 a) exp(-a), a = (pi f0 (t dt - tau_j))^2, of the scatterers
 * within 2/f0 (the wavelet is below 1e-13 there), in scatterer order. */
#define SYN_NT 256
#define SYN_MAXSCAT 1024
__global__ void __launch_bounds__(SYN_NT)
k_synth_traces(const ktm_trace_geom *__restrict__ geom, int64_t ntr, int ns, float dt,
               const float *__restrict__ scat, int nscat, float f0, float v0, float vt, float vx,
               float *__restrict__ traces)
{
    __shared__ float s_tau[SYN_MAXSCAT], s_amp[SYN_MAXSCAT];
    const int64_t i = blockIdx.x;
    if (i >= ntr) return;
    const ktm_trace_geom g = geom[i];
    for (int j = threadIdx.x; j < nscat; j += SYN_NT) {
        const double x = scat[4 * j], y = scat[4 * j + 1], t0 = scat[4 * j + 2];
        const double v = (double)v0 + (double)vt * t0 + (double)vx * x;
        const double h = 0.25 * t0 * t0, iv2 = 1.0 / (v * v);
        const double ds2 = (g.sxt - x) * (g.sxt - x) + (g.syt - y) * (g.syt - y);
        const double dg2 = (g.gxt - x) * (g.gxt - x) + (g.gyt - y) * (g.gyt - y);
        s_tau[j] = (float)(sqrt(h + ds2 * iv2) + sqrt(h + dg2 * iv2));
        s_amp[j] = scat[4 * j + 3];
    }
    __syncthreads();
    const float win = 2.0f / f0;
    const float pf = 3.14159265358979f * f0;
    float *o = traces + i * (int64_t)ns;
    for (int t = threadIdx.x; t < ns; t += SYN_NT) {
        const float tt = (float)t * dt;
        float acc = 0.0f;
        for (int j = 0; j < nscat; ++j) {
            const float d = tt - s_tau[j];
            if (fabsf(d) < win) {
                const float a = (pf * d) * (pf * d);
                acc += s_amp[j] * (1.0f - 2.0f * a) * expf(-a);
            }
        }
        o[t] = acc;
    }
}

/* counter-based Gaussian noise: SplitMix64 of (seed, index) -> two uniforms
 * -> Box-Muller; the value at an index does not depend on the launch shape */
__device__ __forceinline__ uint64_t splitmix64(uint64_t z)
{
    z += 0x9E3779B97F4A7C15ull;
    z = (z ^ (z >> 30)) * 0xBF58476D1CE4E5B9ull;
    z = (z ^ (z >> 27)) * 0x94D049BB133111EBull;
    return z ^ (z >> 31);
}

__global__ void k_synth_noise(float *traces, int64_t n, float sigma, uint64_t seed)
{
    for (int64_t i = (int64_t)blockIdx.x * blockDim.x + threadIdx.x; i < n;
         i += (int64_t)gridDim.x * blockDim.x) {
        const uint64_t r = splitmix64(seed ^ splitmix64((uint64_t)i));
        const double u1 = ((double)(r >> 40) + 0.5) * (1.0 / 16777216.0);
        const double u2 = (double)((r >> 16) & 0xffffffull) * (1.0 / 16777216.0);
        const double z = sqrt(-2.0 * log(u1)) * cos(6.283185307179586 * u2);
        traces[i] += (float)(sigma * z);
    }
}

extern "C" int ktm_synth_traces(const ktm_plan *p, const ktm_trace_geom *geom, int64_t ntr,
                                const float *scat, int nscat, float f0, float v0, float vt,
                                float vx, float *traces, void *stream)
{
    if (!p || ntr < 0 || nscat < 0 || nscat > SYN_MAXSCAT || !(f0 > 0.0f)) return KTM_ERR_ARG;
    if (!ntr) return KTM_OK;
    if (!geom || !traces || (nscat && !scat)) return KTM_ERR_ARG;
    for (int64_t i0 = 0; i0 < ntr; i0 += 1 << 30) {
        const int64_t n = ntr - i0 < (1 << 30) ? ntr - i0 : (1 << 30);
        hipLaunchKernelGGL(k_synth_traces, dim3((unsigned)n), dim3(SYN_NT), 0, (hipStream_t)stream,
                           geom + i0, n, p->tb.s.ns, p->tb.s.dt, scat, nscat, f0, v0, vt, vx,
                           traces + i0 * p->tb.s.ns);
    }
    HIP_TRY(hipGetLastError());
    return KTM_OK;
}

extern "C" int ktm_synth_noise(float *traces, int64_t n, float sigma, uint64_t seed, void *stream)
{
    if (n < 0 || (n && !traces)) return KTM_ERR_ARG;
    if (!n) return KTM_OK;
    hipLaunchKernelGGL(k_synth_noise, dim3(4096), dim3(256), 0, (hipStream_t)stream, traces, n,
                       sigma, seed);
    HIP_TRY(hipGetLastError());
    return KTM_OK;
}

/* LDS read bandwidth probe: every workgroup fills 16 KB of LDS, then each
 * lane reads 16 bytes per step from a conflict-free rotating position,
 * folded with xor so the VALU work per byte stays well below the LDS rate;
 * one word per lane is written so the reads are not removed.  Reads
 * nblocks * 256 * iters * 4 * 16 bytes of LDS. */
__global__ void __launch_bounds__(256) k_probe_lds(uint32_t *__restrict__ out, int iters)
{
    __shared__ uint4 buf[4 * 256];
    const int tid = threadIdx.x;
    for (int k = 0; k < 4; ++k)
        buf[k * 256 + tid] = make_uint4(tid * 4u + k, tid ^ 0x55u, blockIdx.x, k * 7u + 1u);
    __syncthreads();
    uint32_t a = 0u, b = 0u;
    for (int i = 0; i < iters; ++i) {
#pragma unroll
        for (int k = 0; k < 4; ++k) {
            const uint4 v = buf[k * 256 + ((tid + i) & 255)];
            a ^= v.x ^ v.y;
            b ^= v.z ^ v.w;
        }
    }
    out[(int64_t)blockIdx.x * 256 + tid] = a ^ b;
}

extern "C" int ktm_probe_lds(uint32_t *out, int nblocks, int iters, void *stream)
{
    if (!out || nblocks < 1 || iters < 1) return KTM_ERR_ARG;
    hipLaunchKernelGGL(k_probe_lds, dim3((unsigned)nblocks), dim3(256), 0, (hipStream_t)stream,
                       out, iters);
    HIP_TRY(hipGetLastError());
    return KTM_OK;
}

/* ------------------------------------------------------------------------ */
/* self-tests of the numerical assumptions of the fast traveltime path       */
/* ------------------------------------------------------------------------ */
/* out_hw[i] = v_sqrt_f32(x[i]) */
__global__ void k_selftest_sqrt(const float *__restrict__ x, int64_t n, float *__restrict__ out_hw)
{
    const int64_t i = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
    if (i < n) out_hw[i] = __builtin_amdgcn_sqrtf(x[i]);
}

/* For in[i] = (dIS2, dIR2, t, v2) (t a sample index held as a float):
 * q_fast of the migration kernel, the reference Q = RN(dsr/dt) (float,
 * before the floor), and the T each selects; t2 is the table's value. */
__global__ void k_selftest_traveltime(const float4 *__restrict__ in, int64_t n, float dt, int ns,
                                      float *__restrict__ q_fast, float *__restrict__ q_ref,
                                      int *__restrict__ t_kernel, int *__restrict__ t_ref)
{
    const int64_t i = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
    if (i >= n) return;
    const float4 in4 = in[i];
    const int t = (int)in4.z;
    const float4 v = make_float4(in4.x, in4.y, (float)(0.25 * t * t * dt * dt), in4.w);
    const float r = 1.0f / v.w, rdt = 1.0f / dt;
    const float q = fast_qs(v.x, v.y, fast_rs(v.w, dt), fast_tq(t));
    /* the decision procedure of k_migrate */
    const float nsf = (float)ns, qc = fminf(q, nsf);
    int T = fast_T(qc);
    if (fast_ambiguous(qc, KTM_FAST_EPS_S))
        T = q - nsf >= __builtin_fmaf(q, KTM_FAST_EPS_S, KTM_FAST_EPS_S)
                ? ns : exact_T(v.x, v.y, v.z, v.w, r, dt, rdt, ns);
    /* the reference computed independently of exact_T: the compiler's
     * correctly rounded division and double square root */
    const float s1 = v.z + v.x / v.w, s2 = v.z + v.y / v.w;
    const float dsr = (float)(__builtin_sqrt((double)s1) + __builtin_sqrt((double)s2));
    const float qq = dsr / dt;
    q_fast[i] = q;
    q_ref[i] = qq;
    t_kernel[i] = T;
    t_ref[i] = qq < nsf ? (int)qq : ns;
}

extern "C" int ktm_selftest_sqrt(const float *x, int64_t n, float *out_hw, void *stream)
{
    if (n < 0 || (n && (!x || !out_hw))) return KTM_ERR_ARG;
    if (!n) return KTM_OK;
    hipLaunchKernelGGL(k_selftest_sqrt, dim3((unsigned)((n + 255) / 256)), dim3(256), 0,
                       (hipStream_t)stream, x, n, out_hw);
    HIP_TRY(hipGetLastError());
    return KTM_OK;
}

extern "C" int ktm_selftest_traveltime(const float *in4, int64_t n, float dt, int ns, float *q_fast,
                                       float *q_ref, int32_t *t_kernel, int32_t *t_ref, void *stream)
{
    if (n < 0 || (n && (!in4 || !q_fast || !q_ref || !t_kernel || !t_ref))) return KTM_ERR_ARG;
    if (!n) return KTM_OK;
    hipLaunchKernelGGL(k_selftest_traveltime, dim3((unsigned)((n + 255) / 256)), dim3(256), 0,
                       (hipStream_t)stream, (const float4 *)in4, n, dt, ns, q_fast, q_ref,
                       t_kernel, t_ref);
    HIP_TRY(hipGetLastError());
    return KTM_OK;
}


/* ------------------------------------------------------------------------ */
/* adjoint pair: demigration (L) and the post-stack transpose                */
/* ------------------------------------------------------------------------ */
/* Per-pixel velocity tables for trace-parallel kernels: v2 and rs =
 * RN(1/(v2 dt^2)) of every (column, sample), the values k_migrate derives
 * per lane. */
__global__ void k_pixel_velocity(const float *__restrict__ vel2, const int32_t *__restrict__ velx,
                                 const int32_t *__restrict__ vely, const int32_t *__restrict__ t1,
                                 int nxp, int nyp, int ns, int nvel_xl, int nsm, float dt,
                                 float *__restrict__ pv2, float *__restrict__ prs)
{
    const int64_t n = (int64_t)nxp * nyp * ns;
    for (int64_t e = (int64_t)blockIdx.x * blockDim.x + threadIdx.x; e < n;
         e += (int64_t)gridDim.x * blockDim.x) {
        const int t = (int)(e % ns);
        const int64_t c = e / ns;
        const int X = (int)(c / nyp), Y = (int)(c % nyp);
        const float v2 = vel2[((int64_t)velx[X] * nvel_xl + vely[Y]) * nsm + t1[t]];
        pv2[e] = v2;
        prs[e] = fast_rs(v2, dt);
    }
}

#define DM_TB 8                      /* traces per demigration workgroup, at most */

struct DemigArgs {
    const float *km;                 /* image [nxp][nyp][nh][ns] */
    const ktm_trace_geom *geom;
    const int32_t *order;            /* trace order (workgroup b takes order[b*tb ...]) or NULL */
    int64_t ntr;
    float *traces;                   /* out [ntr][ns] */
    const float *pv2, *prs, *t2, *w, *colx, *coly, *slab_cx, *slab_cy, *slab_x2thr;
    const uint32_t *amax;            /* bits of max |km| (k_absmax) */
    float ap_x2thr, aperture, dt, rdt;
    int ns, nh, nxp, nyp, nx, ny, jobs_y, tb;
    int lg_terms;                    /* 2^lg_terms >= nxp * nyp * ns, terms per g[T] at most */
};

/* *out = bits of max |x| over n floats (out zeroed by k_zero_u32 first;
 * non-negative floats order like their bit patterns) */
__global__ void k_zero_u32(uint32_t *out) { *out = 0u; }
__global__ void k_absmax(const float *__restrict__ x, int64_t n, uint32_t *out)
{
    uint32_t m = 0u;
    for (int64_t e = (int64_t)blockIdx.x * blockDim.x + threadIdx.x; e < n;
         e += (int64_t)gridDim.x * blockDim.x)
        m = max(m, __float_as_uint(x[e]) & 0x7fffffffu);
    for (int d = KTM_WAVE / 2; d >= 1; d >>= 1) m = max(m, (uint32_t)__shfl_xor((int)m, d));
    if ((threadIdx.x & (KTM_WAVE - 1)) == 0) atomicMax(out, m);
}

/* Column index range whose centres can be within the aperture of a CMP at
 * coordinate c on an axis of n columns from col[0] to col[n-1]. */
__device__ __forceinline__ void aperture_span(const float *col, int n, float c, float aperture,
                                              int *lo, int *hi)
{
    *lo = 0;
    *hi = n - 1;
    if (n > 1) {
        const float step = (col[n - 1] - col[0]) / (float)(n - 1);
        if (step != 0.0f) {
            const float xc = (c - col[0]) / step, half = aperture / fabsf(step) + 2.0f;
            *lo = max(0, (int)floorf(xc - half));
            *hi = min(n - 1, (int)ceilf(xc + half));
        }
    }
}

/* Sort keys of the demigration trace order: offset class (out-of-class
 * traces last), then the CMP's image cell, row by row, so the traces one
 * workgroup takes lie close together and share their columns. */
__global__ void k_demig_keys(const ktm_trace_geom *__restrict__ geom, int64_t ntr, int nh,
                             const float *__restrict__ colx, int nxp, const float *__restrict__ coly,
                             int nyp, uint64_t *__restrict__ keys, int32_t *__restrict__ idx)
{
    const int64_t i = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
    if (i >= ntr) return;
    const ktm_trace_geom g = geom[i];
    const uint64_t cls = g.off >= 0 && g.off < nh ? (uint64_t)g.off : (uint64_t)nh;
    uint64_t cell[2] = {0, 0};
    const float *col[2] = {colx, coly};
    const int n[2] = {nxp, nyp};
    const float c[2] = {g.cmpx, g.cmpy};
    for (int d = 0; d < 2; ++d) {
        if (n[d] < 2) continue;
        const float step = (col[d][n[d] - 1] - col[d][0]) / (float)(n[d] - 1);
        if (step == 0.0f) continue;
        /* fminf/fmaxf return the number for a NaN operand */
        const float f = fminf(fmaxf(floorf((c[d] - col[d][0]) / step), -1.0f), (float)n[d]);
        cell[d] = (uint64_t)((int)f + 1);
    }
    keys[i] = (cls << 42) | (cell[1] << 21) | cell[0];
    idx[i] = (int32_t)i;
}

/* L, the exact transpose of migration with the CLOSED filter (km = L^T d).
 * A workgroup takes up to DM_TB traces, neighbours in the sort order of
 * k_demig_keys: the waves walk the image columns of the union of their
 * aperture boxes, gate every trace for the column exactly like k_migrate
 * (lane j = trace j), and for each passing trace and every sample t with
 * T(t) < ns add km[column][off][t] into that trace's g[T] (LDS atomics).
 * The column's image samples and rs are loaded once for all its traces.
 * g is 64-bit fixed point, km scaled by 2^E with max |km| < 2^(59 - E -
 * lg_terms), so no sum of at most 2^lg_terms terms passes 2^59 and the
 * filter's combinations of six stay below 2^62: integer LDS
 * atomics run at the rate of plain stores where float ones serialise over
 * the lanes (measured 22 vs 193 ms on the 48,000-trace line), and the
 * integer sums make L deterministic.  The quantum is 2^-E, at most
 * max |km| * 2^(lg_terms - 59).
 * Each trace is then the transpose of the 3-tap filter and of the sample
 * weights applied to its g.
 * Lane l owns samples l*K + ((k + l) mod K) of its wave's 64K-sample
 * segment, not k*64 + l: far from the CMP, T(t) grows slowly at early
 * times, and 64 consecutive samples land on a few T, whose atomics then
 * serialise; samples ~13 apart land on distinct T in distinct banks.
 * A sample whose fast-form q lies within the form's error bound of an
 * integer needs the reference arithmetic: each lane defers its first such
 * sample of a column and the wave resolves them after the column's traces,
 * every lane's together (the integer sums do not depend on the order of the
 * additions); a lane's further ones in the column take it inline.
 * FASTIDX: the index floor(min(q, ns + 1/2)) of the fast form alone, the
 * transpose of k_migrate's FASTIDX variant (floor(min(q, ns)) there: both
 * are ns for q >= ns, so they index the same samples). */
template <int K, int W, bool FASTIDX>
__global__ void __launch_bounds__(8 * KTM_WAVE) __attribute__((amdgpu_waves_per_eu(4, 4)))
k_demigrate(const DemigArgs a)
{
    constexpr int NCOL = 8 / W;
    extern __shared__ unsigned long long sg[];     /* [tb][ns + 4] fixed point */
    __shared__ int32_t sIdx[DM_TB];
    const int tid = threadIdx.x, lane = tid & (KTM_WAVE - 1), wv = tid / KTM_WAVE;
    const int cw = wv / W, seg = wv % W;
    const int ns = a.ns, ldg = ns + 4;
    const int64_t b0 = (int64_t)blockIdx.x * a.tb;
    const int nb = (int)min((int64_t)a.tb, a.ntr - b0);
    for (int e = tid; e < a.tb * ldg; e += 8 * KTM_WAVE) sg[e] = 0ull;
    const float amax = __uint_as_float(*a.amax);
    int emax = 0;                          /* max |km| < 2^emax */
    if (amax > 0.0f) (void)frexpf(amax, &emax);
    const int E = 59 - emax - a.lg_terms;
    /* lane j < nb of every wave holds trace j of the workgroup */
    ktm_trace_geom g;
    memset(&g, 0, sizeof g);
    bool valid = false;
    int gi = 0;
    int xlo = a.nxp, xhi = -1, ylo = a.nyp, yhi = -1;
    if (lane < nb) {
        gi = a.order ? a.order[b0 + lane] : (int)(b0 + lane);
        g = a.geom[gi];
        valid = g.off >= 0 && g.off < a.nh;
        if (valid) {
            aperture_span(a.colx, a.nxp, g.cmpx, a.aperture, &xlo, &xhi);
            aperture_span(a.coly, a.nyp, g.cmpy, a.aperture, &ylo, &yhi);
        }
    }
    if (tid < nb) sIdx[tid] = gi;
    for (int m = KTM_WAVE / 2; m >= 1; m >>= 1) {
        xlo = min(xlo, __shfl_xor(xlo, m));
        xhi = max(xhi, __shfl_xor(xhi, m));
        ylo = min(ylo, __shfl_xor(ylo, m));
        yhi = max(yhi, __shfl_xor(yhi, m));
    }
    __syncthreads();
    const float nsh = (float)ns + 0.5f;   /* clamp of q */
    int rot = lane % K;
    const int tbase = seg * K * KTM_WAVE + lane * K;
    auto slot_t = [&](int k) __attribute__((always_inline)) {   /* sample of slot k */
        return tbase + (k + rot < K ? k + rot : k + rot - K);
    };
    /* past the trace, tq puts q at ns + 1: clamped to ns + 0.5, T = ns, no add */
    const float tq_past = 0.25f * ((float)ns + 1.0f) * ((float)ns + 1.0f);
    float tq[K];
#pragma unroll
    for (int k = 0; k < K; ++k) {
        const int t = slot_t(k);
        tq[k] = t < ns ? fast_tq(t) : tq_past;
    }
    const int nbx = xhi - xlo + 1, nby = yhi - ylo + 1;
    const int64_t ncols = nbx > 0 && nby > 0 ? (int64_t)nbx * nby : 0;
    for (int64_t cidx = cw; cidx < ncols; cidx += NCOL) {
        const int X = xlo + (int)(cidx / nby), Y = ylo + (int)(cidx % nby);
        const float IX = a.colx[X], IY = a.coly[Y];
        const int job = (X / a.nx) * a.jobs_y + (Y / a.ny);
        float g_is = 0.0f, g_ir = 0.0f;
        bool gp = false;
        if (valid) {
            const float ex = a.slab_cx[job] - g.cmpx, ey = a.slab_cy[job] - g.cmpy;
            const float cx = IX - g.cmpx, cy = IY - g.cmpy;
            gp = ex * ex + ey * ey < a.slab_x2thr[job] && cx * cx + cy * cy < a.ap_x2thr;
            const float ax = IX - g.sxt, ay = IY - g.syt, bx = IX - g.gxt, by = IY - g.gyt;
            g_is = ax * ax + ay * ay;
            g_ir = bx * bx + by * by;
        }
        uint64_t pm = __ballot(gp);
        if (!pm) continue;
        /* the slot samples and their addresses are recomputed where used
         * rather than held in registers across the loops (an empty asm
         * hides that rot is loop-invariant) */
        asm volatile("" : "+v"(rot));
        const int64_t pixc = ((int64_t)X * a.nyp + Y) * ns;
        float rs[K];
        long long mq[K];                   /* km samples, fixed point */
#pragma unroll
        for (int k = 0; k < K; ++k) {
            const int t = slot_t(k);
            rs[k] = t < ns ? a.prs[pixc + t] : 0.0f;
            mq[k] = 0;
        }
        int cur = -1;                      /* offset class whose samples mq holds */
        /* one ambiguous sample per lane deferred to the end of the column:
         * its slot | trace << 8 (-1: none) and its fixed-point value */
        int qkj = -1;
        long long qv = 0;
        while (pm) {
            const int j = __builtin_ctzll(pm);
            pm &= pm - 1ull;
            const float dIS2 = __int_as_float(__builtin_amdgcn_readlane(__float_as_int(g_is), j));
            const float dIR2 = __int_as_float(__builtin_amdgcn_readlane(__float_as_int(g_ir), j));
            const int offj = __builtin_amdgcn_readlane(g.off, j);
            asm volatile("" : "+v"(rot));
            if (offj != cur) {
                cur = offj;
                const float *mk = a.km + (((int64_t)X * a.nyp + Y) * a.nh + offj) * ns;
#pragma unroll
                for (int k = 0; k < K; ++k) {
                    const int t = slot_t(k);
                    mq[k] = t < ns ? (long long)ldexpf(mk[t], E) : 0;
                }
            }
            unsigned long long *gl = sg + j * ldg;
#pragma unroll
            for (int k = 0; k < K; ++k) {
                const float qc = fast_clamp(fast_qs(dIS2, dIR2, rs[k], tq[k]), nsh);
                int T = fast_T(qc);
                if (!FASTIDX && fast_ambiguous(qc, KTM_FAST_EPS_S)) {   /* never past the trace */
                    if (qkj < 0) {
                        /* deferred: the integer sums do not depend on the
                         * order of their additions */
                        qkj = k | (j << 8);
                        qv = mq[k];
                        T = ns;
                    } else {
                        const int t = slot_t(k);
                        const float v2 = a.pv2[pixc + t];
                        T = exact_T(dIS2, dIR2, a.t2[t], v2, 1.0f / v2, a.dt, a.rdt, ns);
                    }
                }
                /* T = ns (past the trace, or deferred) adds into the pad
                 * g[ns], which the transpose below never reads: no compare
                 * and no branch per slot */
                atomicAdd(&gl[T], (unsigned long long)mq[k]);
            }
        }
        if (!FASTIDX && __ballot(qkj >= 0)) {
            /* the deferred samples of the column, every lane's together:
             * one reference-arithmetic pass instead of one per ambiguous
             * lane and slot (the trace's distances from its lane, all lanes
             * active for the shuffle) */
            const int jq = qkj >= 0 ? qkj >> 8 : 0;
            const float dIS2 = __shfl(g_is, jq), dIR2 = __shfl(g_ir, jq);
            if (qkj >= 0) {
                const int t = slot_t(qkj & 0xff);
                const float v2 = a.pv2[pixc + t];
                const int T = exact_T(dIS2, dIR2, a.t2[t], v2, 1.0f / v2, a.dt, a.rdt, ns);
                if (T < ns) atomicAdd(&sg[jq * ldg + T], (unsigned long long)qv);
            }
        }
    }
    __syncthreads();
    /* transpose of F = closed_tap(x), x[t] = tr[min(t, ns-1)] * w[t], in
     * double from the fixed-point sums */
    const double q = ldexp(1.0, -E);
    for (int e = tid; e < nb * ns; e += 8 * KTM_WAVE) {
        const int j = e / ns, t = e - j * ns;
        const long long *gl = reinterpret_cast<const long long *>(sg + j * ldg);
        double acc = 0.0;
        const int nt_ = t == ns - 1 ? 3 : 1;   /* tr[ns-1] also feeds x[ns], x[ns+1] */
        for (int u = 0; u < nt_; ++u) {
            const int tt = t + u;
            long long xb = 0;                  /* exact: |g| < 2^59, weights summing to <= 6 */
            if (tt >= 2 && tt < ns) xb += gl[tt];
            if (tt - 1 >= 2 && tt - 1 < ns) xb += 2 * gl[tt - 1];
            if (tt - 2 >= 2 && tt - 2 < ns) xb += gl[tt - 2];
            if (tt == 0) xb += 4 * gl[0] + 2 * gl[1];
            if (tt == 1) xb += 4 * gl[0] + 2 * gl[1];
            if (tt == 2) xb += 2 * gl[0] + 2 * gl[1];
            if (tt == 3) xb += gl[1];
            acc += (double)a.w[tt] * ((double)xb * q);
        }
        a.traces[(int64_t)sIdx[j] * ns + t] = (float)acc;
    }
}

/* Transpose of the post-stack chain image = P0 H C D km of Kirchhoff.c:
 * 1457-1479 (D: first difference with d[0] = 0, C: -dt^2 t / fold[t],
 * H: the even circular low-pass, P0: out[0] = 0):  km_bar = D^T C H P0 img.
 * H is symmetric (h even), so w = H P0 img is the post-stack convolution of
 * img itself (k_postfft<PostImg> or k_postmx<PostImg>, written into out);
 * this kernel finishes a
 * trace per workgroup in place: uu[t] = C[t] w[t] (FP64, uu[0] = uu[ns] = 0),
 * out[t] = uu[t] - uu[t + 1]. */
struct PostImg {
    static __device__ __forceinline__ double x(const float *im, const float *, int u, float)
    {
        return (double)im[u];
    }
};

__global__ void __launch_bounds__(PS_NT) k_poststack_adjoint_tail(const float *__restrict__ fold,
                                                                  int ns, float dt, float *out)
{
    extern __shared__ double uu[];                     /* [ns + 1] */
    const int64_t tr = blockIdx.x;
    float *o = out + tr * ns;
    const float *f = fold + tr * ns;
    for (int t = threadIdx.x; t <= ns; t += PS_NT) {
        double v = 0.0;
        if (t >= 1 && t < ns) v = -(double)dt * (double)dt * (double)t / (double)f[t] * (double)o[t];
        uu[t] = v;
    }
    __syncthreads();
    for (int t = threadIdx.x; t < ns; t += PS_NT) o[t] = (float)(uu[t] - uu[t + 1]);
}

/* ------------------------------------------------------------------------ */
/* host glue                                                                 */
/* The velocity check of ktm_plan_create on the device: every sample an image
 * column reads (column x, y at grid node velx[x], vely[y], time index t1[t])
 * must be > 0 and < 3e19, whose square stays finite (a velocity <= 0 or NaN
 * would divide by zero at Kirchhoff.c:1381); a failing sample sets *bad. */
__global__ void k_velocity_check(const float *__restrict__ vel, const int32_t *__restrict__ velx,
                                 const int32_t *__restrict__ vely, const int32_t *__restrict__ t1,
                                 int nxp, int nyp, int ns, int nvel_xl, int nsm, uint32_t *bad)
{
    const int64_t n = (int64_t)nxp * nyp * ns;
    for (int64_t e = (int64_t)blockIdx.x * blockDim.x + threadIdx.x; e < n;
         e += (int64_t)gridDim.x * blockDim.x) {
        const int t = (int)(e % ns);
        const int64_t c = e / ns;
        const int x = (int)(c / nyp), y = (int)(c % nyp);
        const float v = vel[((int64_t)velx[x] * nvel_xl + vely[y]) * nsm + t1[t]];
        if (!(v > 0.0f) || !(v < 3.0e19f)) *bad = 1u;
    }
}

/* v -> v^2 in place (Kirchhoff.c:617), one FP32 product each */
__global__ void k_square(float *__restrict__ v, int64_t n)
{
    for (int64_t e = (int64_t)blockIdx.x * blockDim.x + threadIdx.x; e < n;
         e += (int64_t)gridDim.x * blockDim.x)
        v[e] = v[e] * v[e];
}

/* ------------------------------------------------------------------------ */
template <typename T>
static int upload(T **dst, const T *src, size_t n)
{
    *dst = nullptr;
    if (hipMalloc((void **)dst, sizeof(T) * (n ? n : 1)) != hipSuccess) return KTM_ERR_NOMEM;
    if (n && hipMemcpy(*dst, src, sizeof(T) * n, hipMemcpyHostToDevice) != hipSuccess)
        return KTM_ERR_DEVICE;
    return KTM_OK;
}

extern "C" void ktm_plan_destroy(ktm_plan *p)
{
    if (!p) return;
    void *bufs[] = {p->d_t2, p->d_t1, p->d_colx, p->d_coly, p->d_velx, p->d_vely,
                    p->d_slab_cx, p->d_slab_cy, p->d_slab_x2thr, p->d_vel2, p->d_w, p->d_hconv2,
                    p->d_stage, p->d_pix_v2, p->d_pix_rs, p->d_dm_keys, p->d_dm_idx,
                    p->d_dm_tmp, p->d_dm_amax, p->d_parts, p->d_pf_g, p->d_pf_tw};
    for (void *b : bufs)
        if (b) (void)hipFree(b);
    ktm_tables_free(&p->tb);
    free(p);
}

extern "C" int ktm_plan_create(ktm_plan **out, const ktm_config *cfg, int ns, float dt, int nsm,
                               float dtm, const float *vel)
{
    if (!out || !cfg || !vel) return KTM_ERR_ARG;
    *out = nullptr;
    if (ns < 4 || ns > KTM_MAX_NS) return KTM_ERR_UNSUPPORTED;
    ktm_plan *p = (ktm_plan *)calloc(1, sizeof(ktm_plan));
    if (!p) return KTM_ERR_NOMEM;
    p->cfg = *cfg;
    int st = ktm_tables_build(cfg, ns, dt, nsm, dtm, &p->tb);
    if (st != KTM_OK) { free(p); return st; }
    if (hipGetDevice(&p->device) != hipSuccess) { ktm_plan_destroy(p); return KTM_ERR_DEVICE; }
    if ((st = scratch_address(p->device, &p->d_scratch)) != KTM_OK) {
        ktm_plan_destroy(p);
        return st;
    }
    p->row = ((int64_t)ns + 1 + 7) / 8 * 8;
    const ktm_tables *tb = &p->tb;
    st = upload(&p->d_t2, tb->t2, (size_t)ns);
    if (st == KTM_OK) st = upload(&p->d_t1, tb->t1, (size_t)ns);
    if (st == KTM_OK) st = upload(&p->d_colx, tb->colx, (size_t)tb->s.nxp);
    if (st == KTM_OK) st = upload(&p->d_coly, tb->coly, (size_t)tb->s.nyp);
    if (st == KTM_OK) st = upload(&p->d_velx, tb->velx, (size_t)tb->s.nxp);
    if (st == KTM_OK) st = upload(&p->d_vely, tb->vely, (size_t)tb->s.nyp);
    if (st == KTM_OK) st = upload(&p->d_slab_cx, tb->slab_cx, (size_t)tb->s.no_job);
    if (st == KTM_OK) st = upload(&p->d_slab_cy, tb->slab_cy, (size_t)tb->s.no_job);
    if (st == KTM_OK) st = upload(&p->d_slab_x2thr, tb->slab_x2thr, (size_t)tb->s.no_job);
    /* the velocity model: checked where the image reads it, then squared
     * (Kirchhoff.c:617), on the device */
    const size_t nv = (size_t)cfg->nvel_ilines * cfg->nvel_xlines * nsm;
    if (st == KTM_OK) st = upload(&p->d_vel2, vel, nv);
    if (st == KTM_OK) {
        uint32_t *d_bad = nullptr, bad = 0;
        if (hipMalloc((void **)&d_bad, sizeof(uint32_t)) != hipSuccess) {
            st = KTM_ERR_NOMEM;
        } else {
            const int64_t nchk = (int64_t)tb->s.nxp * tb->s.nyp * ns;
            const unsigned g = (unsigned)(nchk / 256 < 8192 ? (nchk + 255) / 256 : 8192);
            hipLaunchKernelGGL(k_zero_u32, dim3(1), dim3(1), 0, 0, d_bad);
            hipLaunchKernelGGL(k_velocity_check, dim3(g), dim3(256), 0, 0, p->d_vel2, p->d_velx,
                               p->d_vely, p->d_t1, tb->s.nxp, tb->s.nyp, ns, cfg->nvel_xlines, nsm,
                               d_bad);
            hipLaunchKernelGGL(k_square, dim3(8192), dim3(256), 0, 0, p->d_vel2, (int64_t)nv);
            if (hipGetLastError() != hipSuccess ||
                hipMemcpy(&bad, d_bad, sizeof(uint32_t), hipMemcpyDeviceToHost) != hipSuccess)
                st = KTM_ERR_DEVICE;
            (void)hipFree(d_bad);
            if (st == KTM_OK && bad) st = KTM_ERR_VELOCITY;
        }
    }
    if (st == KTM_OK) st = upload(&p->d_w, tb->w, (size_t)ns + 2);
    if (st == KTM_OK) {
        double *h2 = (double *)malloc(sizeof(double) * 2 * (size_t)ns);
        if (!h2) st = KTM_ERR_NOMEM;
        for (int m = 0; st == KTM_OK && m < 2 * ns; ++m) h2[m] = tb->hconv[m < ns ? m : m - ns];
        if (st == KTM_OK) st = upload(&p->d_hconv2, h2, 2 * (size_t)ns);
        free(h2);
    }
    if (st == KTM_OK && (p->pf_n = postfft_size(ns)) != 0) {
        const size_t n = (size_t)p->pf_n;
        double *g = (double *)malloc(sizeof(double) * 4 * n);
        if (!g) st = KTM_ERR_NOMEM;
        if (st == KTM_OK) {
            postfft_host_tables(tb->hconv, ns, p->pf_n, g, g + 2 * n);
            st = upload(&p->d_pf_g, reinterpret_cast<const double2 *>(g), n);
            if (st == KTM_OK) st = upload(&p->d_pf_tw, reinterpret_cast<const double2 *>(g + 2 * n), n);
        }
        free(g);
    }
    if (st == KTM_OK) st = parts_workspace(p);
    if (st != KTM_OK) { ktm_plan_destroy(p); return st; }
    *out = p;
    return KTM_OK;
}

extern "C" int ktm_plan_setup(const ktm_plan *p, ktm_setup *out)
{
    if (!p || !out) return KTM_ERR_ARG;
    *out = p->tb.s;
    return KTM_OK;
}

extern "C" int64_t ktm_plan_trace_row(const ktm_plan *p) { return p ? p->row : 0; }


extern "C" int64_t ktm_plan_image_size(const ktm_plan *p)
{
    if (!p) return 0;
    return (int64_t)p->tb.s.nxp * p->tb.s.nyp * p->tb.s.nh * p->tb.s.ns;
}

static GeoArgs geo_args(const ktm_plan *p)
{
    GeoArgs g;
    g.ox = p->cfg.ox; g.oy = p->cfg.oy; g.scale = p->cfg.scale_xy;
    g.oh = p->cfg.oh; g.dh = p->cfg.dh;
    g.sinA = p->tb.s.sinA; g.cosA = p->tb.s.cosA;
    return g;
}

extern "C" int ktm_geometry(const ktm_plan *p, const int32_t *coords, int64_t ntr,
                            ktm_trace_geom *geom, void *stream)
{
    if (!p || ntr < 0 || (ntr && (!coords || !geom))) return KTM_ERR_ARG;
    if (!ntr) return KTM_OK;
    hipStream_t s = (hipStream_t)stream;
    const int nt = 256;
    hipLaunchKernelGGL(k_geometry, dim3((unsigned)((ntr + nt - 1) / nt)), dim3(nt), 0, s,
                       geo_args(p), coords, ntr, geom);
    HIP_TRY(hipGetLastError());
    return KTM_OK;
}

/* prep kernels of one trace source (geometry already written) */
template <int SRC>
static int prep_launch(const ktm_plan *p, const TraceSrc src, int64_t ntr, int prep_mode,
                       int dtype, void *prepared, ktm_trace_geom *geom, hipStream_t s)
{
    const int ns = p->tb.s.ns;
    if (dtype == KTM_TRACE_F32) {
        if (prep_mode == KTM_PREP_REFERENCE) {
            hipLaunchKernelGGL(k_prep_ref<SRC>, dim3((unsigned)((ntr + KTM_WAVE - 1) / KTM_WAVE)),
                               dim3(KTM_WAVE), 0, s, src, ntr, ns, p->row, p->tb.s.dt, p->d_t2,
                               (float *)prepared);
        } else {
            if (ntr > 65535 * 1024LL) return KTM_ERR_UNSUPPORTED;
            const int nt = 256;
            for (int64_t i0 = 0; i0 < ntr; i0 += 65535) {
                const int64_t n = ntr - i0 < 65535 ? ntr - i0 : 65535;
                const TraceSrc si = {reinterpret_cast<const uint32_t *>(src.p) + i0 * src.rec_words,
                                     src.rec_words};
                hipLaunchKernelGGL(k_prep_closed<SRC>,
                                   dim3((unsigned)((p->row + nt - 1) / nt), (unsigned)n), dim3(nt),
                                   0, s, si, geom + i0, n, ns, p->row, p->d_w,
                                   (float *)prepared + i0 * p->row);
            }
        }
    } else {
        const int nt = KTM_WAVE;
        const dim3 grid((unsigned)((ntr + nt - 1) / nt));
        if (dtype == KTM_TRACE_BF16)
            hipLaunchKernelGGL((k_quantize<KTM_TRACE_BF16, SRC>), grid, dim3(nt), 0, s, src, ntr, ns,
                               p->row, (uint16_t *)prepared, geom);
        else
            hipLaunchKernelGGL((k_quantize<KTM_TRACE_F16, SRC>), grid, dim3(nt), 0, s, src, ntr, ns,
                               p->row, (uint16_t *)prepared, geom);
    }
    HIP_TRY(hipGetLastError());
    return KTM_OK;
}

static int prep_check(const ktm_plan *p, int prep_mode, int dtype)
{
    if (prep_mode != KTM_PREP_REFERENCE && prep_mode != KTM_PREP_CLOSED) return KTM_ERR_ARG;
    if (dtype != KTM_TRACE_F32 && dtype != KTM_TRACE_BF16 && dtype != KTM_TRACE_F16)
        return KTM_ERR_ARG;
    /* reduced-precision storage keeps the raw trace; its filter (the closed
     * form) is applied when a block is migrated */
    if (dtype != KTM_TRACE_F32 && prep_mode != KTM_PREP_CLOSED) return KTM_ERR_UNSUPPORTED;
    (void)p;
    return KTM_OK;
}

extern "C" int ktm_prep(const ktm_plan *p, const float *traces, const int32_t *coords,
                        int64_t ntr, int prep_mode, int dtype, void *prepared,
                        ktm_trace_geom *geom, void *stream)
{
    if (!p || ntr < 0) return KTM_ERR_ARG;
    if (!ntr) return KTM_OK;
    if (!traces || !coords || !prepared || !geom) return KTM_ERR_ARG;
    int st = prep_check(p, prep_mode, dtype);
    if (st != KTM_OK) return st;
    st = ktm_geometry(p, coords, ntr, geom, stream);
    if (st != KTM_OK) return st;
    const TraceSrc src = {traces, p->tb.s.ns};
    return prep_launch<SRC_F32>(p, src, ntr, prep_mode, dtype, prepared, geom, (hipStream_t)stream);
}

extern "C" int ktm_ingest(const ktm_plan *p, const void *raw, int64_t ntr, int format,
                          int prep_mode, int dtype, void *prepared, ktm_trace_geom *geom,
                          int32_t *coords, void *stream)
{
    if (!p || ntr < 0) return KTM_ERR_ARG;
    if (!ntr) return KTM_OK;
    if (!raw || !prepared || !geom || ((uintptr_t)raw & 3u)) return KTM_ERR_ARG;
    if (format != KTM_SEGY_IBM && format != KTM_SEGY_IEEE) return KTM_ERR_ARG;
    int st = prep_check(p, prep_mode, dtype);
    if (st != KTM_OK) return st;
    hipStream_t s = (hipStream_t)stream;
    const TraceSrc src = {raw, 60 + (int64_t)p->tb.s.ns};
    const int nt = 256;
    hipLaunchKernelGGL(k_geometry_segy, dim3((unsigned)((ntr + nt - 1) / nt)), dim3(nt), 0, s,
                       geo_args(p), (const uint32_t *)raw, src.rec_words, ntr, geom, coords);
    HIP_TRY(hipGetLastError());
    if (format == KTM_SEGY_IBM)
        return prep_launch<SRC_SEGY_IBM>(p, src, ntr, prep_mode, dtype, prepared, geom, s);
    return prep_launch<SRC_SEGY_IEEE>(p, src, ntr, prep_mode, dtype, prepared, geom, s);
}

/* raw SEG-Y records -> FP32 traces and header words: one workgroup per
 * record (the sample loop is contiguous, no index division) */
template <int SRC>
__global__ void k_segy_decode(const TraceSrc src, int64_t ntr, int ns, float *__restrict__ traces,
                              int32_t *__restrict__ coords)
{
    const int64_t i = blockIdx.x;
    if (i >= ntr) return;
    if (traces)
        for (int t = threadIdx.x; t < ns; t += blockDim.x)
            traces[i * ns + t] = load_sample<SRC>(src, i, t);
    if (coords && threadIdx.x < 4)
        coords[4 * i + threadIdx.x] = (int32_t)__builtin_bswap32(
            reinterpret_cast<const uint32_t *>(src.p)[i * src.rec_words + 18 + threadIdx.x]);
}

extern "C" int ktm_segy_decode(const void *raw, int64_t ntr, int ns, int format, float *traces,
                               int32_t *coords, void *stream)
{
    if (ntr < 0 || ns < 4) return KTM_ERR_ARG;
    if (!ntr) return KTM_OK;
    if (!raw || ((uintptr_t)raw & 3u)) return KTM_ERR_ARG;
    if (format != KTM_SEGY_IBM && format != KTM_SEGY_IEEE) return KTM_ERR_ARG;
    const TraceSrc src = {raw, 60 + (int64_t)ns};
    const int nt = 256;
    if (ntr > 0x7fffffffLL) return KTM_ERR_UNSUPPORTED;
    const dim3 grid((unsigned)ntr);
    hipStream_t s = (hipStream_t)stream;
    if (format == KTM_SEGY_IBM)
        hipLaunchKernelGGL(k_segy_decode<SRC_SEGY_IBM>, grid, dim3(nt), 0, s, src, ntr, ns, traces,
                           coords);
    else
        hipLaunchKernelGGL(k_segy_decode<SRC_SEGY_IEEE>, grid, dim3(nt), 0, s, src, ntr, ns, traces,
                           coords);
    HIP_TRY(hipGetLastError());
    return KTM_OK;
}

/* ------------------------------------------------------------------------ */
/* SEG-Y output on the device                                                */
/* ------------------------------------------------------------------------ */

/* IEEE -> IBM hex float, truncating (the encoder of the host's SEG-Y
 * writer): v = M 2^s with a 24-bit M (top bit 23), e2 = s + 24 (frexp's
 * exponent), e16 = ceil(e2 / 4), mantissa M >> (4 e16 - e2).  Zero, NaN and
 * inf encode as 0. */
__device__ __forceinline__ uint32_t float_to_ibm(float v)
{
    const uint32_t b = __float_as_uint(v);
    const uint32_t E = (b >> 23) & 0xffu, F = b & 0x7fffffu;
    if (E == 0xffu || (E == 0u && F == 0u)) return 0u;
    uint32_t M;
    int e2;
    if (E) {
        M = F | 0x800000u;
        e2 = (int)E - 126;
    } else {                                /* subnormal: normalise */
        const int sh = __builtin_clz(F) - 8;
        M = F << sh;
        e2 = -125 - sh;
    }
    const int e16 = (e2 + 3) >> 2;
    const uint32_t mant = M >> (4 * e16 - e2);
    return (b & 0x80000000u) | ((uint32_t)(e16 + 64) << 24) | mant;
}

/* samples of ntr records: records[r][60 + t] = big-endian word of
 * traces[r][t]; one workgroup per record */
template <int FMT>
__global__ void k_encode_samples(const float *__restrict__ traces, int64_t ntr, int ns,
                                 uint32_t *__restrict__ rec)
{
    const int64_t r = blockIdx.x;
    if (r >= ntr) return;
    for (int t = threadIdx.x; t < ns; t += blockDim.x) {
        const float v = traces[r * ns + t];
        const uint32_t w = FMT == KTM_SEGY_IBM ? float_to_ibm(v) : __float_as_uint(v);
        rec[r * (60 + (int64_t)ns) + 60 + t] = __builtin_bswap32(w);
    }
}

/* headers [ntr][60 words] into the records (or zeros) */
__global__ void k_copy_headers(const uint32_t *__restrict__ hdr, int64_t ntr, int ns,
                               uint32_t *__restrict__ rec)
{
    const int64_t e = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
    if (e >= ntr * 60) return;
    const int64_t r = e / 60;
    rec[r * (60 + (int64_t)ns) + (e - r * 60)] = hdr ? hdr[e] : 0u;
}

struct ImgHdrArgs {
    double start_il, start_xl, inc_il, inc_xl, il_dist, xl_dist, oh, dh;
    int ox_i, oy_i;                 /* (int)ox, (int)oy of Kirchhoff.c:1492-1493 */
    float cosA, sinA;
    int nx, ny, nyp, Nx, nh, ns, job, jobs_y;
    uint32_t dt_us;                 /* (unsigned short)(dt * 1e6), Kirchhoff.c:1447 */
};

/* image<job>.su, Kirchhoff.c:1437-1598: record r = (xc ny + yc) nh + ih of
 * the job's slab; one lane per record writes its 240-byte header (the
 * fields the reference sets, big-endian, the rest zero) */
__global__ void k_image_headers(const ImgHdrArgs a, uint32_t *__restrict__ rec)
{
    const int64_t r = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
    const int64_t nrec = (int64_t)a.nx * a.ny * a.nh;
    if (r >= nrec) return;
    const int ih = (int)(r % a.nh);
    const int64_t c = r / a.nh;
    const int xc = (int)(c / a.ny), yc = (int)(c % a.ny);
    const int X = xc + (a.job / a.jobs_y) * a.nx;
    const int Y = yc + (a.job % a.jobs_y) * a.ny;
    const double il = a.start_il + X * a.inc_il;                 /* Kirchhoff.c:1482 */
    const double xl = a.start_xl + Y * a.inc_xl;                 /* Kirchhoff.c:1483 */
    const int xm = (int)(il * a.il_dist);                        /* Kirchhoff.c:1485 */
    const int ym = (int)(xl * a.xl_dist);                        /* Kirchhoff.c:1486 */
    const float cx = (float)a.ox_i + a.cosA * (float)xm - a.sinA * (float)ym;
    const float cy = (float)a.oy_i + a.sinA * (float)xm + a.cosA * (float)ym;
    const int tracl = a.nx * a.ny * a.nh * a.job + 1 + (xc * a.ny + yc) * a.nh + ih;
    uint32_t *h = rec + r * (60 + (int64_t)a.ns);
    for (int w = 0; w < 60; ++w) {
        uint32_t v = 0u;
        switch (w) {
        case 0: v = __builtin_bswap32((uint32_t)tracl); break;                  /* tracl */
        case 5: v = __builtin_bswap32((uint32_t)(X + a.Nx * Y + 1)); break;     /* cdp */
        case 9: v = __builtin_bswap32((uint32_t)(int)ceil(a.oh + ih * a.dh)); break;  /* offset */
        case 18: case 19: case 20: case 21: v = __builtin_bswap32(1u); break;   /* sx..gy */
        case 28: v = __builtin_bswap32((uint32_t)(a.ns & 0xffff)); break;       /* ns, 114 */
        case 29: v = __builtin_bswap32(a.dt_us << 16); break;                   /* dt, 116 */
        case 45: v = __builtin_bswap32((uint32_t)(int)cx); break;               /* cdpx */
        case 46: v = __builtin_bswap32((uint32_t)(int)cy); break;               /* cdpy */
        case 47: v = __builtin_bswap32((uint32_t)(int)il); break;               /* iline */
        case 48: v = __builtin_bswap32((uint32_t)(int)xl); break;               /* xline */
        case 51: case 55: case 56: case 58:     /* mess: 0 in the first record, */
            v = r > 0 ? __builtin_bswap32(9999u) : 0u; break;   /* then 9999 (:1224, 1589) */
        case 59: v = __builtin_bswap32(9999u); break;                           /* :1589-1590 */
        default: break;
        }
        h[w] = v;
    }
}

/* samples of image<job>.su: record r of the slab <- image[X][Y][ih][t] */
/* the samples of image<job>.su: one workgroup per record r, its image row
 * byte-swapped into place; 16-byte vectors when ns % 4 == 0 (image rows and
 * records then start 16-byte aligned: 240-byte header, 4 (60 + ns)-byte
 * records) */
template <bool V4>
__global__ void __launch_bounds__(128) k_image_samples(const float *__restrict__ image,
                                                        const ImgHdrArgs a,
                                                        uint32_t *__restrict__ rec)
{
    const int64_t r = blockIdx.x;
    const int ih = (int)(r % a.nh);
    const int64_t c = r / a.nh;
    const int64_t X = c / a.ny + (int64_t)(a.job / a.jobs_y) * a.nx;
    const int64_t Y = c % a.ny + (int64_t)(a.job % a.jobs_y) * a.ny;
    const float *src = image + ((X * a.nyp + Y) * a.nh + ih) * a.ns;
    uint32_t *dst = rec + r * (60 + (int64_t)a.ns) + 60;
    if (V4) {
        const int n4 = a.ns / 4;
        for (int i = threadIdx.x; i < n4; i += blockDim.x) {
            uint4 w = reinterpret_cast<const uint4 *>(src)[i];
            w.x = __builtin_bswap32(w.x);
            w.y = __builtin_bswap32(w.y);
            w.z = __builtin_bswap32(w.z);
            w.w = __builtin_bswap32(w.w);
            reinterpret_cast<uint4 *>(dst)[i] = w;
        }
    } else {
        for (int t = threadIdx.x; t < a.ns; t += blockDim.x)
            dst[t] = __builtin_bswap32(__float_as_uint(src[t]));
    }
}

extern "C" int ktm_segy_encode(const float *traces, int64_t ntr, int ns, int format,
                               const void *headers, void *records, void *stream)
{
    if (ntr < 0 || ns < 1) return KTM_ERR_ARG;
    if (!ntr) return KTM_OK;
    if (!traces || !records || ((uintptr_t)records & 3u) || ((uintptr_t)headers & 3u))
        return KTM_ERR_ARG;
    if (format != KTM_SEGY_IBM && format != KTM_SEGY_IEEE) return KTM_ERR_ARG;
    hipStream_t s = (hipStream_t)stream;
    const int nt = 256;
    if (ntr > 0x7fffffffLL || (ntr * 60 + nt - 1) / nt > 0x7fffffffLL) return KTM_ERR_UNSUPPORTED;
    hipLaunchKernelGGL(k_copy_headers, dim3((unsigned)((ntr * 60 + nt - 1) / nt)), dim3(nt), 0, s,
                       (const uint32_t *)headers, ntr, ns, (uint32_t *)records);
    if (format == KTM_SEGY_IBM)
        hipLaunchKernelGGL(k_encode_samples<KTM_SEGY_IBM>, dim3((unsigned)ntr), dim3(nt), 0, s,
                           traces, ntr, ns, (uint32_t *)records);
    else
        hipLaunchKernelGGL(k_encode_samples<KTM_SEGY_IEEE>, dim3((unsigned)ntr), dim3(nt), 0, s,
                           traces, ntr, ns, (uint32_t *)records);
    HIP_TRY(hipGetLastError());
    return KTM_OK;
}

extern "C" int ktm_image_records(const ktm_plan *p, const float *image, int job, void *records,
                                 void *stream)
{
    if (!p || !image || !records || ((uintptr_t)records & 3u)) return KTM_ERR_ARG;
    const ktm_setup *st = &p->tb.s;
    if (job < 0 || job >= st->no_job) return KTM_ERR_ARG;
    const ktm_config *c = &p->cfg;
    ImgHdrArgs a;
    a.start_il = c->start_iline; a.start_xl = c->start_xline;
    a.inc_il = c->inc_iline; a.inc_xl = c->inc_xline;
    a.il_dist = c->iline_distance; a.xl_dist = c->xline_distance;
    a.oh = c->oh; a.dh = c->dh;
    a.ox_i = (int)c->ox; a.oy_i = (int)c->oy;
    a.cosA = st->cosA; a.sinA = st->sinA;
    a.nx = st->nx; a.ny = st->ny; a.nyp = st->nyp; a.Nx = st->Nx; a.nh = st->nh; a.ns = st->ns;
    a.job = job; a.jobs_y = c->jobs_y;
    a.dt_us = (uint32_t)(uint16_t)(st->dt * 1000000);
    const int64_t nrec = (int64_t)a.nx * a.ny * a.nh;
    const int nt = 256;
    if (nrec > 0x7fffffffLL) return KTM_ERR_UNSUPPORTED;
    hipStream_t s = (hipStream_t)stream;
    hipLaunchKernelGGL(k_image_headers, dim3((unsigned)((nrec + nt - 1) / nt)), dim3(nt), 0, s, a,
                       (uint32_t *)records);
    const bool v4 = a.ns % 4 == 0 && ((uintptr_t)records & 15u) == 0 && ((uintptr_t)image & 15u) == 0;
    if (v4)
        hipLaunchKernelGGL(k_image_samples<true>, dim3((unsigned)nrec), dim3(128), 0, s, image, a,
                           (uint32_t *)records);
    else
        hipLaunchKernelGGL(k_image_samples<false>, dim3((unsigned)nrec), dim3(128), 0, s, image, a,
                           (uint32_t *)records);
    HIP_TRY(hipGetLastError());
    return KTM_OK;
}

extern "C" int ktm_image_reset(const ktm_plan *p, float *km, float *fold, float *comp, void *stream)
{
    if (!p || !km || !fold) return KTM_ERR_ARG;
    const int64_t n = ktm_plan_image_size(p);
    hipLaunchKernelGGL(k_image_reset, dim3(2048), dim3(256), 0, (hipStream_t)stream, km, fold, comp, n);
    HIP_TRY(hipGetLastError());
    return KTM_OK;
}

/* kernel choice: K sample slots per lane and W waves per column cover ns */
typedef void (*mig_fn)(const MigArgs);

/* B traces staged per batch (8: two workgroups per CU fit the LDS); accum
 * is KTM_ACCUM_PLAIN or KTM_ACCUM_KAHAN, optionally with KTM_INDEX_FAST */
template <int K, int W, int B_ = 8>
static mig_fn pick_accum(int accum, int *B)
{
    *B = B_;
    const bool kahan = (accum & ~KTM_INDEX_FAST) == KTM_ACCUM_KAHAN;
    if (accum & KTM_INDEX_FAST)
        return kahan ? k_migrate<K, W, B_, true, true> : k_migrate<K, W, B_, false, true>;
    return kahan ? k_migrate<K, W, B_, true, false> : k_migrate<K, W, B_, false, false>;
}

/* W waves share one image column (each owns K*64 consecutive samples);
 * a workgroup is 8 waves, so it covers 8/W columns. */
static mig_fn pick_migrate(int ns, int accum, int64_t wgs8, int *W, int *K, int *B)
{
    /* wgs8: workgroups of 8 columns (x offset classes) in the region.  An
     * image too small to give every CU two workgroups at the default layout
     * takes twice the waves per column with half the samples per lane: on
     * the 800-column, 1001-sample line 3.77 ms against 4.79 ms (the 1536-
     * sample layout gained nothing from 6 x 4 on a 1600-column line). */
    if (ns <= 512) {
        if (wgs8 < 512) { *W = 2; *K = 4; return pick_accum<4, 2>(accum, B); }
        *W = 1; *K = 8; return pick_accum<8, 1>(accum, B);
    }
    if (ns <= 1024) {
        if (2 * wgs8 < 512) { *W = 4; *K = 4; return pick_accum<4, 4>(accum, B); }
        *W = 2; *K = 8; return pick_accum<8, 2>(accum, B);
    }
    /* (the Kahan K = 12 variant runs at 0.97x the plain one; four waves of
     * K = 6 per column measured 0.72x, of K = 8 0.63x, before the Kahan sums
     * were blocked per batch) */
    if (ns <= 1536) { *W = 2; *K = 12; return pick_accum<12, 2>(accum, B); }
    if (ns <= 2048) { *W = 4; *K = 8; return pick_accum<8, 4>(accum, B); }
    if (ns <= 3072) { *W = 4; *K = 12; return pick_accum<12, 4>(accum, B); }
    /* one column per workgroup (8 waves of K*64 samples) for long traces,
     * with fewer traces staged per batch so the rows fit the LDS */
    if (ns <= 4096) { *W = 8; *K = 8; return pick_accum<8, 8>(accum, B); }
    *W = 8; *K = 12;
    return pick_accum<12, 8, 5>(accum, B);
}

/* Tiles of the region [rx0,rx1) x [ry0,ry1) at W waves per column: 8/W
 * columns, a row on a 2-D line, as square as possible in 3-D. */
static void tile_layout(const ktm_setup *s, int W, MigArgs *a)
{
    const int ncol = 8 / W;
    a->tx_cols = s->nyp == 1 ? ncol : (ncol == 8 ? 4 : ncol == 4 ? 2 : ncol);
    a->ty_cols = ncol / a->tx_cols;
    a->tiles_x = (a->rx1 - a->rx0 + a->tx_cols - 1) / a->tx_cols;
    const int tiles_y = (a->ry1 - a->ry0 + a->ty_cols - 1) / a->ty_cols;
    a->ntiles = a->tiles_x * tiles_y;
}

/* Compensated sums do not need the reference's order of additions, so a
 * block whose image region gives fewer than 10240 workgroups (20 per
 * workgroup slot, 256 CUs x 2) is split along its traces into parts, each
 * part a workgroup per tile with its own sums, merged afterwards.  On the
 * 1600-column, 1501-sample line (400 tiles) L^T took 33.7 ms unsplit, 27.6
 * with 7 parts, 25.2 with 12 and 24.2 with 24. */
#define KTM_PART_TARGET_WGS 10240
#define KTM_MAX_PARTS 32
#define KTM_PARTS_MAX_BYTES (1ull << 30)

static int parts_for(int64_t wgs)
{
    if (wgs >= KTM_PART_TARGET_WGS) return 1;
    const int64_t n = (KTM_PART_TARGET_WGS + wgs - 1) / wgs;
    return n > KTM_MAX_PARTS ? KTM_MAX_PARTS : (int)n;
}

/* the partial images of the compensated split, sized for the whole image
 * (a job region needs at most that), allocated once with the plan so that
 * neither migration entry point allocates them; none for a large image */
static int parts_workspace(ktm_plan *p)
{
    const ktm_setup *s = &p->tb.s;
    MigArgs a;
    memset(&a, 0, sizeof a);
    a.rx1 = s->nxp; a.ry1 = s->nyp;
    int W = 1, K = 8, B = 8;
    (void)pick_migrate(s->ns, KTM_ACCUM_KAHAN, ((int64_t)s->nxp * s->nyp + 7) / 8 * s->nh, &W, &K,
                       &B);
    tile_layout(s, W, &a);
    int n = parts_for((int64_t)a.ntiles * s->nh);
    const uint64_t per = 3ull * sizeof(float) * (uint64_t)s->nxp * s->nyp * s->nh * s->ns;
    if (n > 1 && (uint64_t)(n - 1) * per > KTM_PARTS_MAX_BYTES)
        n = 1 + (int)(KTM_PARTS_MAX_BYTES / per);
    p->parts_max = n > 1 ? n : 1;
    if (p->parts_max > 1 &&
        hipMalloc((void **)&p->d_parts, (size_t)(p->parts_max - 1) * per) != hipSuccess) {
        p->d_parts = nullptr;
        return KTM_ERR_NOMEM;
    }
    return KTM_OK;
}

/* FP32 filtered traces of a BF16/FP16 block, into stage (ntr * row floats)
 * or, when stage is NULL, into a workspace the plan owns and grows on demand
 * (the only allocation on the migration path) */
static int decode_block(const ktm_plan *cp, const void *prepared, int dtype,
                        const ktm_trace_geom *geom, int64_t ntr, hipStream_t s, float *stage,
                        const float **out)
{
    ktm_plan *p = const_cast<ktm_plan *>(cp);
    float *dst = stage;
    if (!dst) {
        const size_t need = (size_t)ntr * (size_t)p->row;
        if (need > p->stage_cap) {
            if (p->d_stage) HIP_TRY(hipFree(p->d_stage));
            p->d_stage = nullptr;
            p->stage_cap = 0;
            if (hipMalloc((void **)&p->d_stage, need * sizeof(float)) != hipSuccess)
                return KTM_ERR_NOMEM;
            p->stage_cap = need;
        }
        dst = p->d_stage;
    }
    const int nt = 256;
    for (int64_t i0 = 0; i0 < ntr; i0 += 65535) {
        const int64_t n = ntr - i0 < 65535 ? ntr - i0 : 65535;
        const dim3 grid((unsigned)((p->row + nt - 1) / nt), (unsigned)n);
        const TraceSrc src = {reinterpret_cast<const uint16_t *>(prepared) + i0 * p->row, p->row};
        if (dtype == KTM_TRACE_BF16)
            hipLaunchKernelGGL(k_prep_closed<SRC_Q_BF16>, grid, dim3(nt), 0, s, src, geom + i0, n,
                               p->tb.s.ns, p->row, p->d_w, dst + i0 * p->row);
        else
            hipLaunchKernelGGL(k_prep_closed<SRC_Q_F16>, grid, dim3(nt), 0, s, src, geom + i0, n,
                               p->tb.s.ns, p->row, p->d_w, dst + i0 * p->row);
    }
    HIP_TRY(hipGetLastError());
    *out = dst;
    return KTM_OK;
}

static int migrate_block(const ktm_plan *p, const void *prepared, int dtype, int prep_mode,
                         const ktm_trace_geom *geom, int64_t ntr, int accum, int job, float *km,
                         float *fold, float *comp, uint64_t *counter, void *stream, float *stage);

extern "C" int ktm_migrate_block(const ktm_plan *p, const void *prepared, int dtype, int prep_mode,
                                 const ktm_trace_geom *geom, int64_t ntr, int accum, int job,
                                 float *km, float *fold, float *comp, uint64_t *counter,
                                 void *stream)
{
    return migrate_block(p, prepared, dtype, prep_mode, geom, ntr, accum, job, km, fold, comp,
                         counter, stream, nullptr);
}

/* ---- the allocation-free entry point ----
 * The BF16/FP16 decode workspace of ktm_migrate_block_static is this array:
 * every GPU the library runs on holds its own copy (KTM_STATIC_STAGE_FLOATS
 * floats, 64 MiB) until the process ends.  Host code never passes the name
 * `scratch` to a kernel (on the host it is only a stand-in for the symbol):
 * each plan keeps the device address hipGetSymbolAddress gave for its device
 * at ktm_plan_create, looked up once per device. */
static __device__ __align__(256) float scratch[KTM_STATIC_STAGE_FLOATS];

#define KTM_MAX_DEVICES 64
static float *g_scratch_ptr[KTM_MAX_DEVICES];
static std::mutex g_scratch_mu;

/* the device address of `scratch` on device dev, which must be current */
static int scratch_address(int dev, float **out)
{
    std::lock_guard<std::mutex> lock(g_scratch_mu);
    if (dev < 0 || dev >= KTM_MAX_DEVICES) return KTM_ERR_UNSUPPORTED;
    if (!g_scratch_ptr[dev]) {
        void *ptr = nullptr;
        if (hipGetSymbolAddress(&ptr, HIP_SYMBOL(scratch)) != hipSuccess || !ptr)
            return KTM_ERR_DEVICE;
        g_scratch_ptr[dev] = reinterpret_cast<float *>(ptr);
    }
    *out = g_scratch_ptr[dev];
    return KTM_OK;
}

extern "C" int ktm_migrate_block_static(const ktm_plan *p, const void *prepared, int dtype,
                                        int prep_mode, const ktm_trace_geom *geom, int64_t ntr,
                                        int accum, int job, float *km, float *fold, float *comp,
                                        uint64_t *counter, void *stream)
{
    if (!p || ntr < 0) return KTM_ERR_ARG;
    if (!p->d_scratch) return KTM_ERR_DEVICE;
    if (dtype != KTM_TRACE_F32 &&
        (uint64_t)ntr * (uint64_t)p->row > (uint64_t)KTM_STATIC_STAGE_FLOATS)
        return KTM_ERR_UNSUPPORTED;
    return migrate_block(p, prepared, dtype, prep_mode, geom, ntr, accum, job, km, fold, comp,
                         counter, stream, p->d_scratch);
}

static int migrate_block(const ktm_plan *p, const void *prepared, int dtype, int prep_mode,
                         const ktm_trace_geom *geom, int64_t ntr, int accum, int job, float *km,
                         float *fold, float *comp, uint64_t *counter, void *stream, float *stage)
{
    if (!p || !km || !fold || ntr < 0) return KTM_ERR_ARG;
    if (!ntr) return KTM_OK;
    if (!prepared || !geom) return KTM_ERR_ARG;
    const int acc_mode = accum & ~KTM_INDEX_FAST;
    if (acc_mode != KTM_ACCUM_PLAIN && acc_mode != KTM_ACCUM_KAHAN) return KTM_ERR_ARG;
    if (acc_mode == KTM_ACCUM_KAHAN && !comp) return KTM_ERR_ARG;
    if (dtype != KTM_TRACE_F32 && dtype != KTM_TRACE_BF16 && dtype != KTM_TRACE_F16) return KTM_ERR_ARG;
    if (dtype != KTM_TRACE_F32 && prep_mode != KTM_PREP_CLOSED) return KTM_ERR_UNSUPPORTED;
    /* the relaxed index is an option of the CLOSED operator only: the FP32
     * reference prep needs the reference's index (its taps cancel) */
    if ((accum & KTM_INDEX_FAST) && prep_mode != KTM_PREP_CLOSED) return KTM_ERR_UNSUPPORTED;
    const ktm_setup *s = &p->tb.s;
    if (job >= s->no_job) return KTM_ERR_ARG;
    hipStream_t hs = (hipStream_t)stream;
    const float *F = reinterpret_cast<const float *>(prepared);
    if (dtype != KTM_TRACE_F32) {
        const int st = decode_block(p, prepared, dtype, geom, ntr, hs, stage, &F);
        if (st != KTM_OK) return st;
    }
    MigArgs a;
    memset(&a, 0, sizeof a);
    a.prep = F; a.geom = geom; a.ntr = ntr; a.row = p->row;
    a.km = km; a.fold = fold; a.comp = comp;
    a.vel2 = p->d_vel2; a.t1 = p->d_t1; a.velx = p->d_velx; a.vely = p->d_vely;
    a.t2 = p->d_t2; a.colx = p->d_colx; a.coly = p->d_coly;
    a.slab_cx = p->d_slab_cx; a.slab_cy = p->d_slab_cy; a.slab_x2thr = p->d_slab_x2thr;
    a.counter = (unsigned long long *)counter;
    a.ap_x2thr = p->tb.ap_x2thr;
    a.aperture = (float)p->cfg.aperture;
    a.dt = s->dt;
    a.rdt = 1.0f / s->dt;
    a.ns = s->ns; a.nh = s->nh; a.nxp = s->nxp; a.nyp = s->nyp; a.nx = s->nx; a.ny = s->ny;
    a.jobs_y = p->cfg.jobs_y; a.nvel_xl = p->cfg.nvel_xlines; a.nsm = s->nsm;
    if (job < 0) {
        a.rx0 = 0; a.ry0 = 0; a.rx1 = s->nxp; a.ry1 = s->nyp;
    } else { /* slab of Kirchhoff.c:1341-1347, job -> slice as in 287-295 */
        a.rx0 = (job / p->cfg.jobs_y) * s->nx; a.rx1 = a.rx0 + s->nx;
        a.ry0 = (job % p->cfg.jobs_y) * s->ny; a.ry1 = a.ry0 + s->ny;
    }
    int W = 1, K = 8, B = 8;
    const int64_t ncols_region = (int64_t)(a.rx1 - a.rx0) * (a.ry1 - a.ry0);
    mig_fn fn = pick_migrate(s->ns, accum, (ncols_region + 7) / 8 * s->nh, &W, &K, &B);
    tile_layout(s, W, &a);
    const int nthreads = 8 * KTM_WAVE;
    /* compensated sums of a small region: the traces in parts */
    int nparts = 1;
    a.part_n = ncols_region * s->nh * s->ns;
    a.parts = p->d_parts;
    if (acc_mode == KTM_ACCUM_KAHAN && p->parts_max > 1) {
        nparts = parts_for((int64_t)a.ntiles * s->nh);
        if (nparts > p->parts_max) nparts = p->parts_max;
        const int64_t nchunks = (ntr + nthreads - 1) / nthreads;
        if (nparts > nchunks) nparts = (int)nchunks;     /* no part without a chunk */
    }
    a.nparts = nparts;
    /* staged rows + geometry, selection list and counts, and: v^2 per sample
     * and the speculation queue (exact index) or the fold row V (fast index) */
    const bool fast = (accum & KTM_INDEX_FAST) != 0;
    const size_t tail = fast ? sizeof(float) * (size_t)p->row
                             : sizeof(float) * (size_t)(8 * K * KTM_WAVE) +
                                   sizeof(uint2) * (size_t)(8 * KTM_SPEC_Q * KTM_WAVE);
    const size_t lds = sizeof(float) * (size_t)(B * p->row + B * 8) +
                       sizeof(int) * (size_t)(nthreads + 8) + tail;
    if (lds > 160 * 1024) return KTM_ERR_UNSUPPORTED;
    /* runs of a row of tiles, but at least 8 runs per XCD */
    a.run = a.ntiles / 64 < a.tiles_x ? a.ntiles / 64 : a.tiles_x;
    if (a.run < 1) a.run = 1;
    const int64_t span = 8 * (int64_t)a.run;
    const int64_t nblk = ((int64_t)a.ntiles + span - 1) / span * span;
    if (nblk > 0x7fffffffLL) return KTM_ERR_UNSUPPORTED;
    const dim3 grid((unsigned)nblk, (unsigned)s->nh, (unsigned)nparts);
    HIP_TRY(hipFuncSetAttribute((const void *)fn, hipFuncAttributeMaxDynamicSharedMemorySize, (int)lds));
    hipLaunchKernelGGL(fn, grid, dim3(nthreads), lds, hs, a);
    if (nparts > 1) {
        const int64_t nb = (a.part_n + 255) / 256;
        hipLaunchKernelGGL(k_merge_parts, dim3((unsigned)(nb < 8192 ? nb : 8192)), dim3(256), 0, hs,
                           a, nparts);
    }
    if (counter)
        hipLaunchKernelGGL(k_count_outside, dim3((unsigned)((ntr + 255) / 256)), dim3(256), 0, hs, a);
    HIP_TRY(hipGetLastError());
    return KTM_OK;
}

extern "C" int ktm_poststack(const ktm_plan *p, const float *km, const float *fold, float *out,
                             void *stream)
{
    if (!p || !km || !fold || !out) return KTM_ERR_ARG;
    const ktm_setup *s = &p->tb.s;
    const int64_t ntraces = (int64_t)s->nxp * s->nyp * s->nh;
    const hipError_t e =
        p->pf_n ? postfft_launch<PostKtm>(km, fold, ntraces, s->ns, s->dt, p->d_pf_g, p->d_pf_tw,
                                             out, (hipStream_t)stream)
                   : postmx_launch<PostKtm>(km, fold, ntraces, s->ns, s->dt, p->d_hconv2, out,
                                            (hipStream_t)stream);
    return e == hipSuccess ? KTM_OK : e == hipErrorInvalidValue ? KTM_ERR_UNSUPPORTED : KTM_ERR_DEVICE;
}

extern "C" int ktm_device_alloc(void **ptr, size_t bytes)
{
    if (!ptr) return KTM_ERR_ARG;
    *ptr = nullptr;
    return hipMalloc(ptr, bytes ? bytes : 1) == hipSuccess ? KTM_OK : KTM_ERR_NOMEM;
}

extern "C" int ktm_device_free(void *ptr)
{
    return hipFree(ptr) == hipSuccess ? KTM_OK : KTM_ERR_DEVICE;
}

extern "C" int ktm_copy_to_device(void *dst, const void *src, size_t bytes)
{
    return hipMemcpy(dst, src, bytes, hipMemcpyHostToDevice) == hipSuccess ? KTM_OK : KTM_ERR_DEVICE;
}

extern "C" int ktm_copy_to_host(void *dst, const void *src, size_t bytes)
{
    return hipMemcpy(dst, src, bytes, hipMemcpyDeviceToHost) == hipSuccess ? KTM_OK : KTM_ERR_DEVICE;
}

extern "C" int ktm_device_sync(void)
{
    return hipDeviceSynchronize() == hipSuccess ? KTM_OK : KTM_ERR_DEVICE;
}

/* ---- host pipeline runtime (ktm_internal.h) ---- */
extern "C" int ktm_rt_pinned_alloc(void **p, size_t bytes)
{
    if (!p) return KTM_ERR_ARG;
    *p = nullptr;
    return hipHostMalloc(p, bytes ? bytes : 1, hipHostMallocDefault) == hipSuccess ? KTM_OK
                                                                                  : KTM_ERR_NOMEM;
}
extern "C" void ktm_rt_pinned_free(void *p)
{
    if (p) (void)hipHostFree(p);
}
extern "C" int ktm_rt_stream_create(void **s)
{
    if (!s) return KTM_ERR_ARG;
    *s = nullptr;
    hipStream_t h = nullptr;
    if (hipStreamCreate(&h) != hipSuccess) return KTM_ERR_DEVICE;
    *s = h;
    return KTM_OK;
}
extern "C" void ktm_rt_stream_destroy(void *s)
{
    if (s) (void)hipStreamDestroy((hipStream_t)s);
}
extern "C" int ktm_rt_stream_sync(void *s)
{
    return hipStreamSynchronize((hipStream_t)s) == hipSuccess ? KTM_OK : KTM_ERR_DEVICE;
}
extern "C" int ktm_rt_stream_wait(void *s, void *e)
{
    return hipStreamWaitEvent((hipStream_t)s, (hipEvent_t)e, 0) == hipSuccess ? KTM_OK
                                                                            : KTM_ERR_DEVICE;
}
extern "C" int ktm_rt_event_create(void **e)
{
    if (!e) return KTM_ERR_ARG;
    *e = nullptr;
    hipEvent_t h = nullptr;
    if (hipEventCreate(&h) != hipSuccess) return KTM_ERR_DEVICE;
    *e = h;
    return KTM_OK;
}
extern "C" void ktm_rt_event_destroy(void *e)
{
    if (e) (void)hipEventDestroy((hipEvent_t)e);
}
extern "C" int ktm_rt_event_record(void *e, void *s)
{
    return hipEventRecord((hipEvent_t)e, (hipStream_t)s) == hipSuccess ? KTM_OK : KTM_ERR_DEVICE;
}
extern "C" int ktm_rt_event_sync(void *e)
{
    return hipEventSynchronize((hipEvent_t)e) == hipSuccess ? KTM_OK : KTM_ERR_DEVICE;
}
extern "C" int ktm_rt_event_elapsed_ms(void *e0, void *e1, float *ms)
{
    return hipEventElapsedTime(ms, (hipEvent_t)e0, (hipEvent_t)e1) == hipSuccess ? KTM_OK
                                                                               : KTM_ERR_DEVICE;
}
extern "C" int ktm_rt_copy_h2d_async(void *dst, const void *src, size_t bytes, void *s)
{
    return hipMemcpyAsync(dst, src, bytes, hipMemcpyHostToDevice, (hipStream_t)s) == hipSuccess
               ? KTM_OK : KTM_ERR_DEVICE;
}
extern "C" int ktm_rt_copy_d2h_async(void *dst, const void *src, size_t bytes, void *s)
{
    return hipMemcpyAsync(dst, src, bytes, hipMemcpyDeviceToHost, (hipStream_t)s) == hipSuccess
               ? KTM_OK : KTM_ERR_DEVICE;
}

static int pixel_tables(const ktm_plan *cp, hipStream_t hs)
{
    ktm_plan *p = const_cast<ktm_plan *>(cp);
    if (p->d_pix_v2) return KTM_OK;
    const ktm_setup *s = &p->tb.s;
    const size_t n = (size_t)s->nxp * s->nyp * s->ns;
    if (hipMalloc((void **)&p->d_pix_v2, n * sizeof(float)) != hipSuccess) return KTM_ERR_NOMEM;
    if (hipMalloc((void **)&p->d_pix_rs, n * sizeof(float)) != hipSuccess) {
        (void)hipFree(p->d_pix_v2);
        p->d_pix_v2 = nullptr;
        return KTM_ERR_NOMEM;
    }
    hipLaunchKernelGGL(k_pixel_velocity, dim3(4096), dim3(256), 0, hs, p->d_vel2, p->d_velx,
                       p->d_vely, p->d_t1, s->nxp, s->nyp, s->ns, p->cfg.nvel_xlines, s->nsm,
                       s->dt, p->d_pix_v2, p->d_pix_rs);
    HIP_TRY(hipGetLastError());
    return KTM_OK;
}

/* the demigration trace order of n traces (sorted by k_demig_keys) into
 * plan-owned buffers; *order points at the sorted indices */
static int demig_order(const ktm_plan *cp, const ktm_trace_geom *geom, int64_t n, hipStream_t hs,
                       const int32_t **order)
{
    ktm_plan *p = const_cast<ktm_plan *>(cp);
    const ktm_setup *s = &p->tb.s;
    if (n > 0x7fffffffLL || s->nh >= (1 << 21)) return KTM_ERR_UNSUPPORTED;
    if ((size_t)n > p->dm_cap) {
        if (p->d_dm_keys) HIP_TRY(hipFree(p->d_dm_keys));
        if (p->d_dm_idx) HIP_TRY(hipFree(p->d_dm_idx));
        p->d_dm_keys = nullptr;
        p->d_dm_idx = nullptr;
        p->dm_cap = 0;
        if (hipMalloc((void **)&p->d_dm_keys, 2 * (size_t)n * sizeof(uint64_t)) != hipSuccess)
            return KTM_ERR_NOMEM;
        if (hipMalloc((void **)&p->d_dm_idx, 2 * (size_t)n * sizeof(int32_t)) != hipSuccess)
            return KTM_ERR_NOMEM;
        p->dm_cap = (size_t)n;
    }
    uint64_t *k_in = p->d_dm_keys, *k_out = p->d_dm_keys + n;
    int32_t *i_in = p->d_dm_idx, *i_out = p->d_dm_idx + n;
    size_t need = 0;
    HIP_TRY(rocprim::radix_sort_pairs(nullptr, need, k_in, k_out, i_in, i_out, (size_t)n, 0u, 64u,
                                      hs));
    if (need > p->dm_tmp_bytes) {
        if (p->d_dm_tmp) HIP_TRY(hipFree(p->d_dm_tmp));
        p->d_dm_tmp = nullptr;
        p->dm_tmp_bytes = 0;
        if (hipMalloc(&p->d_dm_tmp, need) != hipSuccess) return KTM_ERR_NOMEM;
        p->dm_tmp_bytes = need;
    }
    hipLaunchKernelGGL(k_demig_keys, dim3((unsigned)((n + 255) / 256)), dim3(256), 0, hs, geom, n,
                       s->nh, p->d_colx, s->nxp, p->d_coly, s->nyp, k_in, i_in);
    HIP_TRY(hipGetLastError());
    size_t have = p->dm_tmp_bytes;
    HIP_TRY(rocprim::radix_sort_pairs(p->d_dm_tmp, have, k_in, k_out, i_in, i_out, (size_t)n, 0u,
                                      64u, hs));
    *order = i_out;
    return KTM_OK;
}

static int demigrate(const ktm_plan *p, const float *km, const ktm_trace_geom *geom, int64_t ntr,
                     float *traces, bool fast, void *stream);

extern "C" int ktm_demigrate(const ktm_plan *p, const float *km, const ktm_trace_geom *geom,
                             int64_t ntr, float *traces, void *stream)
{
    return demigrate(p, km, geom, ntr, traces, false, stream);
}

extern "C" int ktm_demigrate_index(const ktm_plan *p, const float *km, const ktm_trace_geom *geom,
                                   int64_t ntr, float *traces, int index, void *stream)
{
    if (index != 0 && index != KTM_INDEX_FAST) return KTM_ERR_ARG;
    return demigrate(p, km, geom, ntr, traces, index == KTM_INDEX_FAST, stream);
}

static int demigrate(const ktm_plan *p, const float *km, const ktm_trace_geom *geom, int64_t ntr,
                     float *traces, bool fast, void *stream)
{
    if (!p || ntr < 0) return KTM_ERR_ARG;
    if (!ntr) return KTM_OK;
    if (!km || !geom || !traces) return KTM_ERR_ARG;
    hipStream_t hs = (hipStream_t)stream;
    int st = pixel_tables(p, hs);
    if (st != KTM_OK) return st;
    const ktm_setup *s = &p->tb.s;
    /* cell indices of the sort key take 21 bits per axis */
    if (s->nxp >= (1 << 21) - 2 || s->nyp >= (1 << 21) - 2) return KTM_ERR_UNSUPPORTED;
    DemigArgs a;
    memset(&a, 0, sizeof a);
    a.km = km; a.geom = geom; a.ntr = ntr; a.traces = traces;
    a.pv2 = p->d_pix_v2; a.prs = p->d_pix_rs; a.t2 = p->d_t2; a.w = p->d_w;
    a.colx = p->d_colx; a.coly = p->d_coly;
    a.slab_cx = p->d_slab_cx; a.slab_cy = p->d_slab_cy; a.slab_x2thr = p->d_slab_x2thr;
    a.ap_x2thr = p->tb.ap_x2thr; a.aperture = (float)p->cfg.aperture;
    a.dt = s->dt; a.rdt = 1.0f / s->dt;
    a.ns = s->ns; a.nh = s->nh; a.nxp = s->nxp; a.nyp = s->nyp; a.nx = s->nx; a.ny = s->ny;
    a.jobs_y = p->cfg.jobs_y;
    /* fixed-point scale of g: max |km| and a bound on the terms per g[T] */
    a.lg_terms = 0;
    while ((1ull << a.lg_terms) < (unsigned long long)s->nxp * s->nyp * s->ns) ++a.lg_terms;
    if (!p->d_dm_amax) {
        ktm_plan *pm = const_cast<ktm_plan *>(p);
        if (hipMalloc((void **)&pm->d_dm_amax, sizeof(uint32_t)) != hipSuccess) return KTM_ERR_NOMEM;
    }
    a.amax = p->d_dm_amax;
    const int64_t nimg = (int64_t)s->nxp * s->nyp * s->nh * s->ns;
    hipLaunchKernelGGL(k_zero_u32, dim3(1), dim3(1), 0, hs, p->d_dm_amax);
    hipLaunchKernelGGL(k_absmax, dim3(1024), dim3(256), 0, hs, km, nimg, p->d_dm_amax);
    /* traces per workgroup: DM_TB, fewer so that two workgroups share a CU
     * (g rows of 8 bytes per sample within 76 KB), or one beyond that */
    const size_t row = sizeof(unsigned long long) * (size_t)(s->ns + 4);
    a.tb = (int)((76 * 1024) / row);
    if (a.tb < 2) a.tb = (int)((152 * 1024) / row);
    if (a.tb > DM_TB) a.tb = DM_TB;
    if (a.tb < 1) return KTM_ERR_UNSUPPORTED;
    typedef void (*fn_t)(const DemigArgs);
    fn_t fn;
    if (s->ns <= 512) fn = fast ? k_demigrate<8, 1, true> : k_demigrate<8, 1, false>;
    else if (s->ns <= 1024) fn = fast ? k_demigrate<8, 2, true> : k_demigrate<8, 2, false>;
    else if (s->ns <= 1536) fn = fast ? k_demigrate<12, 2, true> : k_demigrate<12, 2, false>;
    else if (s->ns <= 2048) fn = fast ? k_demigrate<8, 4, true> : k_demigrate<8, 4, false>;
    else if (s->ns <= 3072) fn = fast ? k_demigrate<12, 4, true> : k_demigrate<12, 4, false>;
    else if (s->ns <= 4096) fn = fast ? k_demigrate<8, 8, true> : k_demigrate<8, 8, false>;
    else fn = fast ? k_demigrate<12, 8, true> : k_demigrate<12, 8, false>;
    const size_t lds = row * (size_t)a.tb;
    HIP_TRY(hipFuncSetAttribute((const void *)fn, hipFuncAttributeMaxDynamicSharedMemorySize, (int)lds));
    for (int64_t i0 = 0; i0 < ntr; i0 += 1 << 30) {
        const int64_t n = ntr - i0 < (1 << 30) ? ntr - i0 : (1 << 30);
        DemigArgs b = a;
        b.geom = geom + i0;
        b.traces = traces + i0 * s->ns;
        b.ntr = n;
        if ((st = demig_order(p, b.geom, n, hs, &b.order)) != KTM_OK) return st;
        const int64_t nblk = (n + a.tb - 1) / a.tb;
        hipLaunchKernelGGL(fn, dim3((unsigned)nblk), dim3(8 * KTM_WAVE), lds, hs, b);
    }
    HIP_TRY(hipGetLastError());
    return KTM_OK;
}

extern "C" int ktm_poststack_adjoint(const ktm_plan *p, const float *img, const float *fold,
                                     float *km_bar, void *stream)
{
    if (!p || !img || !fold || !km_bar) return KTM_ERR_ARG;
    const ktm_setup *s = &p->tb.s;
    const int64_t ntraces = (int64_t)s->nxp * s->nyp * s->nh;
    if (ntraces > 0x7fffffffLL) return KTM_ERR_UNSUPPORTED;
    hipStream_t hs = (hipStream_t)stream;
    const hipError_t e =
        p->pf_n ? postfft_launch<PostImg>(img, fold, ntraces, s->ns, s->dt, p->d_pf_g,
                                             p->d_pf_tw, km_bar, hs)
                   : postmx_launch<PostImg>(img, fold, ntraces, s->ns, s->dt, p->d_hconv2, km_bar,
                                            hs);
    if (e != hipSuccess) return e == hipErrorInvalidValue ? KTM_ERR_UNSUPPORTED : KTM_ERR_DEVICE;
    hipLaunchKernelGGL(k_poststack_adjoint_tail, dim3((unsigned)ntraces), dim3(PS_NT),
                       sizeof(double) * (size_t)(s->ns + 1), hs, fold, s->ns, s->dt, km_bar);
    HIP_TRY(hipGetLastError());
    return KTM_OK;
}
