/*
 * ktm_postfft.h -- the real circular convolution of ktm_postmx.h,
 *   out[j][t] = sum_u h[(t - u) mod ns] x[j][u]   (t = 0 stored as 0),
 * by FFT in FP64: with N the power of two >= max(256, 2 ns - 1), the kernel
 * wrapped into N points, g[d] = h[d] (0 <= d < ns) and g[N - d] = h[ns - d]
 * (0 < d < ns), zero between, makes the length-N circular convolution of the
 * zero-padded trace equal the length-ns one on t < ns:
 *   out = IDFT_N(G . DFT_N(x)) / N,   G = DFT_N(g),
 * the O(ns log ns) form of Kirchhoff.c:1465-1478 (FFTW forward, filter,
 * backward, / ns: here the 1/ns is in h) and EOM.c:1386-1413.  g is real,
 * so two traces go through one complex transform, x_a + i x_b, and come
 * out as its real and imaginary parts.
 *
 * k_postfft: radix-16 Stockham passes (the last one radix 2, 4 or 8 when
 * N is not a power of 16), in place in LDS: every thread loads the 16
 * points of its butterflies into registers, the workgroup synchronises,
 * and the results go to their Stockham positions.  N / 16 threads per
 * transform; a workgroup of 256 threads holds 4096 / N transforms (N <=
 * 4096) or one transform of 512 threads (N = 8192).  The first forward pass
 * reads the traces (the prologue Pro of ktm_postmx.h), the first inverse
 * pass multiplies by G, the last inverse pass writes the output (both with
 * consecutive threads on consecutive samples).  LDS rows are padded by one
 * double2 per 16 so the stride-16 stores of the first pass spread over the
 * banks.
 */
#ifndef KTM_POSTFFT_H
#define KTM_POSTFFT_H

#include <hip/hip_runtime.h>
#include <math.h>
#include <stdint.h>
#include <stdlib.h>

#define PF_MIN_LOGN 8
#define PF_MAX_LOGN 13

__device__ __forceinline__ double2 pf_add(double2 a, double2 b) { return make_double2(a.x + b.x, a.y + b.y); }
__device__ __forceinline__ double2 pf_sub(double2 a, double2 b) { return make_double2(a.x - b.x, a.y - b.y); }
__device__ __forceinline__ double2 pf_mul(double2 a, double2 w)
{
    return make_double2(__builtin_fma(a.x, w.x, -(a.y * w.y)), __builtin_fma(a.x, w.y, a.y * w.x));
}
/* a * (-i) forward, a * (+i) inverse */
template <bool INV> __device__ __forceinline__ double2 pf_rot(double2 a)
{
    return INV ? make_double2(-a.y, a.x) : make_double2(a.y, -a.x);
}
template <bool INV> __device__ __forceinline__ double2 pf_conj(double2 w)
{
    return INV ? make_double2(w.x, -w.y) : w;
}

template <bool INV> __device__ __forceinline__ void pf_dft2(double2 &a0, double2 &a1)
{
    const double2 t = a0;
    a0 = pf_add(t, a1);
    a1 = pf_sub(t, a1);
}

template <bool INV>
__device__ __forceinline__ void pf_dft4(double2 &a0, double2 &a1, double2 &a2, double2 &a3)
{
    const double2 t0 = pf_add(a0, a2), t1 = pf_sub(a0, a2);
    const double2 t2 = pf_add(a1, a3), t3 = pf_rot<INV>(pf_sub(a1, a3));
    a0 = pf_add(t0, t2);
    a2 = pf_sub(t0, t2);
    a1 = pf_add(t1, t3);
    a3 = pf_sub(t1, t3);
}

/* exp(-2 pi i m / 16), m = 0..9 */
__device__ __forceinline__ double2 pf_w16(int m)
{
    const double c1 = 0.92387953251128673848, s1 = 0.38268343236508978178;
    const double c2 = 0.70710678118654752440;
    switch (m) {
    case 0: return make_double2(1.0, 0.0);
    case 1: return make_double2(c1, -s1);
    case 2: return make_double2(c2, -c2);
    case 3: return make_double2(s1, -c1);
    case 4: return make_double2(0.0, -1.0);
    case 6: return make_double2(-c2, -c2);
    default: return make_double2(-c1, s1);   /* 9 */
    }
}

/* In-place DFT of R = 16, 8, 4 or 2 points v[o], v[o + 1], ...; the result
 * X[k] ends in v[o + pf_slot<R>(k)]. */
template <int R> __device__ __forceinline__ constexpr int pf_slot(int k)
{
    return R == 16 ? 4 * (k & 3) + (k >> 2) : R == 8 ? 4 * (k & 1) + (k >> 1) : k;
}

template <int R, bool INV> __device__ __forceinline__ void pf_dft(double2 *v)
{
    if (R == 2) {
        pf_dft2<INV>(v[0], v[1]);
    } else if (R == 4) {
        pf_dft4<INV>(v[0], v[1], v[2], v[3]);
    } else if (R == 8) {
        /* n = 4 n1 + n2: 2-point DFTs over n1, twiddle W8^(n2 k1), 4-point over n2 */
#pragma unroll
        for (int n2 = 0; n2 < 4; ++n2) pf_dft2<INV>(v[n2], v[4 + n2]);
#pragma unroll
        for (int n2 = 1; n2 < 4; ++n2) v[4 + n2] = pf_mul(v[4 + n2], pf_conj<INV>(pf_w16(2 * n2)));
#pragma unroll
        for (int k1 = 0; k1 < 2; ++k1) pf_dft4<INV>(v[4 * k1], v[4 * k1 + 1], v[4 * k1 + 2], v[4 * k1 + 3]);
    } else {
        /* n = 4 n1 + n2: 4-point DFTs over n1, twiddle W16^(n2 k1), 4-point over n2 */
#pragma unroll
        for (int n2 = 0; n2 < 4; ++n2) pf_dft4<INV>(v[n2], v[4 + n2], v[8 + n2], v[12 + n2]);
#pragma unroll
        for (int k1 = 1; k1 < 4; ++k1)
#pragma unroll
            for (int n2 = 1; n2 < 4; ++n2)
                v[4 * k1 + n2] = pf_mul(v[4 * k1 + n2], pf_conj<INV>(pf_w16(n2 * k1)));
#pragma unroll
        for (int k1 = 0; k1 < 4; ++k1) pf_dft4<INV>(v[4 * k1], v[4 * k1 + 1], v[4 * k1 + 2], v[4 * k1 + 3]);
    }
}

__device__ __forceinline__ int pf_pad(int i) { return i + (i >> 4); }

/* One Stockham pass of radix R, stride NS (points already combined), on the
 * transform of N points whose TPF = N / 16 threads include this one (index
 * tj).  MODE 0: LDS -> LDS; 1: the first forward pass, input from Pro;
 * 2: the first inverse pass, input times G; 3: the last inverse pass,
 * output to out. */
template <class Pro, int N, int R, int NS, bool INV, int MODE>
__device__ __forceinline__ void pf_pass(double2 *buf, int tj, const double2 *__restrict__ tw,
                                        const double2 *__restrict__ G, const float *in,
                                        const float *fold, int64_t tra, int64_t ntraces, int ns,
                                        float dt, float *out)
{
    constexpr int TPF = N / 16, NB = 16 / R, STEP = N / R;
    double2 v[16];
#pragma unroll
    for (int b = 0; b < NB; ++b) {
        const int j = tj + b * TPF;
#pragma unroll
        for (int r = 0; r < R; ++r) {
            const int u = j + r * STEP;
            if (MODE == 1) {
                double xa = 0.0, xb = 0.0;
                if (u >= 1 && u < ns) {
                    if (tra < ntraces) xa = Pro::x(in + tra * ns, fold + tra * ns, u, dt);
                    if (tra + 1 < ntraces) xb = Pro::x(in + (tra + 1) * ns, fold + (tra + 1) * ns, u, dt);
                }
                v[b * R + r] = make_double2(xa, xb);
            } else {
                v[b * R + r] = buf[pf_pad(u)];
                if (MODE == 2) v[b * R + r] = pf_mul(v[b * R + r], G[u]);
            }
        }
        if (NS > 1) {
            /* twiddle W_N^(r * (j mod NS) * N / (NS R)) of element r: the
             * powers 1, 2, 4, 8 from the table, the others their products */
            const int c = (j % NS) * (N / (NS * R));
            double2 w[16];
            w[1] = tw[c];
            if (R > 2) w[2] = tw[2 * c];
            if (R > 4) w[4] = tw[4 * c];
            if (R > 8) w[8] = tw[8 * c];
#pragma unroll
            for (int r = 3; r < R; ++r) {
                const int hi = r & (r - 1) ? (r & 8 ? 8 : r & 4 ? 4 : 2) : 0;
                if (hi) w[r] = pf_mul(w[hi], w[r - hi]);
            }
#pragma unroll
            for (int r = 1; r < R; ++r) v[b * R + r] = pf_mul(v[b * R + r], pf_conj<INV>(w[r]));
        }
        pf_dft<R, INV>(v + b * R);
    }
    if (MODE != 3) {
        __syncthreads();                    /* every load of the pass done */
#pragma unroll
        for (int b = 0; b < NB; ++b) {
            const int j = tj + b * TPF;
            const int d0 = (j / NS) * NS * R + j % NS;
#pragma unroll
            for (int r = 0; r < R; ++r) buf[pf_pad(d0 + r * NS)] = v[b * R + pf_slot<R>(r)];
        }
        __syncthreads();                    /* every store of the pass done */
    } else {
        /* last pass: NS R = N, so the destination is j + r N / R */
        constexpr double inv_n = 1.0 / (double)N;
#pragma unroll
        for (int b = 0; b < NB; ++b) {
            const int j = tj + b * TPF;
#pragma unroll
            for (int r = 0; r < R; ++r) {
                const int t = j + r * STEP;
                if (t < ns) {
                    const double2 y = v[b * R + pf_slot<R>(r)];
                    if (tra < ntraces) out[tra * ns + t] = t == 0 ? 0.0f : (float)(y.x * inv_n);
                    if (tra + 1 < ntraces) out[(tra + 1) * ns + t] = t == 0 ? 0.0f : (float)(y.y * inv_n);
                }
            }
        }
    }
}

/* the passes of one direction: radix 16 with strides 1, 16, 256, ... and a
 * last pass of radix N / 16^k when N is not a power of 16 */
template <class Pro, int N, int NS, bool INV, int FIRST, bool LAST_OUT>
__device__ __forceinline__ void pf_passes(double2 *buf, int tj, const double2 *tw,
                                          const double2 *G, const float *in, const float *fold,
                                          int64_t tra, int64_t ntraces, int ns, float dt,
                                          float *out)
{
    constexpr int REM = N / NS;             /* points left to combine */
    constexpr int R = REM >= 16 ? 16 : REM;
    constexpr bool LAST = NS * R == N;
    constexpr int MODE = (LAST && LAST_OUT) ? 3 : FIRST;
    pf_pass<Pro, N, R, NS, INV, MODE>(buf, tj, tw, G, in, fold, tra, ntraces, ns, dt, out);
    if constexpr (!LAST)
        pf_passes<Pro, N, NS * R, INV, 0, LAST_OUT>(buf, tj, tw, G, in, fold, tra, ntraces, ns, dt,
                                                    out);
}

template <class Pro, int LOGN>
__global__ void __launch_bounds__(LOGN == 13 ? 512 : 256)
k_postfft(const float *__restrict__ in, const float *__restrict__ fold, int64_t ntraces, int ns,
          float dt, const double2 *__restrict__ G, const double2 *__restrict__ tw,
          float *__restrict__ out)
{
    constexpr int N = 1 << LOGN, TPF = N / 16;
    constexpr int NT = LOGN == 13 ? 512 : 256;
    constexpr int FPW = NT / TPF;           /* transforms per workgroup */
    constexpr int ROW = N + N / 16;         /* padded LDS row (double2) */
    extern __shared__ double2 pf_lds[];
    const int tid = threadIdx.x;
    const int f = tid / TPF, tj = tid % TPF;
    double2 *buf = pf_lds + f * ROW;
    const int64_t tra = ((int64_t)blockIdx.x * FPW + f) * 2;
    pf_passes<Pro, N, 1, false, 1, false>(buf, tj, tw, G, in, fold, tra, ntraces, ns, dt, out);
    pf_passes<Pro, N, 1, true, 2, true>(buf, tj, tw, G, in, fold, tra, ntraces, ns, dt, out);
}

/* log2 of the transform length for ns samples, 0 if ns is beyond the FFT
 * kernels (the caller then takes the direct product of ktm_postmx.h) */
static inline int postfft_logn(int ns)
{
    int logn = PF_MIN_LOGN;
    while ((1 << logn) < 2 * ns - 1) ++logn;
    return logn <= PF_MAX_LOGN ? logn : 0;
}

/* Host tables of the FFT form for the kernel h[0..ns): G = DFT_N(g) and the
 * twiddles exp(-2 pi i m / N), m < N, each N double2 ({re, im}) into
 * caller arrays.  G by an iterative radix-2 transform in double with
 * twiddles from cos/sin of the exact angle. */
static inline void postfft_host_tables(const double *h, int ns, int logn, double *G, double *tw)
{
    const int n = 1 << logn;
    const double two_pi = 6.283185307179586476925286766559;
    for (int m = 0; m < n; ++m) {
        tw[2 * m] = cos(two_pi * m / n);
        tw[2 * m + 1] = -sin(two_pi * m / n);
    }
    for (int j = 0; j < n; ++j) G[2 * j] = G[2 * j + 1] = 0.0;
    for (int d = 0; d < ns; ++d) G[2 * d] = h[d];
    for (int d = 1; d < ns; ++d) G[2 * (n - d)] = h[ns - d];
    /* bit-reversal permutation, then log2 N butterfly stages */
    for (int i = 1, j = 0; i < n; ++i) {
        int bit = n >> 1;
        for (; j & bit; bit >>= 1) j ^= bit;
        j ^= bit;
        if (i < j) {
            const double tr = G[2 * i], ti = G[2 * i + 1];
            G[2 * i] = G[2 * j]; G[2 * i + 1] = G[2 * j + 1];
            G[2 * j] = tr; G[2 * j + 1] = ti;
        }
    }
    for (int len = 2; len <= n; len <<= 1) {
        const int half = len >> 1, stride = n / len;
        for (int s = 0; s < n; s += len)
            for (int k = 0; k < half; ++k) {
                const double wr = tw[2 * k * stride], wi = tw[2 * k * stride + 1];
                double *a = G + 2 * (s + k), *b = G + 2 * (s + k + half);
                const double br = b[0] * wr - b[1] * wi, bi = b[0] * wi + b[1] * wr;
                b[0] = a[0] - br; b[1] = a[1] - bi;
                a[0] += br; a[1] += bi;
            }
    }
}

/* Launch k_postfft<Pro> over ntraces traces of ns samples; G and tw are the
 * device copies of postfft_host_tables for logn = postfft_logn(ns). */
template <class Pro>
static hipError_t postfft_launch(const float *in, const float *fold, int64_t ntraces, int ns,
                                 float dt, const double2 *G, const double2 *tw, float *out,
                                 hipStream_t stream)
{
    typedef void (*fn_t)(const float *, const float *, int64_t, int, float, const double2 *,
                         const double2 *, float *);
    const int logn = postfft_logn(ns);
    if (!logn) return hipErrorInvalidValue;
    static const fn_t fns[] = {k_postfft<Pro, 8>, k_postfft<Pro, 9>, k_postfft<Pro, 10>,
                               k_postfft<Pro, 11>, k_postfft<Pro, 12>, k_postfft<Pro, 13>};
    const fn_t fn = fns[logn - PF_MIN_LOGN];
    const int n = 1 << logn, nt = logn == 13 ? 512 : 256;
    const int fpw = nt / (n / 16);
    const size_t lds = sizeof(double2) * (size_t)fpw * (size_t)(n + n / 16);
    const int64_t nblk = (ntraces + 2 * fpw - 1) / (2 * fpw);
    if (nblk > 0x7fffffffLL) return hipErrorInvalidValue;
    hipError_t e = hipFuncSetAttribute((const void *)fn, hipFuncAttributeMaxDynamicSharedMemorySize,
                                       (int)lds);
    if (e != hipSuccess) return e;
    hipLaunchKernelGGL(fn, dim3((unsigned)nblk), dim3(nt), lds, stream, in, fold, ntraces, ns, dt,
                       G, tw, out);
    return hipGetLastError();
}

#endif
