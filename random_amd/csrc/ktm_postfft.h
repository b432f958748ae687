/*
 * ktm_postfft.h -- the real circular convolution of ktm_postmx.h,
 *   out[j][t] = sum_u h[(t - u) mod ns] x[j][u]   (t = 0 stored as 0),
 * by FFT in FP64: with N the smallest of 256, 512, ..., 8192 and 3072 that is
 * >= 2 ns - 1, the kernel
 * wrapped into N points, g[d] = h[d] (0 <= d < ns) and g[N - d] = h[ns - d]
 * (0 < d < ns), zero between, makes the length-N circular convolution of the
 * zero-padded trace equal the length-ns one on t < ns:
 *   out = IDFT_N(G . DFT_N(x)) / N,   G = DFT_N(g),
 * the O(ns log ns) form of Kirchhoff.c:1465-1478 (FFTW forward, filter,
 * backward, / ns: here the 1/ns is in h) and EOM.c:1386-1413.  g is real,
 * so two traces go through one complex transform, x_a + i x_b, and come
 * out as its real and imaginary parts.
 *
 * k_postfft: radix-16 Stockham passes (the last one radix 2, 4, 8 or 12
 * when N is not a power of 16), in place in LDS: every thread loads the
 * points of its butterflies (at most 16) into registers, the workgroup
 * synchronises, and the results go to their Stockham positions.  N / 16
 * threads per transform; a workgroup of 256 threads holds 4096 / N
 * transforms (N <= 4096) or one transform of 512 threads (N = 8192); N =
 * 3072 = 16 x 16 x 12 (traces of 1025 to 1536 samples) takes 256 threads,
 * 192 of them busy in the radix-16 passes, and 52 KB of LDS, so three
 * workgroups share a CU where two of 4096 points would.  The first forward pass
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

template <bool INV>
__device__ __forceinline__ void pf_dft3(double2 &a0, double2 &a1, double2 &a2)
{
    /* X0 = a0 + a1 + a2, X1,2 = a0 - (a1 + a2)/2 -+ i sqrt(3)/2 (a1 - a2) (forward) */
    constexpr double S3 = 0.86602540378443864676;
    const double2 t1 = pf_add(a1, a2);
    const double2 t2 = make_double2(a0.x - 0.5 * t1.x, a0.y - 0.5 * t1.y);
    const double2 d = pf_sub(a1, a2);
    const double2 t3 = INV ? make_double2(-S3 * d.y, S3 * d.x) : make_double2(S3 * d.y, -S3 * d.x);
    a0 = pf_add(a0, t1);
    a1 = pf_add(t2, t3);
    a2 = pf_sub(t2, t3);
}

/* exp(-2 pi i m / 12), m = 1, 2, 3, 4, 6 */
__device__ __forceinline__ double2 pf_w12(int m)
{
    constexpr double S3 = 0.86602540378443864676;
    switch (m) {
    case 1: return make_double2(S3, -0.5);
    case 2: return make_double2(0.5, -S3);
    case 3: return make_double2(0.0, -1.0);
    case 4: return make_double2(-0.5, -S3);
    default: return make_double2(-1.0, 0.0);   /* 6 */
    }
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

/* In-place DFT of R = 16, 12, 8, 4 or 2 points v[o], v[o + 1], ...; the
 * result X[k] ends in v[o + pf_slot<R>(k)]. */
template <int R> __device__ __forceinline__ constexpr int pf_slot(int k)
{
    return R == 16 ? 4 * (k & 3) + (k >> 2)
           : R == 12 ? 4 * (k % 3) + k / 3
           : R == 8 ? 4 * (k & 1) + (k >> 1) : k;
}

template <int R, bool INV> __device__ __forceinline__ void pf_dft(double2 *v)
{
    if (R == 2) {
        pf_dft2<INV>(v[0], v[1]);
    } else if (R == 4) {
        pf_dft4<INV>(v[0], v[1], v[2], v[3]);
    } else if (R == 12) {
        /* n = 4 n1 + n2: 3-point DFTs over n1, twiddle W12^(n2 k1), 4-point over n2 */
#pragma unroll
        for (int n2 = 0; n2 < 4; ++n2) pf_dft3<INV>(v[n2], v[4 + n2], v[8 + n2]);
#pragma unroll
        for (int k1 = 1; k1 < 3; ++k1)
#pragma unroll
            for (int n2 = 1; n2 < 4; ++n2)
                v[4 * k1 + n2] = pf_mul(v[4 * k1 + n2], pf_conj<INV>(pf_w12(n2 * k1)));
#pragma unroll
        for (int k1 = 0; k1 < 3; ++k1) pf_dft4<INV>(v[4 * k1], v[4 * k1 + 1], v[4 * k1 + 2], v[4 * k1 + 3]);
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
 * transform of N points whose TPF threads include this one (index tj): the
 * N / R butterflies go round-robin over the threads.  MODE 0: LDS -> LDS;
 * 1: the first forward pass, input from Pro; 2: the first inverse pass,
 * input times G; 3: the last inverse pass, output to out. */
template <class Pro, int N, int R, int NS, int TPF, bool INV, int MODE>
__device__ __forceinline__ void pf_pass(double2 *buf, int tj, const double2 *__restrict__ tw,
                                        const double2 *__restrict__ G, const float *in,
                                        const float *fold, int64_t tra, int64_t ntraces, int ns,
                                        float dt, float *out)
{
    constexpr int NBF = N / R, NB = (NBF + TPF - 1) / TPF, STEP = N / R;
    static_assert(NB * R <= 16, "at most 16 points per thread");
    constexpr bool GUARD = NBF % TPF != 0;  /* some threads idle in this pass */
    double2 v[16];
#pragma unroll
    for (int b = 0; b < NB; ++b) {
        const int j = tj + b * TPF;
        if (GUARD && j >= NBF) continue;
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
            if (GUARD && j >= NBF) continue;
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
            if (GUARD && j >= NBF) continue;
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
 * last pass of radix N / 16^k (2, 4, 8 or 12) when N is not a power of 16 */
template <class Pro, int N, int NS, int TPF, bool INV, int FIRST, bool LAST_OUT>
__device__ __forceinline__ void pf_passes(double2 *buf, int tj, const double2 *tw,
                                          const double2 *G, const float *in, const float *fold,
                                          int64_t tra, int64_t ntraces, int ns, float dt,
                                          float *out)
{
    constexpr int REM = N / NS;             /* points left to combine */
    constexpr int R = REM >= 16 ? 16 : REM;
    constexpr bool LAST = NS * R == N;
    constexpr int MODE = (LAST && LAST_OUT) ? 3 : FIRST;
    pf_pass<Pro, N, R, NS, TPF, INV, MODE>(buf, tj, tw, G, in, fold, tra, ntraces, ns, dt, out);
    if constexpr (!LAST)
        pf_passes<Pro, N, NS * R, TPF, INV, 0, LAST_OUT>(buf, tj, tw, G, in, fold, tra, ntraces,
                                                         ns, dt, out);
}

/* threads per workgroup and per transform for an N-point transform */
template <int N> struct PfShape {
    static constexpr int NT = N == 8192 ? 512 : 256;
    static constexpr int TPF = N == 3072 ? 256 : N / 16;
    static constexpr int FPW = NT / TPF;   /* transforms per workgroup */
    static constexpr int ROW = N + N / 16; /* padded LDS row (double2) */
    static constexpr int WAVES_PER_EU = N == 3072 ? 3 : 2;
};

template <class Pro, int N>
__global__ void __launch_bounds__(PfShape<N>::NT)
    __attribute__((amdgpu_waves_per_eu(PfShape<N>::WAVES_PER_EU)))
k_postfft(const float *__restrict__ in, const float *__restrict__ fold, int64_t ntraces, int ns,
          float dt, const double2 *__restrict__ G, const double2 *__restrict__ tw,
          float *__restrict__ out)
{
    typedef PfShape<N> S;
    extern __shared__ double2 pf_lds[];
    const int tid = threadIdx.x;
    const int f = tid / S::TPF, tj = tid % S::TPF;
    double2 *buf = pf_lds + f * S::ROW;
    const int64_t tra = ((int64_t)blockIdx.x * S::FPW + f) * 2;
    pf_passes<Pro, N, 1, S::TPF, false, 1, false>(buf, tj, tw, G, in, fold, tra, ntraces, ns, dt,
                                                  out);
    pf_passes<Pro, N, 1, S::TPF, true, 2, true>(buf, tj, tw, G, in, fold, tra, ntraces, ns, dt,
                                                out);
}

/* The transform length for ns samples: the smallest of 256, 512, 1024, 2048,
 * 3072, 4096, 8192 that is >= 2 ns - 1, or 0 beyond (the caller then takes
 * the direct product of ktm_postmx.h). */
static inline int postfft_size(int ns)
{
    static const int sizes[] = {256, 512, 1024, 2048, 3072, 4096, 8192};
    for (int n : sizes)
        if (n >= 2 * ns - 1) return n;
    return 0;
}

/* Host tables of the FFT form for the kernel h[0..ns): G = DFT_N(g) and the
 * twiddles exp(-2 pi i m / N), m < N, each N double2 ({re, im}) into
 * caller arrays.  G in double: by an iterative radix-2 transform for a
 * power of two, directly (the sum over g's 2 ns - 1 nonzero points, angles
 * from the twiddle table) otherwise; twiddles from cos/sin of the exact
 * angle. */
static inline void postfft_host_tables(const double *h, int ns, int n, double *G, double *tw)
{
    const double two_pi = 6.283185307179586476925286766559;
    for (int m = 0; m < n; ++m) {
        tw[2 * m] = cos(two_pi * m / n);
        tw[2 * m + 1] = -sin(two_pi * m / n);
    }
    if (n & (n - 1)) {
        for (int k = 0; k < n; ++k) {
            double re = 0.0, im = 0.0;
            /* g[d] = h[d]: angle index d k mod n, stepped by k */
            for (int d = 0, m = 0; d < ns; ++d, m = m + k >= n ? m + k - n : m + k) {
                re += h[d] * tw[2 * m];
                im += h[d] * tw[2 * m + 1];
            }
            /* g[n - d] = h[ns - d]: angle index (n - d) k mod n = -d k mod n */
            for (int d = 1, m = k ? n - k : 0; d < ns; ++d, m = m >= k ? m - k : m - k + n) {
                re += h[ns - d] * tw[2 * m];
                im += h[ns - d] * tw[2 * m + 1];
            }
            G[2 * k] = re;
            G[2 * k + 1] = im;
        }
        return;
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
 * device copies of postfft_host_tables for n = postfft_size(ns). */
template <class Pro>
static hipError_t postfft_launch(const float *in, const float *fold, int64_t ntraces, int ns,
                                 float dt, const double2 *G, const double2 *tw, float *out,
                                 hipStream_t stream)
{
    typedef void (*fn_t)(const float *, const float *, int64_t, int, float, const double2 *,
                         const double2 *, float *);
    const int n = postfft_size(ns);
    fn_t fn = nullptr;
    int nt = 256, fpw = 1;
    switch (n) {
    case 256: fn = k_postfft<Pro, 256>; fpw = PfShape<256>::FPW; break;
    case 512: fn = k_postfft<Pro, 512>; fpw = PfShape<512>::FPW; break;
    case 1024: fn = k_postfft<Pro, 1024>; fpw = PfShape<1024>::FPW; break;
    case 2048: fn = k_postfft<Pro, 2048>; fpw = PfShape<2048>::FPW; break;
    case 3072: fn = k_postfft<Pro, 3072>; fpw = PfShape<3072>::FPW; break;
    case 4096: fn = k_postfft<Pro, 4096>; fpw = PfShape<4096>::FPW; break;
    case 8192: fn = k_postfft<Pro, 8192>; fpw = PfShape<8192>::FPW; nt = 512; break;
    default: return hipErrorInvalidValue;
    }
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
