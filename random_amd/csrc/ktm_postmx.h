/*
 * ktm_postmx.h -- a real even-kernel circular convolution of many traces in
 * FP64 on the matrix cores, shared by the post-stack filter of Kirchhoff.c
 * (1457-1479) and the post filter of EOM.c (1386-1413):
 *   out[j][t] = sum_u H[t][u] x[j][u],  H[t][u] = h[(t - u) mod ns],
 * t = 0 stored as 0 (both references zero the first output sample).  h2[m] =
 * h[m mod ns] for m < 2 ns; x[j][u] is formed by the prologue Pro from the
 * two float inputs (x[j][0] = 0).
 *
 * k_postmx: a workgroup = 16 traces (the 16 columns of the MFMA tiles) x up
 * to PM_WAVES * NTW * 16 samples (gridDim.y groups cover longer traces);
 * wave w owns NTW 16-row tiles.  x streams through LDS in chunks of PM_UC
 * samples; each v_mfma_f64_16x16x4f64 takes a 4 x 16 block of x (one LDS
 * read shared by the wave's NTW tiles) and a 16 x 4 block of H read from a
 * window of h in LDS (the NTW reads of a k-step at one address plus
 * immediate offsets), ~0.5 B of LDS per FMA.  Operand layout of
 * v_mfma_f64_16x16x4f64 for lane l (measured with one-hot operands):
 * A[l % 16][l / 16], B[l / 16][l % 16], D[4 r + l / 16][l % 16] in result
 * register r.
 */
#ifndef KTM_POSTMX_H
#define KTM_POSTMX_H

#include <hip/hip_runtime.h>
#include <stdint.h>

#define PM_WAVES 8
#define PM_WAVE 64
#define PM_UC 256                    /* samples of x per LDS chunk */
#define PM_LD 17                     /* padded row of the chunk (16 traces) */
typedef double pm_d4 __attribute__((ext_vector_type(4)));

template <class Pro, int NTW>
__global__ void __launch_bounds__(PM_WAVES * PM_WAVE)
k_postmx(const float *__restrict__ in, const float *__restrict__ fold, int64_t ntraces, int ns,
         float dt, const double *__restrict__ h2, float *__restrict__ out)
{
    constexpr int RW = PM_WAVES * NTW * 16;            /* rows of a group */
    extern __shared__ double pm_lds[];
    double *xs = pm_lds;                               /* [PM_UC][PM_LD] */
    double *hp = pm_lds + PM_UC * PM_LD;               /* [RW + ns + 3] window of h */
    const int tid = threadIdx.x, lane = tid & (PM_WAVE - 1), wv = tid / PM_WAVE;
    const int64_t tr0 = (int64_t)blockIdx.x * 16;
    const int g0 = blockIdx.y * RW;                    /* first row of the group */
    const int tw0 = g0 + wv * NTW * 16;                /* first row of the wave */
    /* hp[k] = h[(k + g0 - 3) mod ns]: H[t][u] = h[(t - u) mod ns] =
     * hp[t - g0 - u + ns + 3] for every row t of the group and every u < ns + 3,
     * so the operand reads need no clamp */
    for (int k = tid; k < RW + ns + 3; k += PM_WAVES * PM_WAVE) {
        const int m = (k + g0 - 3) % ns;
        hp[k] = h2[m < 0 ? m + ns : m];
    }
    pm_d4 acc[NTW];
#pragma unroll
    for (int i = 0; i < NTW; ++i) acc[i] = (pm_d4){0.0, 0.0, 0.0, 0.0};
    const int ar = lane & 15, ak = lane >> 4;          /* A row, k of this lane */
    /* A[ar][ak] of tile i at u = u0 + ak: hp[hbase - u0 + 16 i] */
    const int hbase = (tw0 - g0) + ar - ak + ns + 3;
    for (int c0 = 0; c0 < ns; c0 += PM_UC) {
        __syncthreads();                               /* previous chunk consumed */
        /* x[u][j] of the chunk; 0 for u = 0, past the trace or past the image */
        for (int e = tid; e < PM_UC * 16; e += PM_WAVES * PM_WAVE) {
            const int j = e / PM_UC, ul = e - j * PM_UC, u = c0 + ul;
            const int64_t tr = tr0 + j;
            double v = 0.0;
            if (tr < ntraces && u >= 1 && u < ns) v = Pro::x(in + tr * ns, fold + tr * ns, u, dt);
            xs[ul * PM_LD + j] = v;
        }
        __syncthreads();
        const int nu = min(PM_UC, ((ns - c0) + 3) & ~3);
        for (int k0 = 0; k0 < nu; k0 += 4) {
            const double b = xs[(k0 + ak) * PM_LD + ar];  /* B[ak][ar] = x[j=ar][c0+k0+ak] */
            const double *ha = hp + (hbase - c0 - k0);
#pragma unroll
            for (int i = 0; i < NTW; ++i)
                acc[i] = __builtin_amdgcn_mfma_f64_16x16x4f64(ha[16 * i], b, acc[i], 0, 0, 0);
        }
    }
    /* D[4 r + l / 16][l % 16]: rows t, column j = trace */
    const int64_t tr = tr0 + ar;
    if (tr < ntraces) {
#pragma unroll
        for (int i = 0; i < NTW; ++i) {
#pragma unroll
            for (int r = 0; r < 4; ++r) {
                const int t = tw0 + i * 16 + 4 * r + ak;
                if (t < ns) out[tr * ns + t] = t == 0 ? 0.0f : (float)acc[i][r];
            }
        }
    }
}

/* Launch k_postmx<Pro> over ntraces traces of ns samples on stream. */
template <class Pro>
static hipError_t postmx_launch(const float *in, const float *fold, int64_t ntraces, int ns,
                                float dt, const double *h2, float *out, hipStream_t stream)
{
    typedef void (*fn_t)(const float *, const float *, int64_t, int, float, const double *, float *);
    /* rows per wave: the smallest of 4, 8, 12 tiles of 16 covering ns over
     * the 8 waves, else 12 and several row groups */
    const int ntw = ns <= PM_WAVES * 4 * 16 ? 4 : ns <= PM_WAVES * 8 * 16 ? 8 : 12;
    const int rows = PM_WAVES * ntw * 16;
    const int groups = (ns + rows - 1) / rows;
    const size_t lds = sizeof(double) * ((size_t)PM_UC * PM_LD + (size_t)rows + ns + 3);
    if (lds > 160 * 1024) return hipErrorInvalidValue;
    fn_t fn = ntw == 4 ? k_postmx<Pro, 4> : ntw == 8 ? k_postmx<Pro, 8> : k_postmx<Pro, 12>;
    const int64_t nblk = (ntraces + 15) / 16;
    if (nblk > 0x7fffffffLL || groups > 65535) return hipErrorInvalidValue;
    hipError_t e = hipFuncSetAttribute((const void *)fn, hipFuncAttributeMaxDynamicSharedMemorySize,
                                       (int)lds);
    if (e != hipSuccess) return e;
    hipLaunchKernelGGL(fn, dim3((unsigned)nblk, (unsigned)groups), dim3(PM_WAVES * PM_WAVE), lds,
                       stream, in, fold, ntraces, ns, dt, h2, out);
    return hipGetLastError();
}

#endif
