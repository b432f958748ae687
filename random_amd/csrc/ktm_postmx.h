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
 * read shared by the wave's NTW tiles) and a 16 x 4 block of H read from h,
 * ~0.5 B of LDS per FMA.  Operand layout of v_mfma_f64_16x16x4f64 for lane
 * l (measured with one-hot operands): A[l % 16][l / 16], B[l / 16][l % 16],
 * D[4 r + l / 16][l % 16] in result register r.
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

template <class Pro, int NTW, bool HG>
__global__ void __launch_bounds__(PM_WAVES * PM_WAVE)
k_postmx(const float *__restrict__ in, const float *__restrict__ fold, int64_t ntraces, int ns,
         float dt, const double *__restrict__ h2, float *__restrict__ out)
{
    extern __shared__ double pm_lds[];
    double *xs = pm_lds;                               /* [PM_UC][PM_LD] */
    double *hl = pm_lds + PM_UC * PM_LD;               /* [2 ns] when !HG */
    const double *hh = HG ? h2 : hl;
    const int tid = threadIdx.x, lane = tid & (PM_WAVE - 1), wv = tid / PM_WAVE;
    const int64_t tr0 = (int64_t)blockIdx.x * 16;
    const int tw0 = (blockIdx.y * PM_WAVES + wv) * NTW * 16;   /* first row of the wave */
    if (!HG)
        for (int m = tid; m < 2 * ns; m += PM_WAVES * PM_WAVE) hl[m] = h2[m];
    pm_d4 acc[NTW];
#pragma unroll
    for (int i = 0; i < NTW; ++i) acc[i] = (pm_d4){0.0, 0.0, 0.0, 0.0};
    const int ar = lane & 15, ak = lane >> 4;          /* A row, k of this lane */
    const int n2 = 2 * ns - 1;
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
            const int u = c0 + k0 + ak;
#pragma unroll
            for (int i = 0; i < NTW; ++i) {
                /* A[ar][ak] = H[t][u] = h2[t - u + ns], clamped into h2 (rows
                 * past ns are not stored; columns past ns meet x = 0) */
                const int m = min(max(tw0 + i * 16 + ar - u + ns, 0), n2);
                acc[i] = __builtin_amdgcn_mfma_f64_16x16x4f64(hh[m], b, acc[i], 0, 0, 0);
            }
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
    /* h in LDS when it fits beside the x chunk, else read through L1/L2 */
    const size_t lds_x = sizeof(double) * PM_UC * PM_LD;
    const bool hg = lds_x + sizeof(double) * 2 * (size_t)ns > 150 * 1024;
    const size_t lds = lds_x + (hg ? 0 : sizeof(double) * 2 * (size_t)ns);
    fn_t fn;
    if (ntw == 4) fn = hg ? k_postmx<Pro, 4, true> : k_postmx<Pro, 4, false>;
    else if (ntw == 8) fn = hg ? k_postmx<Pro, 8, true> : k_postmx<Pro, 8, false>;
    else fn = hg ? k_postmx<Pro, 12, true> : k_postmx<Pro, 12, false>;
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
