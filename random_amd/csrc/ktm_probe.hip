/*
 * ktm_probe.hip -- VALU issue-rate probe of gfx950 (a measurement helper with
 * no counterpart in the reference).
 *
 * The migration kernel k_migrate issues about one VALU instruction per
 * contribution and waits on nothing else (LDS and HBM are far from their
 * peaks), so its roofline is the rate at which the compute units issue VALU
 * instructions of its mix.  This probe measures that rate instead of
 * assuming it: every lane runs a fixed sequence of VALU instructions written
 * as inline assembly (so the count per iteration is exact and the compiler
 * adds no VALU work), over 8 independent register chains interleaved so that
 * no instruction waits on its predecessor, at full occupancy.  Each
 * workgroup also records the shader clock (clock64, s_memtime) and the
 * 100 MHz constant clock (wall_clock64) around its loop, so the host can
 * express the rate per CU-cycle at the clock the chip actually ran.
 *
 * Modes (VALU instructions per lane per loop iteration in brackets):
 *   0 v_fma_f32, independent                                    [32]
 *   1 v_add_f32, independent                                    [32]
 *   2 v_sqrt_f32, independent (transcendental)                  [32]
 *   3 v_cvt_u32_f32 / v_cvt_f32_u32 pairs                       [32]
 *   4 v_fma_f32, one dependent chain                            [32]
 *   5 the migration tap: 2 v_fma_f32, 2 v_sqrt_f32, v_add_f32,
 *     v_min_f32, v_cvt_u32_f32, v_lshl_add_u32, v_add_f32       [72]
 *   6 one v_sqrt_f32 per four v_fma_f32                          [40]
 *   7 one v_sqrt_f32 per v_fma_f32                               [32]
 *   8 v_min_f32                                                 [32]
 *   9 v_rndne_f32                                               [32]
 *  10 v_lshl_add_u32                                            [32]
 *  11 v_cmp_le_f32 into an SGPR pair                            [32]
 *  12 v_cndmask_b32                                             [32]
 *  13 v_add_u32                                                 [32]
 *  14 v_fma_f64                                                 [32]
 *  15 v_readlane_b32                                            [32]
 *  16 v_rsq_f64 (transcendental)                                [32]
 *  17 v_lshlrev_b64                                             [32]
 *  18 v_mov_b32                                                 [32]
 * For modes 11, 15 and 18 the compiler inserts hazard waits between the
 * inline-assembly instructions: their rates are lower bounds.
 */
#include <hip/hip_runtime.h>

#include <stdint.h>

#include "ktm_internal.h"

#define PV_NT 256

#define V_FMA(x, a, b) asm volatile("v_fma_f32 %0, %0, %1, %2" : "+v"(x) : "v"(a), "v"(b))
#define V_FMA3(d, a, b, c) asm volatile("v_fma_f32 %0, %1, %2, %3" : "=v"(d) : "v"(a), "v"(b), "v"(c))
#define V_ADD(x, a) asm volatile("v_add_f32 %0, %0, %1" : "+v"(x) : "v"(a))
#define V_SQRT(x) asm volatile("v_sqrt_f32 %0, %0" : "+v"(x))
#define V_CVT_U(u, x) asm volatile("v_cvt_u32_f32 %0, %1" : "=v"(u) : "v"(x))
#define V_CVT_F(x, u) asm volatile("v_cvt_f32_u32 %0, %1" : "=v"(x) : "v"(u))

/* 8 chains, each op applied to all 8 before the next op */
#define ALL8(OP) do { OP(0); OP(1); OP(2); OP(3); OP(4); OP(5); OP(6); OP(7); } while (0)

template <int MODE>
__global__ void __launch_bounds__(PV_NT) k_probe_valu(uint32_t *__restrict__ out, int iters,
                                                      float seed)
{
    const int tid = threadIdx.x;
    float x[8], y[8];
    for (int k = 0; k < 8; ++k) {
        x[k] = seed + 0.001f * (float)(tid + k);
        y[k] = x[k] + 1.0f;
    }
    const float d1 = 3.5f, d2 = 7.25f, rs = 0.25f, nsf = 1500.0f;
    const uint32_t base = 64u;
    uint32_t w[8];
    double z[8];
    for (int k = 0; k < 8; ++k) {
        w[k] = (uint32_t)tid + 3u * (uint32_t)k;
        z[k] = (double)x[k];
    }
    const double za = 0.9999, zb = 1e-4;
    uint64_t m[8] = {0, 0, 0, 0, 0, 0, 0, 0};
    uint32_t sr[8] = {0, 0, 0, 0, 0, 0, 0, 0};
    const uint32_t lanesel = 5u;
    const float ca = 0.9999f, cb = 1e-4f;
    const uint64_t c0 = clock64();
    const uint64_t r0 = wall_clock64();
    for (int it = 0; it < iters; ++it) {
        if (MODE == 0) {
#define OP(k) V_FMA(x[k], ca, cb)
            for (int u = 0; u < 4; ++u) ALL8(OP);
#undef OP
        } else if (MODE == 1) {
#define OP(k) V_ADD(x[k], cb)
            for (int u = 0; u < 4; ++u) ALL8(OP);
#undef OP
        } else if (MODE == 2) {
#define OP(k) V_SQRT(x[k])
            for (int u = 0; u < 4; ++u) ALL8(OP);
#undef OP
        } else if (MODE == 3) {
#define OPU(k) V_CVT_U(w[k], x[k])
#define OPF(k) V_CVT_F(x[k], w[k])
            for (int u = 0; u < 2; ++u) {
                ALL8(OPU);
                ALL8(OPF);
            }
#undef OPU
#undef OPF
        } else if (MODE == 4) {
            for (int u = 0; u < 32; ++u) V_FMA(x[0], ca, cb);
        } else if (MODE == 5) {
            /* per chain: a1 = fma(d1, rs, tq), a2 = fma(d2, rs, tq), q =
             * sqrt(a1) + sqrt(a2), qc = min(q, ns), T = cvt(qc), addr =
             * T * 4 + base, acc += bits(addr) -- the dependency pattern of
             * one tap of the migration loop, without the LDS read */
            float a1[8], a2[8];
            uint32_t t[8];
            {
#define OP1(k) V_FMA3(a1[k], d1, rs, x[k])
#define OP2(k) V_FMA3(a2[k], d2, rs, x[k])
#define OP3(k) V_SQRT(a1[k])
#define OP4(k) V_SQRT(a2[k])
#define OP5(k) V_ADD(a1[k], a2[k])
#define OP6(k) asm volatile("v_min_f32 %0, %0, %1" : "+v"(a1[k]) : "v"(nsf))
#define OP7(k) V_CVT_U(t[k], a1[k])
#define OP8(k) asm volatile("v_lshl_add_u32 %0, %0, 2, %1" : "+v"(t[k]) : "v"(base))
#define OP9(k) asm volatile("v_add_f32 %0, %0, %1" : "+v"(x[k]) : "v"(t[k]))
                ALL8(OP1); ALL8(OP2); ALL8(OP3); ALL8(OP4); ALL8(OP5);
                ALL8(OP6); ALL8(OP7); ALL8(OP8); ALL8(OP9);
#undef OP1
#undef OP2
#undef OP3
#undef OP4
#undef OP5
#undef OP6
#undef OP7
#undef OP8
#undef OP9
            }
        } else if (MODE == 6) {
#define OPS(k) V_SQRT(y[k])
#define OPA(k) V_FMA(x[k], ca, cb)
            ALL8(OPS); ALL8(OPA); ALL8(OPA); ALL8(OPA); ALL8(OPA);
#undef OPS
#undef OPA
        } else if (MODE == 7) {
#define OPS(k) V_SQRT(y[k])
#define OPA(k) V_FMA(x[k], ca, cb)
            for (int u = 0; u < 2; ++u) {
                ALL8(OPS); ALL8(OPA);
            }
#undef OPS
#undef OPA
        } else if (MODE == 8) {
#define OP(k) asm volatile("v_min_f32 %0, %0, %1" : "+v"(x[k]) : "v"(nsf))
            for (int u = 0; u < 4; ++u) ALL8(OP);
#undef OP
        } else if (MODE == 9) {
#define OP(k) asm volatile("v_rndne_f32 %0, %0" : "+v"(x[k]))
            for (int u = 0; u < 4; ++u) ALL8(OP);
#undef OP
        } else if (MODE == 10) {
#define OP(k) asm volatile("v_lshl_add_u32 %0, %0, 2, %1" : "+v"(w[k]) : "v"(base))
            for (int u = 0; u < 4; ++u) ALL8(OP);
#undef OP
        } else if (MODE == 11) {
            /* independent compares into 8 SGPR pairs (not read) */
#define OP(k) asm volatile("v_cmp_le_f32 %0, %1, %2" : "=s"(m[k]) : "v"(x[k]), "v"(y[k]))
            for (int u = 0; u < 4; ++u) ALL8(OP);
#undef OP
        } else if (MODE == 12) {
#define OP(k) asm volatile("v_cndmask_b32 %0, %0, %1, %2" : "+v"(x[k]) : "v"(y[k]), "s"(m[k]))
            for (int u = 0; u < 4; ++u) ALL8(OP);
#undef OP
        } else if (MODE == 13) {
#define OP(k) asm volatile("v_add_u32 %0, %0, %1" : "+v"(w[k]) : "v"(base))
            for (int u = 0; u < 4; ++u) ALL8(OP);
#undef OP
        } else if (MODE == 14) {
#define OP(k) asm volatile("v_fma_f64 %0, %0, %1, %2" : "+v"(z[k]) : "v"(za), "v"(zb))
            for (int u = 0; u < 4; ++u) ALL8(OP);
#undef OP
        } else if (MODE == 15) {
#define OP(k) asm volatile("v_readlane_b32 %0, %1, %2" : "=s"(sr[k]) : "v"(x[k]), "s"(lanesel))
            for (int u = 0; u < 4; ++u) ALL8(OP);
#undef OP
        } else if (MODE == 16) {
#define OP(k) asm volatile("v_rsq_f64 %0, %0" : "+v"(z[k]))
            for (int u = 0; u < 4; ++u) ALL8(OP);
#undef OP
        } else if (MODE == 17) {
#define OP(k) asm volatile("v_lshlrev_b64 %0, 1, %0" : "+v"(z[k]))
            for (int u = 0; u < 4; ++u) ALL8(OP);
#undef OP
        } else if (MODE == 18) {
#define OP(k) asm volatile("v_mov_b32 %0, %1" : "=v"(w[k]) : "v"(x[k]))
            for (int u = 0; u < 4; ++u) ALL8(OP);
#undef OP
        }
    }
    const uint64_t c1 = clock64();
    const uint64_t r1 = wall_clock64();
    float s = 0.0f;
    for (int k = 0; k < 8; ++k) s += x[k] + y[k] + (float)w[k] + (float)z[k];
    for (int k = 0; k < 8; ++k) s += (float)(uint32_t)m[k] + (float)sr[k];
    if (tid == 0) {
        out[3 * blockIdx.x + 0] = (uint32_t)(c1 - c0);
        out[3 * blockIdx.x + 1] = (uint32_t)(r1 - r0);
    }
    /* keep every chain live: one word per workgroup that depends on all */
    __shared__ float red[PV_NT];
    red[tid] = s;
    __syncthreads();
    if (tid == 0) {
        float t = 0.0f;
        for (int k = 0; k < PV_NT; ++k) t += red[k];
        out[3 * blockIdx.x + 2] = __float_as_uint(t);
    }
}

/* VALU instructions per lane per loop iteration of each mode (the inline
 * assembly above) */
extern "C" int ktm_probe_valu_ops(int mode)
{
    static const int ops[19] = {32, 32, 32, 32, 32, 72, 40, 32, 32, 32,
                                32, 32, 32, 32, 32, 32, 32, 32, 32};
    return mode >= 0 && mode < 19 ? ops[mode] : -1;
}

/* out (device): 3 words per workgroup -- shader-clock cycles and 100 MHz
 * ticks of its loop, and a checksum */
extern "C" int ktm_probe_valu(int mode, uint32_t *out, int nblocks, int iters, void *stream)
{
    if (!out || nblocks < 1 || iters < 1 || mode < 0 || mode > 18) return KTM_ERR_ARG;
    typedef void (*fn_t)(uint32_t *, int, float);
    static const fn_t fns[19] = {k_probe_valu<0>, k_probe_valu<1>, k_probe_valu<2>, k_probe_valu<3>,
                                 k_probe_valu<4>, k_probe_valu<5>, k_probe_valu<6>, k_probe_valu<7>,
                                 k_probe_valu<8>, k_probe_valu<9>, k_probe_valu<10>, k_probe_valu<11>,
                                 k_probe_valu<12>, k_probe_valu<13>, k_probe_valu<14>, k_probe_valu<15>,
                                 k_probe_valu<16>, k_probe_valu<17>, k_probe_valu<18>};
    hipLaunchKernelGGL(fns[mode], dim3((unsigned)nblocks), dim3(PV_NT), 0, (hipStream_t)stream,
                       out, iters, 1.5f);
    return hipGetLastError() == hipSuccess ? KTM_OK : KTM_ERR_DEVICE;
}
