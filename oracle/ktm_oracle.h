/*
 * ktm_oracle.h -- CPU oracle for the Kirchhoff prestack time-migration hot path.
 *
 * TEST INFRASTRUCTURE ONLY.  This is a from-scratch C restatement of the
 * reference program's `kirchhoff()` path (Kirchhoff.c, obriengareth/Random),
 * written to be the checker of the HIP product path.  Only tests/, the smoke
 * check and bench.py's cpu_baseline leg may load it; the product library
 * never links or calls it.
 *
 * Arithmetic follows the C source of the reference exactly as the C abstract
 * machine defines it (FLT_EVAL_METHOD == 0, IEEE round-to-nearest, no FMA
 * contraction): every float/double promotion and every rounding point of
 * Kirchhoff.c is reproduced.  The reference itself was built with
 * -ffast-math (Kirchhoff.c:3); what that flag did to a given build is
 * compiler dependent, so the strict-IEEE reading is the canonical semantics.
 *
 * Parity pins (tests/test_oracle_*.py):
 *   orc_ibm_to_float      exact known-answer words + exhaustive exponent sweep
 *   orc_alias_filter      closed-form values, trangeMax table, odd-ns zero bins
 *   orc_trace_prep_*      closed form F(T)=x[T]+2x[T+1]+x[T+2] and the mirror /
 *                         tail special cases in exact rational arithmetic
 *   orc_migrate_job       single-spike impulse response on the DSR surface,
 *                         point-scatterer focusing, linearity, zero-in/zero-out,
 *                         closed-form fold count, slab-union == whole image,
 *                         brute-force pure-Python loop on a tiny survey
 *   orc_poststack         direct DFT in exact rational/np.fft agreement,
 *                         linear-phase impulse response
 *   orc_segy_*            byte layout of READ_SEGY.c:29-115 (field order), the
 *                         write sequence of Kirchhoff.c:1499-1590 word by word
 *   orc_eom_*             brute-force pure-Python loop, equivalent-offset closed
 *                         forms, linearity, fold count, np.fft for the filter
 *   orc_vgrad             FP64 central differences of J at every node, spline
 *                         partition of unity, Catmull-Rom reproduction of
 *                         quadratics, zero-offset identity I = F
 */
#ifndef KTM_ORACLE_H
#define KTM_ORACLE_H

#include <stdint.h>

#ifdef __cplusplus
extern "C" {
#endif

/* Mirrors the compile-time configuration macros of Kirchhoff.c:17-65 plus the
 * three geometry variants of Kirchhoff_GHTest1.c.  Types follow the macros:
 * integer macros are int, the ones written with a decimal point are double. */
typedef struct orc_params {
    int32_t start_iline, end_iline;      /* Kirchhoff.c:34-35 */
    int32_t start_xline, end_xline;      /* Kirchhoff.c:36-37 */
    double  inc_iline, inc_xline;        /* Kirchhoff.c:40-41 */
    double  dh;                          /* Kirchhoff.c:44 */
    int32_t nh;                          /* Kirchhoff.c:45 */
    int32_t pad0;
    double  oh;                          /* Kirchhoff.c:46 */
    double  ox, oy, angle_e_to_n;        /* Kirchhoff.c:49-51 */
    double  iline_distance;              /* Kirchhoff.c:54 */
    double  xline_distance;              /* Kirchhoff.c:55 */
    double  aperture;                    /* Kirchhoff.c:58 */
    double  anti_alias_cutoff;           /* Kirchhoff.c:59 */
    int32_t nvel_ilines, nvel_xlines;    /* Kirchhoff.c:22-23 */
    int32_t dvel_ilines, dvel_xlines;    /* Kirchhoff.c:24-25 */
    int32_t ovel_ilines, ovel_xlines;    /* Kirchhoff.c:26-27 */
    int32_t ntblock;                     /* Kirchhoff.c:63 */
    int32_t jobs_x, jobs_y;              /* Kirchhoff.c:64-65 */
    int32_t image_origin_zero;           /* Kirchhoff_GHTest1.c:1346-1348,1365-1366 */
    int32_t vel_lower_clamp;             /* Kirchhoff_GHTest1.c:1395-1399 */
    int32_t pad1;
    double  scale_xy;                    /* 0: Kirchhoff.c; >0: Kirchhoff_GHTest1.c:56 */
} orc_params;

/* Slab decomposition of Kirchhoff.c:266-300. */
typedef struct orc_grid {
    int32_t Nx, Ny, nx, ny, no_job;
    float   sinA, cosA;                  /* Kirchhoff.c:312-313 */
} orc_grid;

void  orc_grid_init(const orc_params *p, orc_grid *g);

/* Kirchhoff.c:882-895 + 1284-1290: one big-endian IBM word as stored in the
 * file (four bytes in file order) -> IEEE float. */
float orc_ibm_to_float(const uint8_t be[4]);
void  orc_ibm_block(const uint8_t *bytes, int64_t nwords, float *out);

/* Kirchhoff.c:380-441.  Returns trangeMax, fills anti_alias[ns]. */
int   orc_alias_filter(const orc_params *p, int ns, float dt, float *anti_alias);

/* Kirchhoff.c:1246-1247 + 1312-1328: double anti-alias integral of one trace
 * into dataP[ns + 2*tm] (FP32 as written; FP64 variant widens every float). */
void  orc_trace_prep_fp32(const float *tr, int ns, float dt, int tm, float *dataP);
void  orc_trace_prep_fp64(const float *tr, int ns, float dt, int tm, double *dataP);

/* Kirchhoff.c:1296-1310: rotated source/receiver, CMP and offset class.
 * geo[0..5] = sxt, syt, gxt, gyt, cmpX, cmpY; returns off. */
int   orc_trace_geometry(const orc_params *p, const orc_grid *g,
                         const int32_t coords[4], float geo[6]);

/* Migration modes. */
enum { ORC_FP32 = 0,        /* Kirchhoff.c exactly: km/fold are float         */
       ORC_FP64_DATA = 1,   /* FP32 traveltime index T, FP64 data path + sums */
       ORC_FP64_FULL = 2 }; /* every float of kirchhoff() widened to double   */

/* One job of Kirchhoff.c:1104-1404 (the migration part, before the
 * post-stack filter).  traces[ntr][ns] are IEEE samples, coords[ntr][4] the
 * header words sx,sy,gx,gy (bytes 73-88) as host integers, vel[nvel_il]
 * [nvel_xl][nsm] the RMS velocity as read (Kirchhoff.c:545), squared inside
 * like Kirchhoff.c:617.  km/fold are the job slab [nx][ny][nh][ns], float for
 * ORC_FP32 and double otherwise; fold starts at 1 (Kirchhoff.c:1270).
 * Returns the number of contributions: sum over traces of the columns that
 * pass the aperture gate of Kirchhoff.c:1354 (and the slab prefilter 1338),
 * times ns.  Returns -1 on a velocity index outside the model (a read the
 * reference performs out of bounds). */
int64_t orc_migrate_job(const orc_params *p, int ns, float dt, int nsm, float dtm,
                        int job, const float *traces, const int32_t *coords,
                        int64_t ntr, const float *vel, int mode,
                        void *km, void *fold);

/* All jobs (Kirchhoff.c:347-350 on one rank), written into the padded whole
 * image [nx*jobs_x][ny*jobs_y][nh][ns]; jobs run in parallel on nthreads
 * OpenMP threads (the reference's one-rank-per-slab decomposition). */
int64_t orc_migrate_all(const orc_params *p, int ns, float dt, int nsm, float dtm,
                        const float *traces, const int32_t *coords, int64_t ntr,
                        const float *vel, int mode, void *km, void *fold,
                        int nthreads);

/* Kirchhoff.c:1423-1479: derivative, fold scaling, FFT low-pass, for
 * ntraces output traces of ns samples.  FP32 semantics of the reference
 * (FFT in double, spectrum rounded through float a,c,d as at 1470-1472).
 * The DFT is the textbook definition evaluated directly in double. */
void  orc_poststack_fp32(int ns, float dt, const float *anti_alias,
                         const float *km, const float *fold, int64_t ntraces,
                         float *out);
void  orc_poststack_fp64(int ns, float dt, const float *anti_alias,
                         const double *km, const double *fold, int64_t ntraces,
                         double *out);

/* Kirchhoff.c:1437-1593: 240-byte header of output trace (xc,yc,ih) of a job,
 * tracl is the running trace number (Kirchhoff.c:1438) and mess the value the
 * function-local `mess` holds when the record is written: 0 (Kirchhoff.c:1224)
 * for the first record of a kirchhoff() call, 9999 (left by Kirchhoff.c:1589)
 * for every later one.  It lands in bytes 205, 221, 225 and 233. */
void  orc_segy_trace_header(const orc_params *p, const orc_grid *g, int ns, float dt,
                            int job, int xc, int yc, int ih, int tracl, int32_t mess,
                            uint8_t out[240]);
/* The headers of image<job>.su in write order (Kirchhoff.c:1437-1598), with
 * the running tracl and the carried `mess`: out[nx*ny*nh][240]. */
void  orc_segy_job_headers(const orc_params *p, const orc_grid *g, int ns, float dt,
                           int job, uint8_t *out);
/* Kirchhoff.c:729-814: 3200-byte text + 400-byte binary header; fix_count
 * writes the trace count the reference meant to write at byte 12. */
void  orc_segy_file_header(const orc_params *p, const orc_grid *g, int ns, float dt,
                           int fix_count, uint8_t out[3600]);

/* Kirchhoff.c:633-722 / READ_SEGY.c:169-209: ns, dt, nt of a file. */
int   orc_count_size(const uint8_t *head3720, int64_t length, int *ns, float *dt,
                     int *nt);

/* Differentiable offset-class migration (an extension with no counterpart in
 * Kirchhoff.c, whose nearest-neighbour velocity 1357-1358,1376 and floor(T)
 * 1384 have zero derivative almost everywhere).  FP64 throughout.
 *   v(t,X) = sum_ij V[i][j] B((t/dtn) + 1 - i) B((X/dxn) + 1 - j), B the centred
 *            cubic B-spline, dtn = (ns-1)/(nt_nodes-3), dxn = (nxp-1)/(nx_nodes-3)
 *   q      = (sqrt(t2 + a/v^2) + sqrt(t2 + b/v^2)) / dt, t2 = (t dt / 2)^2,
 *            a, b the squared source/receiver distances (Kirchhoff.c:1348-1351)
 *   I[X][h][t] = sum over traces of class h passing the aperture gate of
 *            Kirchhoff.c:1354 of the Catmull-Rom interpolant of F at q
 *            (F zero outside [0, ns), q clamped to ns + 2)
 *   J      = sum_{X,t} sum_{h < nh-1} (I[X][h+1][t] - I[X][h][t])^2
 * F: [ntr][ns] filtered traces; geo[ntr][6] and off[ntr] from
 * orc_trace_geometry; colx[nxp] the column coordinates, coly the line's.
 * Fills cig[nxp][nh][ns] and dJdV[nt_nodes][nx_nodes] when non-NULL (the
 * analytic gradient); returns J. */
double orc_vgrad(int ns, double dt, double aperture, int nh, int nxp, const float *colx,
                 float coly, const double *F, const float *geo, const int32_t *off,
                 int64_t ntr, int nt_nodes, int nx_nodes, const double *V, double *cig,
                 double *dJdV);
/* v(t, X) of the spline above, [nxp][ns]. */
void  orc_vgrad_velocity(int ns, int nxp, int nt_nodes, int nx_nodes, const double *V,
                         double *v);

/* ---- Equivalent-offset migration (EOM.c, csp(), the common-scatter-point
 * gather operator).  Macros of EOM.c:60-99 with their types. ---- */
typedef struct orc_eom_params {
    double  ox, oy, angle_e_to_n;        /* EOM.c:66-68 */
    double  oxcsp, oycsp, excsp, eycsp;  /* EOM.c:73-76 */
    double  scale_xy;                    /* EOM.c:77 */
    double  dcsp_x, dcsp_y, dh;          /* EOM.c:79-81 */
    double  aperture, low_pass_cutoff;   /* EOM.c:85-86 */
    int32_t nh, oh;                      /* EOM.c:82, 99 (oh is an integer macro) */
    int32_t jobs_x, jobs_y;              /* EOM.c:91-92 */
} orc_eom_params;

typedef struct orc_eom_grid {
    int32_t Nx, Ny, nx, ny, no_job;
    float   sinA, cosA, start_x, start_y;  /* EOM.c:282-286 */
} orc_eom_grid;

void  orc_eom_grid_init(const orc_eom_params *p, orc_eom_grid *g);
/* EOM.c:759-845: low-pass and sqrt-rho spectra (rootrhoR == rootrhoI). */
void  orc_eom_filters(int ns, float dt, double cutoff, float *lowpass, float *rootrho);
/* EOM.c:1157-1316 for one job: csp/fold [nx][ny][nh][ns] (csp = 0, fold = 1
 * first), traces[ntr][ns] IEEE samples, coords[ntr][4] header words 18-21,
 * vel[ns] the RMS velocity (squared inside, EOM.c:877).  Returns the number
 * of evaluated samples (sum over passing (trace, CSP) of ns - talpha), -1 on
 * a bad job. */
int64_t orc_eom_job(const orc_eom_params *p, int ns, float dt, int job, const float *traces,
                    const int32_t *coords, int64_t ntr, const float *vel, float *csp,
                    float *fold);
/* EOM.c:1386-1413 for ntraces gathers traces: out = Re IFFT(rootrho (1+i)
 * lowpass FFT((csp/fold) dt t)) / ns, out[0] = 0. */
void  orc_eom_post(int ns, float dt, const float *lowpass, const float *rootrho,
                   const float *csp, const float *fold, int64_t ntraces, float *out);
/* EOM.c:1414-1500: 240-byte header of gather trace (xc,yc,ih) of a job; it is
 * the running trace number, mess the carried value as in
 * orc_segy_trace_header (EOM.c:1182 sets 0, EOM.c:1492 leaves 9999). */
void  orc_eom_trace_header(const orc_eom_params *p, int ns, float dt, int job, int xc, int yc,
                           int ih, int it, int32_t mess, uint8_t out[240]);
/* image<job>.su of csp(job_id) (EOM.c:1364-1505): for (xc, yc, ih) in order,
 * header and ns big-endian floats of post[nx][ny][nh][ns] (the post filter
 * output), only while it < Nx*Ny*nh + 1 (EOM.c:1417).  out holds up to
 * nx*ny*nh*(240 + 4 ns) bytes; returns the number of records written. */
int64_t orc_eom_job_records(const orc_eom_params *p, int ns, float dt, int job,
                            const float *post, uint8_t *out);
/* EOM.c:450-537: ebdic_header.bin + binary_header.bin as 3600 bytes (the
 * uninitialised short at binary byte 12 written as 0). */
void  orc_eom_file_header(int ns, float dt, uint8_t out[3600]);

#ifdef __cplusplus
}
#endif
#endif
