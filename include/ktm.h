/*
 * ktm.h -- C ABI of the MI355X Kirchhoff prestack time-migration engine.
 *
 * The reference program (obriengareth/Random, Kirchhoff.c) exposes one entry
 * point, `void kirchhoff(int job_id)` (Kirchhoff.c:92, 1104), fed through
 * process globals by main()'s setup (Kirchhoff.c:215-373).  This header keeps
 * that signature as a compatibility shim and adds explicit-argument entry
 * points.  Every function returns a ktm_status (no exit(), unlike the
 * reference's error paths at Kirchhoff.c:1122-1136,1415-1420).
 *
 * Pointers documented as "device" are HIP device memory; `stream` is a
 * hipStream_t passed as void* (NULL = the null stream).  A plan is not
 * thread-safe; distinct plans may be used from distinct threads.  A plan's
 * workspaces (the BF16/FP16 decode buffer and the compensated partial
 * images of ktm_migrate_block, the demigration sort keys and max-|km| word
 * of ktm_demigrate) are reused by
 * every call without stream ordering: all calls on one plan must be issued
 * on a single stream (or the caller orders them), or a later call can
 * overwrite a workspace an earlier call's kernels still read.  ktm_prep,
 * ktm_ingest and ktm_geometry use none of these workspaces, nor does
 * ktm_migrate_block on FP32 traces with plain sums: the block loop of
 * ktm_migrate and kirchhoff() runs the input of one block beside the
 * migration of the previous one on a second stream for that reason.
 */
#ifndef KTM_H
#define KTM_H

#include <stddef.h>
#include <stdint.h>

#ifdef __cplusplus
extern "C" {
#endif

typedef enum ktm_status {
    KTM_OK = 0,
    KTM_ERR_ARG = -1,          /* bad argument or inconsistent configuration   */
    KTM_ERR_IO = -2,           /* fopen/fread/fwrite failure (Kirchhoff.c:1122) */
    KTM_ERR_FORMAT = -3,       /* SEG-Y file shorter than its headers claim     */
    KTM_ERR_VELOCITY = -4,     /* velocity <= 0, or a grid index below zero     */
    KTM_ERR_NOMEM = -5,        /* host or device allocation failed              */
    KTM_ERR_DEVICE = -6,       /* HIP runtime error or no device                */
    KTM_ERR_UNSUPPORTED = -7   /* mode/size outside what the kernels implement  */
} ktm_status;

/* Configuration: the compile-time macros of Kirchhoff.c:17-65 as run-time
 * fields (same names, same types: integer macros are int32, decimal ones
 * double), plus the three geometry variants of Kirchhoff_GHTest1.c. */
typedef struct ktm_config {
    int32_t start_iline, end_iline;      /* Startiline, Endiline   Kirchhoff.c:34-35 */
    int32_t start_xline, end_xline;      /* Startxline, Endxline   Kirchhoff.c:36-37 */
    double  inc_iline, inc_xline;        /* INCiline, INCxline     Kirchhoff.c:40-41 */
    double  dh;                          /* dh                     Kirchhoff.c:44    */
    int32_t nh;                          /* nh                     Kirchhoff.c:45    */
    int32_t reserved0;
    double  oh;                          /* oh                     Kirchhoff.c:46    */
    double  ox, oy, angle_e_to_n;        /* ox, oy, angleEtoN      Kirchhoff.c:49-51 */
    double  iline_distance;              /* iline_distance         Kirchhoff.c:54    */
    double  xline_distance;              /* xline_distance         Kirchhoff.c:55    */
    double  aperture;                    /* aperture               Kirchhoff.c:58    */
    double  anti_alias_cutoff;           /* anti_alias_cutoff      Kirchhoff.c:59    */
    int32_t nvel_ilines, nvel_xlines;    /* NVel_ilines/xlines     Kirchhoff.c:22-23 */
    int32_t dvel_ilines, dvel_xlines;    /* dVel_ilines/xlines     Kirchhoff.c:24-25 */
    int32_t ovel_ilines, ovel_xlines;    /* oVel_ilines/xlines     Kirchhoff.c:26-27 */
    int32_t ntblock;                     /* ntblock                Kirchhoff.c:63    */
    int32_t jobs_x, jobs_y;              /* number_mpi_jobsX/Y     Kirchhoff.c:64-65 */
    int32_t image_origin_zero;           /* Kirchhoff_GHTest1.c:1346-1348,1365-1366  */
    int32_t vel_lower_clamp;             /* Kirchhoff_GHTest1.c:1395-1399            */
    int32_t reserved1;
    double  scale_xy;                    /* 0 = Kirchhoff.c; >0 = scaleXY of
                                            Kirchhoff_GHTest1.c:56,1302-1315        */
} ktm_config;

/* The reference's own macro values (Kirchhoff.c:17-65). */
void ktm_config_default(ktm_config *cfg);

/* What main() derives before the job loop: count_size_file (ns, dt, nt,
 * Kirchhoff.c:633-722), create_aliasfilter (trangeMax, Kirchhoff.c:380-441),
 * the slab decomposition (Kirchhoff.c:266-300) and the survey rotation
 * (Kirchhoff.c:312-313).  nxp/nyp = nx*jobs_x, ny*jobs_y: the padded image
 * that the jobs cover together (the last slab is not clipped at Endiline). */
typedef struct ktm_setup {
    int32_t ns, nsm;
    float   dt, dtm;
    int32_t trange_max;
    int32_t Nx, Ny, nx, ny, no_job;
    int32_t nxp, nyp, nh;
    float   sinA, cosA;
    int32_t reserved;
} ktm_setup;

/* Setup without a device: fills *out; anti_alias (may be NULL) gets ns floats.
 * Replaces create_aliasfilter() Kirchhoff.c:380-441 and main() 266-313. */
int ktm_setup_init(const ktm_config *cfg, int ns, float dt, int nsm, float dtm,
                   ktm_setup *out, float *anti_alias);

/* Per-trace geometry and scale as the kernels consume it (32 bytes).
 * sxt..cmpy: rotated source, receiver and CMP (Kirchhoff.c:1296-1309);
 * off: offset class (Kirchhoff.c:1310); scale: per-trace FP32 scale of a
 * BF16/FP16-stored trace (1 for FP32 storage). */
typedef struct ktm_trace_geom {
    float   sxt, syt, gxt, gyt, cmpx, cmpy;
    int32_t off;
    float   scale;
} ktm_trace_geom;

/* How a trace is turned into the filtered trace F(T) that migration sums.
 * REFERENCE: the FP32 double integral of Kirchhoff.c:1312-1328 followed by
 *   the trange=2 second difference of Kirchhoff.c:1391-1392, bit for bit.
 * CLOSED: the same linear operator evaluated as the 3-tap [1,2,1] filter
 *   F(T) = x[T] + 2x[T+1] + x[T+2] on x[t] = tr[t]/(4 t2[t]) (with the
 *   mirror-pad and tail cases of T < 2 and T >= ns-2), which avoids the
 *   catastrophic cancellation of the FP32 prefix sums. */
enum { KTM_PREP_REFERENCE = 0, KTM_PREP_CLOSED = 1 };
/* Storage type of the traces the migration kernel stages. */
enum { KTM_TRACE_F32 = 0, KTM_TRACE_BF16 = 1, KTM_TRACE_F16 = 2 };
/* Image accumulation: plain FP32 adds in trace order (Kirchhoff.c:1394),
 * or FP32 with a Kahan compensation image (the taps of up to 32 staged
 * traces are summed in FP32, then added with one compensated step, so the
 * error stays of order 2^-24 sum |tap|). */
enum { KTM_ACCUM_PLAIN = 0, KTM_ACCUM_KAHAN = 1 };
/* OR'ed into the accumulation argument of ktm_migrate_block with CLOSED prep
 * (FP32, BF16 or FP16 storage): the traveltime index is floor(q) of the
 * single-precision fast form q = sqrt(dIS2/(v^2 dt^2) + t^2/4) +
 * sqrt(dIR2/(v^2 dt^2) + t^2/4) (hardware square roots, one FMA per
 * radicand), without the reference-exact fix-up of the samples whose q lies
 * within 7 ulp of an integer (Kirchhoff.c:1381-1384 rounds differently
 * there).  No counterpart in the reference; for the modes whose contract is
 * an error bound against FP64 rather than the reference's bits. */
enum { KTM_INDEX_FAST = 0x100 };

/* Device-resident setup: velocity model (squared, Kirchhoff.c:617), traveltime
 * tables and the image-column tables of one configuration.  Samples per
 * trace: 4 <= ns <= KTM_MAX_NS (KTM_ERR_UNSUPPORTED otherwise). */
typedef struct ktm_plan ktm_plan;
#define KTM_MAX_NS 6144

/* vel: host array [nvel_ilines][nvel_xlines][nsm] of RMS velocities as read
 * from the velocity SEG-Y (Kirchhoff.c:545).  Fails with KTM_ERR_VELOCITY if
 * an image column maps below the model (a read out of bounds in the
 * reference, Kirchhoff.c:1367-1372) or a velocity used by the image is <= 0.
 * Replaces build_vel_model() Kirchhoff.c:449-626 plus the per-column index
 * arithmetic of Kirchhoff.c:1346-1365,1376-1378. */
int  ktm_plan_create(ktm_plan **plan, const ktm_config *cfg, int ns, float dt,
                     int nsm, float dtm, const float *vel);
void ktm_plan_destroy(ktm_plan *plan);
int  ktm_plan_setup(const ktm_plan *plan, ktm_setup *out);
/* Row length (elements) of one prepared trace in the buffers below. */
int64_t ktm_plan_trace_row(const ktm_plan *plan);
/* Elements of one image array: nxp*nyp*nh*ns. */
int64_t ktm_plan_image_size(const ktm_plan *plan);

/* Trace preparation (device): traces[ntr][ns] IEEE samples, coords[ntr][4]
 * header words sx,sy,gx,gy (bytes 73-88, host order, >= 0).  Writes geom[ntr]
 * and, for F32 storage, prepared[ntr][trace_row] floats holding F(T) for
 * T < ns (zero beyond).  For BF16/F16 storage (CLOSED prep only), prepared
 * [ntr][trace_row] holds the raw trace divided by its max-abs (geom.scale),
 * rounded with sixth-order error feedback that shapes the rounding noise
 * against the [1,2,1] taps; ktm_migrate_block decodes and
 * filters such a block into an FP32 workspace owned by the plan before
 * migrating it.  Replaces Kirchhoff.c:1296-1328. */
int ktm_prep(const ktm_plan *plan, const float *traces, const int32_t *coords,
             int64_t ntr, int prep_mode, int trace_dtype, void *prepared,
             ktm_trace_geom *geom, void *stream);

/* ---- SEG-Y ingest on the device ----
 * raw: ntr trace records exactly as stored in the file after its 3600-byte
 * header (Kirchhoff.c:1281-1283: 240-byte header + ns big-endian sample
 * words), on the device, 4-byte aligned.  format: the SEG-Y sample format
 * code, 1 (IBM hex float, decoded as Kirchhoff.c:1284-1290) or 5 (IEEE). */
enum { KTM_SEGY_IBM = 1, KTM_SEGY_IEEE = 5 };
/* The block's prep in one pass per kernel: byte swap, IBM decode and header
 * words 18-21 (Kirchhoff.c:1297,1302) inside the geometry and prep kernels;
 * results identical to ktm_segy_read + ktm_prep.  coords ([ntr][4], may be
 * NULL) receives the header words.  Replaces the per-trace fread/decode
 * loop of Kirchhoff.c:1276-1328. */
int ktm_ingest(const ktm_plan *plan, const void *raw, int64_t ntr, int format,
               int prep_mode, int trace_dtype, void *prepared, ktm_trace_geom *geom,
               int32_t *coords, void *stream);
/* Decode raw records into FP32 traces [ntr][ns] and header words 18-21
 * [ntr][4] (either may be NULL), bit for bit as ktm_segy_read does on the
 * host (NaN payloads aside). */
int ktm_segy_decode(const void *raw, int64_t ntr, int ns, int format, float *traces,
                    int32_t *coords, void *stream);

/* Encode FP32 traces [ntr][ns] (device) into raw SEG-Y trace records
 * (device, ntr * (240 + 4 ns) bytes, 4-byte aligned): the 240-byte headers
 * copied from headers ([ntr][240] device bytes, or zeros if NULL) and the
 * samples as big-endian IEEE (format 5) or IBM (format 1, truncating, the
 * same encoding as ktm_segy_write). */
int ktm_segy_encode(const float *traces, int64_t ntr, int ns, int format, const void *headers,
                    void *records, void *stream);

/* km = 0, fold = 1, comp = 0 (Kirchhoff.c:1262-1274); comp may be NULL. */
int ktm_image_reset(const ktm_plan *plan, float *km, float *fold, float *comp,
                    void *stream);

/* Migrate one block of prepared traces into the resident image (device):
 * the trace loop of Kirchhoff.c:1292-1406 for every job slab at once.
 * km/fold/comp are [nxp][nyp][nh][ns]; comp is needed for KAHAN only.
 * Accumulation continues from the current contents, in trace order, so
 * successive blocks equal one pass over their concatenation (bit for bit
 * with PLAIN; with KAHAN within the compensated bound).  With KAHAN a block
 * whose image region is small (fewer than 10240 workgroups of tiles x
 * offset classes) is split along its traces into up to 32 parts (its
 * 512-trace chunks dealt round-robin) summed separately and merged in FP64
 * (deterministic: the same call gives the same bits); the partial images live in the plan, allocated by
 * ktm_plan_create for an image that small (at most 1 GiB).  If counter is
 * non-NULL (device uint64) the number of contributions is added to it: ns per
 * (trace, column) pair through the aperture gates of Kirchhoff.c:1338,1354,
 * traces outside the offset classes included (the reference evaluates their
 * loop body and drops the sum; the kernels skip them). */
int ktm_migrate_block(const ktm_plan *plan, const void *prepared, int trace_dtype,
                      int prep_mode, const ktm_trace_geom *geom, int64_t ntr,
                      int accum, int job, float *km, float *fold, float *comp,
                      uint64_t *counter, void *stream);
/* job < 0 migrates every slab; job >= 0 only the columns of that job's slab
 * (Kirchhoff.c:1341-1347), leaving the rest of the image untouched. */

/* ktm_migrate_block without any allocation, for callers that cannot
 * allocate device memory (code under HIP graph capture, serving loops,
 * harnesses that forbid hipMalloc).  No counterpart in the reference.  Same
 * arguments and results as ktm_migrate_block (bit for bit); the BF16/FP16
 * decode workspace, which ktm_migrate_block allocates in the plan on first
 * use, is instead a fixed device array of KTM_STATIC_STAGE_FLOATS floats
 * (64 MiB) that each GPU holds until the process ends, so a BF16/FP16 call
 * takes at most KTM_STATIC_STAGE_FLOATS / ktm_plan_trace_row(plan) traces
 * (KTM_ERR_UNSUPPORTED beyond; FP32 storage needs no workspace).  The
 * compensated partial images were allocated by ktm_plan_create and serve
 * both entry points alike.  Every call
 * on one GPU shares that array: calls that overlap on a GPU (other streams,
 * other plans, other threads) race and must be serialised by the caller. */
#define KTM_STATIC_STAGE_FLOATS (16u << 20)
int ktm_migrate_block_static(const ktm_plan *plan, const void *prepared, int trace_dtype,
                             int prep_mode, const ktm_trace_geom *geom, int64_t ntr,
                             int accum, int job, float *km, float *fold, float *comp,
                             uint64_t *counter, void *stream);

/* Geometry only (device): coords[ntr][4] -> geom[ntr], Kirchhoff.c:1296-1310. */
int ktm_geometry(const ktm_plan *plan, const int32_t *coords, int64_t ntr,
                 ktm_trace_geom *geom, void *stream);

/* Device memory and stream helpers for C callers (thin wrappers over HIP). */
int ktm_device_alloc(void **ptr, size_t bytes);
int ktm_device_free(void *ptr);
int ktm_copy_to_device(void *dst, const void *src, size_t bytes);
int ktm_copy_to_host(void *dst, const void *src, size_t bytes);
int ktm_device_sync(void);

/* Post-stack (device): out = low-pass(-dt*t*dkm/dt*dt/fold) per image trace,
 * Kirchhoff.c:1423-1479, as an exact circular convolution with the
 * real kernel whose DFT is anti_alias (FP64 arithmetic). */
int ktm_poststack(const ktm_plan *plan, const float *km, const float *fold,
                  float *out, void *stream);


/* ---- adjoint pair (device), for least-squares migration and gradients ----
 * With the CLOSED prep, migration is linear in the traces:
 *   km = L^T d     (ktm_prep CLOSED + ktm_migrate_block into a zeroed km;
 *                   fold is a by-product that depends on geometry only)
 *   d  = L km      (ktm_demigrate: the exact transpose, same traveltime
 *                   index, gates and filter, demigration / modelling)
 * and the post-stack chain image = S km (ktm_poststack, S linear for a fixed
 * fold) has the transpose km = S^T image (ktm_poststack_adjoint).  No
 * counterpart in the reference, which only migrates.
 * ktm_demigrate sums each trace in 64-bit fixed point scaled from max |km|
 * (a finite image): the result is deterministic and exact under scaling of
 * km by powers of two; its quantum is max |km| * 2^(lg - 59), 2^lg >=
 * columns x ns.  It sorts the traces by (offset class, CMP cell) into
 * plan-owned buffers that grow on demand (a first call, or a larger ntr,
 * allocates). */
int ktm_demigrate(const ktm_plan *plan, const float *km, const ktm_trace_geom *geom,
                  int64_t ntr, float *traces, void *stream);
/* ktm_demigrate with the traveltime index of KTM_INDEX_FAST (index =
 * KTM_INDEX_FAST; 0 is ktm_demigrate itself): the exact transpose of
 * ktm_migrate_block with CLOSED prep and accumulation | KTM_INDEX_FAST, so
 * that least-squares migration can run on the relaxed index's cheaper pair.
 * No counterpart in the reference. */
int ktm_demigrate_index(const ktm_plan *plan, const float *km, const ktm_trace_geom *geom,
                        int64_t ntr, float *traces, int index, void *stream);
int ktm_poststack_adjoint(const ktm_plan *plan, const float *image, const float *fold,
                          float *km, void *stream);

/* ---- velocity gradient of a gather-flatness objective (device) ----
 * A differentiable variant of the migration loop (Kirchhoff.c:1332-1404)
 * for velocity analysis; no counterpart in the reference, whose
 * nearest-node velocity (1357-1358,1376) and floor(dsr/dt) (1384) have zero
 * derivative almost everywhere.  Lines only (one crossline column):
 *   v(t,X)     = sum_ij V[i][j] B(t/dtn + 1 - i) B(X/dxn + 1 - j), uniform cubic
 *                B-spline on nt_nodes x nx_nodes, dtn = (ns-1)/(nt_nodes-3),
 *                dxn = (nxp-1)/(nx_nodes-3);
 *   q          = (sqrt(t2 + dIS2/v^2) + sqrt(t2 + dIR2/v^2)) / dt (not floored);
 *   I[X][h][t] = sum over traces of class h in aperture of the CLOSED filtered
 *                trace at q, Catmull-Rom cubic interpolation (C1 in q);
 *   J          = sum_{X,t} sum_{h<nh-1} (I[X][h+1][t] - I[X][h][t])^2.
 * dJ/dV comes from a hand-written backward pass that recomputes traveltimes
 * per time chunk of `chunk` (<= 256) samples instead of storing them. */
typedef struct ktm_vgrad ktm_vgrad;
int ktm_vgrad_create(ktm_vgrad **out, const ktm_config *cfg, int ns, float dt, int nt_nodes,
                     int nx_nodes, int chunk);
int ktm_vgrad_destroy(ktm_vgrad *vg);
/* Use a block of CLOSED-prepared traces (ktm_prep with KTM_PREP_CLOSED,
 * F32; row = ktm_plan_trace_row) and their geometry, both on the device.
 * Bins the traces by (offset class, CMP) on the host and copies them, in
 * that order and zero-padded, into a table of (traces in classes) x (ns + 6)
 * floats that the handle holds with the binned geometry (grown on a larger
 * rebind); prepared and geom may be released on return.  Synchronises the
 * stream. */
int ktm_vgrad_bind(ktm_vgrad *vg, const float *prepared, int64_t row,
                   const ktm_trace_geom *geom, int64_t ntr, void *stream);
/* v(t,X) of node velocities V[nt_nodes][nx_nodes] on the image grid
 * [nxp][ns] (device); KTM_ERR_VELOCITY if some v is not positive and finite.
 * Synchronises the stream. */
int ktm_vgrad_velocity(ktm_vgrad *vg, const float *V, float *v, void *stream);
/* J (host, FP64) for node velocities V (device); if dJdV != NULL also the
 * gradient [nt_nodes][nx_nodes] (device, FP64); if cig != NULL the gathers
 * I [nxp][nh][ns] (device).  Synchronises the stream. */
int ktm_vgrad_eval(ktm_vgrad *vg, const float *V, double *J, double *dJdV, float *cig,
                   void *stream);
/* (trace, column) pairs of the bound block that pass the aperture gate (each
 * contributes ns samples per pass); -1 before ktm_vgrad_bind. */
int64_t ktm_vgrad_pairs(const ktm_vgrad *vg);

/* The trace records of image<job>.su (Kirchhoff.c:1437-1598) on the device:
 * for every (xc, yc, ih) of the job's slab in the reference's order, the
 * 240-byte header of ktm_segy_image_trace_header and ns big-endian IEEE
 * samples of image [nxp][nyp][nh][ns] (the ktm_poststack output).  records:
 * device, nx*ny*nh*(240 + 4 ns) bytes, 4-byte aligned.  The file is these
 * bytes written as they are. */
int ktm_image_records(const ktm_plan *plan, const float *image, int job, void *records,
                      void *stream);

/* ---- equivalent-offset migration (EOM.c): common-scatter-point gathers ----
 * The macros of EOM.c:60-99 with their types (oh is an integer macro). */
typedef struct ktm_eom_config {
    double  ox, oy, angle_e_to_n;        /* EOM.c:66-68 */
    double  oxcsp, oycsp, excsp, eycsp;  /* EOM.c:73-76 */
    double  scale_xy;                    /* EOM.c:77 */
    double  dcsp_x, dcsp_y, dh;          /* EOM.c:79-81 */
    double  aperture, low_pass_cutoff;   /* EOM.c:85-86 */
    int32_t nh, oh;                      /* EOM.c:82, 99 */
    int32_t jobs_x, jobs_y;              /* EOM.c:91-92 */
    int32_t ntblock, reserved;           /* EOM.c:89 */
    double  velocity_start, velocity_end;/* EOM.c:61-62 */
} ktm_eom_config;
void ktm_eom_config_default(ktm_eom_config *cfg);
/* main()'s EOM setup (EOM.c:279-320) and the filters of EOM.c:759-845:
 * lowpass and rootrho (rootrhoR == rootrhoI) get ns floats each if non-NULL. */
typedef struct ktm_eom_setup {
    int32_t ns;
    float   dt;
    int32_t Nx, Ny, nx, ny, no_job, nxp, nyp, nh;
    float   sinA, cosA, start_x, start_y;   /* EOM.c:282-286 */
} ktm_eom_setup;
int ktm_eom_setup_init(const ktm_eom_config *cfg, int ns, float dt, ktm_eom_setup *out,
                       float *lowpass, float *rootrho);
typedef struct ktm_eom_plan ktm_eom_plan;
/* vel: ns RMS velocities (squared inside, EOM.c:877; the reference's own
 * model is velocity_start at every sample, EOM.c:871-875). */
int ktm_eom_plan_create(ktm_eom_plan **plan, const ktm_eom_config *cfg, int ns, float dt,
                        const float *vel);
int ktm_eom_plan_destroy(ktm_eom_plan *plan);
int ktm_eom_plan_setup(const ktm_eom_plan *plan, ktm_eom_setup *out);
/* Elements of one gather volume csp/fold: nxp*nyp*nh*ns, [X][Y][offset bin][t]. */
int64_t ktm_eom_image_size(const ktm_eom_plan *plan);
/* csp = 0, fold = 1 (EOM.c:1202-1211), device. */
int ktm_eom_reset(const ktm_eom_plan *plan, float *csp, float *fold, void *stream);
/* Bin a block of traces (device: traces[ntr][ns] IEEE samples, coords[ntr][4]
 * header words 18-21) into the gathers of every job slab (job < 0) or one,
 * continuing from their contents in trace order: the trace loop of
 * EOM.c:1226-1316.  counter (device uint64, may be NULL) accumulates the
 * evaluated samples, sum over passing (trace, CSP) of ns - talpha. */
int ktm_eom_block(const ktm_eom_plan *plan, const float *traces, const int32_t *coords,
                  int64_t ntr, int job, float *csp, float *fold, uint64_t *counter, void *stream);
/* out = Re IFFT(rootrho (1+i) lowpass FFT((csp/fold) dt t)) / ns per gather
 * trace, out[..][0] = 0 (EOM.c:1386-1413), device; ns <= 8192. */
int ktm_eom_post(const ktm_eom_plan *plan, const float *csp, const float *fold, float *out,
                 void *stream);
/* The same on one job's gathers only, [nx][ny][nh][ns] (csp_data and fold of
 * csp(job_id), EOM.c:1190-1211): element count, reset, block (job >= 0) and
 * post filter, device. */
int64_t ktm_eom_job_size(const ktm_eom_plan *plan);
int ktm_eom_job_reset(const ktm_eom_plan *plan, float *csp, float *fold, void *stream);
int ktm_eom_job_block(const ktm_eom_plan *plan, const float *traces, const int32_t *coords,
                      int64_t ntr, int job, float *csp, float *fold, uint64_t *counter,
                      void *stream);
int ktm_eom_job_post(const ktm_eom_plan *plan, const float *csp, const float *fold, float *out,
                     void *stream);
/* The records of image<job>.su (EOM.c:1346-1537) from the job's post-filtered
 * gathers post [nx][ny][nh][ns] (device): per (xc, yc, ih) in order a 240-byte
 * header (tracl = running number, cdp = iline = CSP number, offset =
 * ceil(oh + ih dh), sx..gy = 13..16, cdpx/cdpy = ox/oy + the rotated CSP
 * position truncated, cline = 1, 9999 at byte 236 and, from the second record
 * on, at 204, 220, 224, 232) and ns big-endian IEEE samples (the first 0).
 * Only records with running number < Nx*Ny*nh + 1 are written (EOM.c:1417),
 * a prefix of the job's: *nrec (host) receives their count.  records:
 * device, room for nx*ny*nh*(240 + 4 ns) bytes, 4-byte aligned. */
int ktm_eom_job_records(const ktm_eom_plan *plan, const float *post, int job, void *records,
                        int64_t *nrec, void *stream);

/* Host convenience: whole reference job loop for all jobs (Kirchhoff.c:
 * 347-350 on one rank): prep + migrate in blocks of cfg->ntblock + post-stack,
 * on host arrays.  Replaces main()'s setup plus its kirchhoff(job) loop
 * (Kirchhoff.c:254-350) with explicit arguments and a status code. */
typedef struct ktm_stats {
    uint64_t contributions;   /* (trace, column) pairs in aperture x ns */
    double   migrate_ms;      /* device time of the migration kernels   */
    double   total_ms;        /* host wall time of the call             */
} ktm_stats;

/* The input traces as the reference reads them (Kirchhoff.c:1276-1310):
 * ntr traces of ns FP32 samples at dt seconds and the header words sx, sy,
 * gx, gy of each (bytes 73-88 of the trace header, as integers). */
typedef struct ktm_traces {
    const float   *samples;   /* [ntr][ns] */
    const int32_t *coords;    /* [ntr][4]  */
    int64_t        ntr;
    int32_t        ns;
    float          dt;        /* s */
} ktm_traces;

/* The RMS velocity cube of build_vel_model (Kirchhoff.c:449-626), in m/s
 * (not squared), [inline][crossline][nsm] at dtm seconds per sample. */
typedef struct ktm_velocity {
    const float *v;
    int32_t      nsm;
    float        dtm;
} ktm_velocity;

/* Outputs, [nxp][nyp][nh][ns] host arrays; any may be NULL: the summed
 * image km, its fold count, and the post-stack image (Kirchhoff.c:1423-1479). */
typedef struct ktm_image {
    float *km;
    float *fold;
    float *image;
} ktm_image;

int ktm_migrate(const ktm_config *cfg, const ktm_traces *traces, const ktm_velocity *vel,
                ktm_image *out, ktm_stats *stats);

/* ---- SEG-Y (host), READ_SEGY.c:29-115,169-450 and Kirchhoff.c:633-874 ---- */

/* IBM hex float words in file byte order -> IEEE (Kirchhoff.c:882-895,
 * 1284-1290, including its overflow-to-inf and 0*inf=NaN cases). */
void ktm_ibm_to_float(const uint8_t *bytes, int64_t nwords, float *out);
/* count_size_file (Kirchhoff.c:633-722): ns, dt from trace 1, nt from size. */
int  ktm_segy_count(const char *path, int *ns, float *dt, int *nt);
/* Read traces [first, first+count) of a SEG-Y file: samples IBM-decoded like
 * Kirchhoff.c:1284-1290 (ieee != 0 reads format-5 IEEE instead), coords gets
 * header words 18-21 (sx,sy,gx,gy), extra (may be NULL) gets header words
 * 47-48 (bytes 189-196: iline/xline, Kirchhoff.c:537-538). */
int  ktm_segy_read(const char *path, int64_t first, int64_t count, int ieee,
                   float *samples, int32_t *coords, int32_t *extra);
/* Write a SEG-Y file from IEEE samples and full 240-byte headers. */
int  ktm_segy_write(const char *path, const uint8_t *file_header3600,
                    const uint8_t *trace_headers, const float *samples,
                    int64_t ntr, int ns, int ieee);
/* build_vel_model (Kirchhoff.c:449-626): velocity SEG-Y -> vel[il][xl][nsm]
 * (not squared; zeros are counted in *nzero like the check at 614-615). */
int  ktm_segy_read_velocity(const char *path, const ktm_config *cfg, float *vel,
                            int *nsm, float *dtm, int64_t *nzero);
/* Headers of the image output (Kirchhoff.c:729-814, 1437-1590). fix_count
 * writes at byte 3212 the trace count the reference meant to write.  A trace
 * header other than the job file's first (xc, yc, ih) = (0, 0, 0) carries
 * 9999 at bytes 204, 220, 224 and 232 as well as 236: the reference's `mess`
 * keeps the value set at Kirchhoff.c:1589 for the following records. */
void ktm_segy_file_header(const ktm_config *cfg, int ns, float dt, int fix_count,
                          uint8_t out[3600]);
void ktm_segy_image_trace_header(const ktm_config *cfg, int ns, float dt, int job,
                                 int xc, int yc, int ih, uint8_t out[240]);

/* ---- reference-signature shim (Kirchhoff.c:92, 215-373, 1104) ---- */

/* Does main()'s setup: count_size_file on input_data, create_aliasfilter,
 * build_vel_model on input_vel, mkdir image_dir, header files.  vel_path may
 * be NULL with vel (host, [il][xl][nsm]) given instead. */
int  ktm_globals_init(const ktm_config *cfg, const char *input_data,
                      const char *input_vel, const float *vel, int nsm, float dtm,
                      const char *image_dir);
/* kirchhoff(job_id): migrates the job's slab on the GPU from the input file
 * and writes image_dir/image<job_id>.su byte-compatible with Kirchhoff.c. */
void kirchhoff(int job_id);
/* Status of the last kirchhoff() call (the reference exits instead). */
int  ktm_last_status(void);
/* Counts and times of the last kirchhoff() call: its contributions, the
 * device time of its migration launches and its host wall time, file
 * input, post-stack and image<job>.su written included (no counterpart in
 * the reference, which prints per-block timings instead). */
int  ktm_last_stats(ktm_stats *out);
void ktm_globals_free(void);

/* combine_all_files() of Kirchhoff.c:821-874 (EOM.c:543-590 is the same):
 * writes image_dir/combine_script -- cat of ebdic_header.bin,
 * binary_header.bin and image0.su .. image<no_job-1>.su into output_data,
 * then rm of the parts -- and makes it executable (a+x); it is not run. */
int  ktm_combine_script(const char *image_dir, const char *output_data, int no_job);
/* What that script's cat lines produce, done here: output_data = the two
 * header files and every job's image in job order (the parts stay). */
int  ktm_combine_files(const char *image_dir, const char *output_data, int no_job);

/* csp(job_id), the entry point of EOM.c (EOM.c:1058), with main()'s setup
 * (EOM.c:274-345) done by ktm_eom_globals_init: count_size_file, the CSP
 * grid, the velocity of velocity_model 0 (EOM.c:869-878), the filters, mkdir
 * image_dir, ebdic_header.bin + binary_header.bin.  csp(job) reads the input
 * SEG-Y (IBM) in blocks of cfg->ntblock, decodes, bins, post-filters and
 * encodes on the device and writes image_dir/image<job>.su. */
int  ktm_eom_globals_init(const ktm_eom_config *cfg, const char *input_data,
                          const char *image_dir);
void csp(int job_id);
/* Status of the last csp() call (the reference exits instead). */
int  ktm_eom_last_status(void);
void ktm_eom_globals_free(void);

/* ---- seeded synthetic surveys (device), for tests and benchmarks ---- */

/* traces[ntr][ns] = sum of Ricker wavelets (peak frequency f0) on the
 * double-square-root diffraction curves of nscat point scatterers
 * scat[nscat][4] = (x, y, t0, amplitude) in the grid frame of geom (from
 * ktm_geometry), with v = v0 + vt*t0 + vx*x at each scatterer. */
int ktm_synth_traces(const ktm_plan *plan, const ktm_trace_geom *geom, int64_t ntr,
                     const float *scat, int nscat, float f0, float v0, float vt, float vx,
                     float *traces, void *stream);
/* traces[i] += sigma * N(0,1), the normal deviate a function of (seed, i). */
int ktm_synth_noise(float *traces, int64_t n, float sigma, uint64_t seed, void *stream);
/* Measurement helper (no reference counterpart): an LDS read-bandwidth probe
 * of nblocks workgroups of 256 lanes, each reading iters * 64 bytes per lane
 * from LDS; out (device) receives nblocks * 256 words. */
int ktm_probe_lds(uint32_t *out, int nblocks, int iters, void *stream);
/* Measurement helper (no reference counterpart): a VALU issue-rate probe.
 * nblocks workgroups of 256 lanes each run iters iterations of a fixed
 * inline-assembly VALU sequence over 8 independent register chains: mode 0
 * v_fma_f32, 1 v_add_f32, 2 v_sqrt_f32, 3 v_cvt_u32_f32/v_cvt_f32_u32,
 * 4 one dependent v_fma_f32 chain, 5 the migration tap (2 fma, 2 sqrt, add,
 * min, cvt, lshl_add, add), 6 one sqrt per four fma, 7 one sqrt per fma,
 * 8 v_min_f32, 9 v_rndne_f32, 10 v_lshl_add_u32, 11 v_cmp_le_f32, 12
 * v_cndmask_b32, 13 v_add_u32, 14 v_fma_f64, 15 v_readlane_b32, 16 v_rsq_f64,
 * 17 v_lshlrev_b64, 18 v_mov_b32.
 * ktm_probe_valu_ops(mode) is the VALU instruction count per lane and
 * iteration (-1 for a bad mode).  out (device) receives 3 words per
 * workgroup: shader-clock cycles (clock64) and 100 MHz ticks (wall_clock64)
 * of its loop, and a checksum. */
int ktm_probe_valu(int mode, uint32_t *out, int nblocks, int iters, void *stream);
int ktm_probe_valu_ops(int mode);

/* The float t such that, for every float x >= 0, x < t exactly when
 * (double)(float)sqrt((double)x) < A -- the form of the aperture and slab
 * gates of Kirchhoff.c:1334-1338,1352-1354 -- so the kernels gate on squared
 * distances with the reference's decisions. */
float ktm_gate_threshold(double A);

/* ---- self-tests of the numerical assumptions of the kernels (device) ---- */

/* out_hw[i] = the hardware square root v_sqrt_f32(x[i]) used by the fast
 * traveltime path (its error bound is an input of that path's proof). */
int ktm_selftest_sqrt(const float *x, int64_t n, float *out_hw, void *stream);
/* in4[n][4] = (dIS2, dIR2, t, v2), t a sample index held as a float (t2
 * is formed as in the traveltime table, Kirchhoff.c:1246-1247): q_fast (fast
 * path), q_ref = RN(dsr/dt) with the arithmetic of Kirchhoff.c:1381-1384,
 * and the sample index the migration kernel selects (t_kernel) vs the
 * reference's (t_ref). */
int ktm_selftest_traveltime(const float *in4, int64_t n, float dt, int ns, float *q_fast,
                            float *q_ref, int32_t *t_kernel, int32_t *t_ref, void *stream);

#ifdef __cplusplus
}
#endif
#endif /* KTM_H */
