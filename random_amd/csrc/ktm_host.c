/*
 * ktm_host.c -- C host layer: SEG-Y input/output, the whole-survey driver
 * ktm_migrate(), and the reference-signature shim kirchhoff(job_id).
 *
 * The reference reads and writes SEG-Y itself inside kirchhoff()
 * (Kirchhoff.c:1121-1136, 1259-1290, 1414-1593) and gets its setup from
 * main() through globals (Kirchhoff.c:103-115, 215-373).  Here the file
 * handling stays on the host in C, in streamed blocks of ntblock traces, and
 * the migration itself goes to the device through the C ABI of ktm.h.
 */
#define _FILE_OFFSET_BITS 64
#define _POSIX_C_SOURCE 200809L
#include "ktm_internal.h"

#include <errno.h>
#include <math.h>
#include <stdio.h>
#include <stdlib.h>
#include <string.h>
#include <sys/stat.h>
#include <sys/types.h>
#include <time.h>

#define SEGY_FILE_HDR 3600
#define SEGY_TRACE_HDR 240

/* ------------------------------------------------------------------------ */
/* IBM hex float -> IEEE, Kirchhoff.c:882-895 + 1284-1290                   */
/* ------------------------------------------------------------------------ */
static float ibm_word(uint32_t w)
{
    /* value = sign * (fraction / 2^24) * (float)16^(exponent-64), evaluated
     * as the reference does: the power of 16 is rounded to float on its own
     * (overflowing to inf above 2^127, flushing to 0 below 2^-149) before the
     * single float product, so a zero fraction with a huge exponent is NaN */
    const int32_t iw = (int32_t)w;
    const int sign = ((iw >> 31) & 1) ? -1 : 1;
    const float frac = (float)(iw & 0x00ffffff) / 16777216.0f;
    const int e4 = 4 * (((iw >> 24) & 0x7f) - 64);
    const float p16 = (float)ldexp(1.0, e4);
    return (float)sign * frac * p16;
}

void ktm_ibm_to_float(const uint8_t *bytes, int64_t nwords, float *out)
{
    for (int64_t i = 0; i < nwords; ++i) out[i] = ibm_word(ktm_get_be32(bytes + 4 * i));
}

static float ieee_word(uint32_t w)
{
    float f;
    memcpy(&f, &w, 4);
    return f;
}

/* ------------------------------------------------------------------------ */
/* SEG-Y reading, READ_SEGY.c:169-209 / Kirchhoff.c:633-722, 1259-1297       */
/* ------------------------------------------------------------------------ */
static int file_size(FILE *f, int64_t *len)
{
    if (fseeko(f, 0, SEEK_END) != 0) return KTM_ERR_IO;
    const off_t n = ftello(f);
    if (n < 0) return KTM_ERR_IO;
    *len = (int64_t)n;
    return KTM_OK;
}

int ktm_segy_count(const char *path, int *ns, float *dt, int *nt)
{
    if (!path || !ns || !dt || !nt) return KTM_ERR_ARG;
    FILE *f = fopen(path, "rb");
    if (!f) return KTM_ERR_IO;
    int64_t len = 0;
    uint8_t b[4];
    int st = file_size(f, &len);
    if (st == KTM_OK && len < SEGY_FILE_HDR + SEGY_TRACE_HDR) st = KTM_ERR_FORMAT;
    /* ns, dt at bytes 115-118 of the first trace header (Kirchhoff.c:652-658) */
    if (st == KTM_OK && (fseeko(f, SEGY_FILE_HDR + 114, SEEK_SET) != 0 || fread(b, 1, 4, f) != 4))
        st = KTM_ERR_IO;
    fclose(f);
    if (st != KTM_OK) return st;
    const uint16_t ns1 = ktm_get_be16(b), dt1 = ktm_get_be16(b + 2);
    if (ns1 == 0) return KTM_ERR_FORMAT;
    *ns = (int)ns1;
    *dt = (float)(dt1 / 1000000.0);
    *nt = (int)((len - SEGY_FILE_HDR) / (SEGY_TRACE_HDR + 4 * (int64_t)ns1)); /* :659 */
    return KTM_OK;
}

int ktm_segy_read(const char *path, int64_t first, int64_t count, int ieee, float *samples,
                  int32_t *coords, int32_t *extra)
{
    int ns, nt;
    float dt;
    int st = ktm_segy_count(path, &ns, &dt, &nt);
    if (st != KTM_OK) return st;
    if (first < 0 || count < 0 || first + count > nt) return KTM_ERR_ARG;
    FILE *f = fopen(path, "rb");
    if (!f) return KTM_ERR_IO;
    const int64_t rec = SEGY_TRACE_HDR + 4 * (int64_t)ns;
    /* trace n starts at 3600 + (240 + 4 ns) (n-1), READ_SEGY.c:227,441 */
    if (fseeko(f, SEGY_FILE_HDR + first * rec, SEEK_SET) != 0) { fclose(f); return KTM_ERR_IO; }
    const int64_t chunk = 4096;
    uint8_t *buf = (uint8_t *)malloc((size_t)(rec * chunk));
    if (!buf) { fclose(f); return KTM_ERR_NOMEM; }
    for (int64_t i0 = 0; i0 < count && st == KTM_OK; i0 += chunk) {
        const int64_t n = count - i0 < chunk ? count - i0 : chunk;
        if (fread(buf, (size_t)rec, (size_t)n, f) != (size_t)n) { st = KTM_ERR_IO; break; }
        for (int64_t k = 0; k < n; ++k) {
            const uint8_t *h = buf + k * rec;
            const int64_t i = i0 + k;
            if (coords) /* header words 18-21: sx, sy, gx, gy (Kirchhoff.c:1297,1302) */
                for (int c = 0; c < 4; ++c) coords[4 * i + c] = (int32_t)ktm_get_be32(h + 72 + 4 * c);
            if (extra) { /* words 47-48: iline, xline (Kirchhoff.c:537-538) */
                extra[2 * i] = (int32_t)ktm_get_be32(h + 188);
                extra[2 * i + 1] = (int32_t)ktm_get_be32(h + 192);
            }
            if (samples) {
                float *o = samples + i * ns;
                const uint8_t *d = h + SEGY_TRACE_HDR;
                if (ieee)
                    for (int t = 0; t < ns; ++t) o[t] = ieee_word(ktm_get_be32(d + 4 * t));
                else
                    ktm_ibm_to_float(d, ns, o);
            }
        }
    }
    free(buf);
    fclose(f);
    return st;
}

int ktm_segy_write(const char *path, const uint8_t *file_header3600, const uint8_t *trace_headers,
                   const float *samples, int64_t ntr, int ns, int ieee)
{
    if (!path || !file_header3600 || (ntr > 0 && (!trace_headers || !samples)) || ns <= 0)
        return KTM_ERR_ARG;
    FILE *f = fopen(path, "wb");
    if (!f) return KTM_ERR_IO;
    int st = fwrite(file_header3600, 1, SEGY_FILE_HDR, f) == SEGY_FILE_HDR ? KTM_OK : KTM_ERR_IO;
    uint8_t *d = (uint8_t *)malloc((size_t)ns * 4);
    if (!d) st = KTM_ERR_NOMEM;
    for (int64_t i = 0; i < ntr && st == KTM_OK; ++i) {
        const float *s = samples + i * ns;
        for (int t = 0; t < ns; ++t) {
            uint32_t u;
            if (ieee) {
                memcpy(&u, s + t, 4);
            } else { /* IEEE -> IBM (truncating), for writing IBM test files */
                const float v = s[t];
                u = 0;
                if (v != 0.0f && isfinite(v)) {
                    int e2;
                    double m = frexp(fabs((double)v), &e2); /* v = m 2^e2, m in [0.5,1) */
                    int e16 = (e2 + 3) >> 2;                /* ceil(e2/4) */
                    double frac = m * ldexp(1.0, e2 - 4 * e16);
                    uint32_t mant = (uint32_t)(frac * 16777216.0);
                    if (mant > 0xffffff) { mant >>= 4; ++e16; }
                    u = ((v < 0.0f) ? 0x80000000u : 0u) | ((uint32_t)(e16 + 64) << 24) | mant;
                }
            }
            ktm_put_be32(d + 4 * t, u);
        }
        if (fwrite(trace_headers + i * SEGY_TRACE_HDR, 1, SEGY_TRACE_HDR, f) != SEGY_TRACE_HDR ||
            fwrite(d, 4, (size_t)ns, f) != (size_t)ns)
            st = KTM_ERR_IO;
    }
    free(d);
    if (fclose(f) != 0 && st == KTM_OK) st = KTM_ERR_IO;
    return st;
}

int ktm_segy_read_velocity(const char *path, const ktm_config *cfg, float *vel, int *nsm,
                           float *dtm, int64_t *nzero)
{
    /* build_vel_model(), Kirchhoff.c:476-546: grid index by integer division
     * of iline/xline (words 47/48), out-of-grid traces ignored */
    int ns1, nt1;
    float dt1;
    int st = ktm_segy_count(path, &ns1, &dt1, &nt1);
    if (st != KTM_OK) return st;
    if (nsm) *nsm = ns1;
    if (dtm) *dtm = dt1;
    if (!vel) return KTM_OK;
    const int nil = cfg->nvel_ilines, nxl = cfg->nvel_xlines;
    memset(vel, 0, sizeof(float) * (size_t)nil * nxl * ns1);
    const int64_t block = 5000; /* ntblock_vel, Kirchhoff.c:503 */
    float *s = (float *)malloc(sizeof(float) * (size_t)(block * ns1));
    int32_t *ex = (int32_t *)malloc(sizeof(int32_t) * (size_t)(2 * block));
    if (!s || !ex) { free(s); free(ex); return KTM_ERR_NOMEM; }
    for (int64_t i0 = 0; i0 < nt1 && st == KTM_OK; i0 += block) {
        const int64_t n = nt1 - i0 < block ? nt1 - i0 : block;
        st = ktm_segy_read(path, i0, n, 0, s, NULL, ex);
        for (int64_t k = 0; k < n && st == KTM_OK; ++k) {
            const int i = (ex[2 * k] - cfg->ovel_ilines) / cfg->dvel_ilines;
            const int j = (ex[2 * k + 1] - cfg->ovel_xlines) / cfg->dvel_xlines;
            if (i >= 0 && i < nil && j >= 0 && j < nxl)
                memcpy(vel + ((size_t)i * nxl + j) * ns1, s + k * ns1, sizeof(float) * (size_t)ns1);
        }
    }
    free(s);
    free(ex);
    if (st == KTM_OK && nzero) { /* the zero report of Kirchhoff.c:614-615 */
        int64_t z = 0;
        for (size_t k = 0; k < (size_t)nil * nxl * ns1; ++k) z += vel[k] == 0.0f;
        *nzero = z;
    }
    return st;
}

/* ------------------------------------------------------------------------ */
/* SEG-Y headers of the image, Kirchhoff.c:729-814 and 1437-1590            */
/* ------------------------------------------------------------------------ */
void ktm_segy_file_header(const ktm_config *cfg, int ns, float dt, int fix_count, uint8_t out[3600])
{
    ktm_setup s;
    memset(out, 0, SEGY_FILE_HDR);
    if (ktm_setup_init(cfg, ns, dt, 1, dt, &s, NULL) != KTM_OK && s.Nx == 0) return;
    /* the second sprintf of Kirchhoff.c:757-758 overwrites the first */
    static const char text[] = "written by G. O'Brien\n";
    memcpy(out, text, sizeof text - 1);
    uint8_t *bin = out + 3200;
    if (fix_count) ktm_put_be16(bin + 12, (uint16_t)(s.Nx * s.Ny * cfg->nh));
    ktm_put_be16(bin + 16, (uint16_t)(dt * 1000000)); /* htons(dt*1000000), :786 */
    ktm_put_be16(bin + 20, (uint16_t)ns);
    ktm_put_be16(bin + 22, (uint16_t)ns);
    ktm_put_be16(bin + 24, 5);                       /* IEEE format, :797-799 */
    ktm_put_be16(bin + 26, 1);
    ktm_put_be16(bin + 28, 1);
    ktm_put_be16(bin + 30, 1);
}

void ktm_segy_image_trace_header(const ktm_config *cfg, int ns, float dt, int job, int xc, int yc,
                                 int ih, uint8_t out[240])
{
    ktm_setup s;
    memset(out, 0, SEGY_TRACE_HDR);
    ktm_setup_init(cfg, ns, dt, 1, dt, &s, NULL);
    const int X = xc + (job / cfg->jobs_y) * s.nx;
    const int Y = yc + (job % cfg->jobs_y) * s.ny;
    const double il = cfg->start_iline + X * cfg->inc_iline; /* Kirchhoff.c:1482 */
    const double xl = cfg->start_xline + Y * cfg->inc_xline; /* Kirchhoff.c:1483 */
    const int xm = (int)(il * cfg->iline_distance);          /* Kirchhoff.c:1485 (int xc1) */
    const int ym = (int)(xl * cfg->xline_distance);          /* Kirchhoff.c:1486 (int yc1) */
    /* running trace number (Kirchhoff.c:1438,1594) */
    const int tracl = s.nx * s.ny * cfg->nh * job + 1 + (xc * s.ny + yc) * cfg->nh + ih;
    ktm_put_be32(out + 0, (uint32_t)tracl);
    ktm_put_be32(out + 20, (uint32_t)(X + s.Nx * Y + 1));                  /* cdp, :1444 */
    ktm_put_be32(out + 36, (uint32_t)(int)ceil(cfg->oh + ih * cfg->dh));   /* offset, :1484 */
    for (int c = 0; c < 4; ++c) ktm_put_be32(out + 72 + 4 * c, 1u);        /* sx..gy = 1 */
    ktm_put_be16(out + 114, (uint16_t)ns);
    ktm_put_be16(out + 116, (uint16_t)(dt * 1000000));
    /* cdpx/cdpy, Kirchhoff.c:1492-1493: (int)ox + float terms, truncated */
    const float cx = (int)cfg->ox + s.cosA * xm - s.sinA * ym;
    const float cy = (int)cfg->oy + s.sinA * xm + s.cosA * ym;
    ktm_put_be32(out + 180, (uint32_t)(int)cx);
    ktm_put_be32(out + 184, (uint32_t)(int)cy);
    ktm_put_be32(out + 188, (uint32_t)(int)il);
    ktm_put_be32(out + 192, (uint32_t)(int)xl);
    /* the local `mess` (zero at Kirchhoff.c:1224) is written at bytes 204,
     * 220, 224 and 232 of every record and holds htonl(9999) from
     * Kirchhoff.c:1589 on: 0 in the job file's first record, 9999 after */
    static const int mess_at[4] = {204, 220, 224, 232};
    if (xc != 0 || yc != 0 || ih != 0)
        for (int k = 0; k < 4; ++k) ktm_put_be32(out + mess_at[k], 9999u);
    ktm_put_be32(out + 236, 9999u); /* Kirchhoff.c:1589-1590 */
}

/* ------------------------------------------------------------------------ */
/* whole-survey driver                                                       */
/* ------------------------------------------------------------------------ */
static double now_ms(void)
{
    struct timespec ts;
    clock_gettime(CLOCK_MONOTONIC, &ts);
    return ts.tv_sec * 1e3 + ts.tv_nsec * 1e-6;
}

typedef struct dev_bufs {
    float *km, *fold, *image;
    uint64_t *counter;
} dev_bufs;

static void free_bufs(dev_bufs *b)
{
    ktm_device_free(b->km); ktm_device_free(b->fold); ktm_device_free(b->image);
    ktm_device_free(b->counter);
    memset(b, 0, sizeof *b);
}

/* ------------------------------------------------------------------------ */
/* host -> device staging                                                    */
/* ------------------------------------------------------------------------ */
/* Two pinned chunks: the host fills one (fread from the file, or memcpy from
 * the caller's arrays) while the DMA out of the other is in flight; each
 * chunk's event marks its copy done before the chunk is refilled.  For raw
 * SEG-Y records a chunk holds whole records, so the coordinate words of
 * each record are checked as it arrives. */
#define KTM_STAGE_BYTES ((size_t)64 << 20)

typedef struct stage_ring {
    uint8_t *buf[2];
    void *done[2];
    int used[2];
    size_t cap;
    int next;
} stage_ring;

static void ring_free(stage_ring *r)
{
    for (int k = 0; k < 2; ++k) {
        if (r->used[k]) (void)ktm_rt_event_sync(r->done[k]);
        ktm_rt_event_destroy(r->done[k]);
        ktm_rt_pinned_free(r->buf[k]);
    }
    memset(r, 0, sizeof *r);
}

static int ring_init(stage_ring *r, size_t cap)
{
    memset(r, 0, sizeof *r);
    r->cap = cap;
    int st = KTM_OK;
    for (int k = 0; k < 2 && st == KTM_OK; ++k) {
        st = ktm_rt_pinned_alloc((void **)&r->buf[k], cap);
        if (st == KTM_OK) st = ktm_rt_event_create(&r->done[k]);
    }
    if (st != KTM_OK) ring_free(r);
    return st;
}

/* bytes into dst (device) on stream s, from src or, when f != NULL, read
 * from f at its position; rec > 0: the bytes are rec-byte SEG-Y trace
 * records (a chunk holds whole records), and with check their source/
 * receiver words 18-21 (Kirchhoff.c:1297,1302) must not be negative
 * (htonl() - ox on such a word is undefined in the reference) */
static int ring_put(stage_ring *r, void *dst, const void *src, FILE *f, size_t bytes, size_t rec,
                    int check, void *s)
{
    const size_t step = rec ? r->cap / rec * rec : r->cap;
    for (size_t off = 0; off < bytes; off += step) {
        const size_t len = bytes - off < step ? bytes - off : step;
        const int k = r->next;
        r->next ^= 1;
        if (r->used[k]) {
            const int st = ktm_rt_event_sync(r->done[k]);
            if (st != KTM_OK) return st;
        }
        uint8_t *b = r->buf[k];
        if (f) {
            if (fread(b, 1, len, f) != len) return KTM_ERR_IO;
        } else {
            memcpy(b, (const uint8_t *)src + off, len);
        }
        for (size_t o = 0; check && rec && o < len; o += rec)
            for (int c = 0; c < 4; ++c)
                if (b[o + 72 + 4 * c] & 0x80u) return KTM_ERR_ARG;
        int st = ktm_rt_copy_h2d_async((uint8_t *)dst + off, b, len, s);
        if (st == KTM_OK) st = ktm_rt_event_record(r->done[k], s);
        if (st != KTM_OK) return st;
        r->used[k] = 1;
    }
    return KTM_OK;
}

/* bytes of src (device) written to f: each chunk is copied into one pinned
 * buffer while the host writes the other (stream s) */
static int ring_write(stage_ring *r, FILE *f, const void *src, size_t bytes, void *s)
{
    const size_t nch = (bytes + r->cap - 1) / r->cap;
    int st = KTM_OK;
    for (size_t c = 0; c <= nch && st == KTM_OK; ++c) {
        /* issue chunk c, then write chunk c - 1 (whose buffer's copy is the
         * older event) */
        if (c < nch) {
            const int k = (int)(c & 1);
            const size_t off = c * r->cap, len = bytes - off < r->cap ? bytes - off : r->cap;
            if (r->used[k]) st = ktm_rt_event_sync(r->done[k]);
            if (st == KTM_OK) st = ktm_rt_copy_d2h_async(r->buf[k], (const uint8_t *)src + off, len, s);
            if (st == KTM_OK) st = ktm_rt_event_record(r->done[k], s);
            if (st == KTM_OK) r->used[k] = 1;
        }
        if (c > 0 && st == KTM_OK) {
            const int k = (int)((c - 1) & 1);
            const size_t off = (c - 1) * r->cap, len = bytes - off < r->cap ? bytes - off : r->cap;
            st = ktm_rt_event_sync(r->done[k]);
            if (st == KTM_OK && fwrite(r->buf[k], 1, len, f) != len) st = KTM_ERR_IO;
        }
    }
    return st;
}

/* the file `name` holding bytes of src (device), through a staging ring */
static int write_device_file(const char *name, const void *src, size_t bytes)
{
    FILE *f = fopen(name, "wb");
    if (!f) return KTM_ERR_IO;
    stage_ring ring;
    void *s = NULL;
    int st = ring_init(&ring, KTM_STAGE_BYTES);
    if (st == KTM_OK) st = ktm_rt_stream_create(&s);
    if (st == KTM_OK && bytes) st = ring_write(&ring, f, src, bytes, s);
    if (s && ktm_rt_stream_sync(s) != KTM_OK && st == KTM_OK) st = KTM_ERR_DEVICE;
    ring_free(&ring);
    ktm_rt_stream_destroy(s);
    if (fclose(f) != 0 && st == KTM_OK) st = KTM_ERR_IO;
    return st;
}

/* One pass over all traces in blocks of ntblock (Kirchhoff.c:1276-1408),
 * for every job (job < 0) or one job.  Traces come from the host arrays
 * (traces/coords) or, when path != NULL, from the SEG-Y file: each block of
 * raw records is read as stored and decoded on the device (ktm_ingest).
 * Two device slots and two streams overlap the input of block i + 1 (host
 * read, copy, decode and prep on the input stream) with the migration of
 * block i (migration stream); blocks are migrated in order, so the image is
 * the sequential one bit for bit.  mig_ms sums the migration launches
 * (events on the migration stream). */
static int run_survey(const ktm_plan *plan, const ktm_config *cfg, const char *path,
                      const float *traces, const int32_t *coords, int64_t ntr, int job,
                      dev_bufs *b, uint64_t *contrib, double *mig_ms)
{
    ktm_setup s;
    ktm_plan_setup(plan, &s);
    const int64_t blk = cfg->ntblock > 0 ? cfg->ntblock : 50000;
    const int64_t row = ktm_plan_trace_row(plan);
    const int64_t nb = ntr < blk ? ntr : blk;
    const int64_t rec = SEGY_TRACE_HDR + 4 * (int64_t)s.ns;
    const int64_t nblocks = ntr > 0 ? (ntr + blk - 1) / blk : 0;
    struct {
        void *raw;
        float *tr, *prepared;
        int32_t *co;
        ktm_trace_geom *geom;
        void *ready, *freed;               /* events: input done, migration done */
    } slot[2];
    memset(slot, 0, sizeof slot);
    void *s_in = NULL, *s_mig = NULL;
    void **tev = NULL;                     /* [2 * nblocks] migration start/end */
    stage_ring ring;
    memset(&ring, 0, sizeof ring);
    FILE *f = NULL;
    int st = KTM_OK;
    if (path && !(f = fopen(path, "rb"))) st = KTM_ERR_IO;
    if (st == KTM_OK)
        st = ring_init(&ring, path && (size_t)rec > KTM_STAGE_BYTES ? (size_t)rec : KTM_STAGE_BYTES);
    if (st == KTM_OK) st = ktm_rt_stream_create(&s_in);
    if (st == KTM_OK) st = ktm_rt_stream_create(&s_mig);
    if (st == KTM_OK && !(tev = (void **)calloc((size_t)(2 * nblocks + 1), sizeof(void *))))
        st = KTM_ERR_NOMEM;
    for (int64_t e = 0; e < 2 * nblocks && st == KTM_OK; ++e) st = ktm_rt_event_create(&tev[e]);
    for (int k = 0; k < 2 && k < nblocks && st == KTM_OK; ++k) {
        if (path) {
            st = ktm_device_alloc(&slot[k].raw, (size_t)(nb * rec));
        } else {
            st = ktm_device_alloc((void **)&slot[k].tr, sizeof(float) * (size_t)(nb * s.ns));
            if (st == KTM_OK)
                st = ktm_device_alloc((void **)&slot[k].co, sizeof(int32_t) * (size_t)(4 * nb));
        }
        if (st == KTM_OK)
            st = ktm_device_alloc((void **)&slot[k].prepared, sizeof(float) * (size_t)(nb * row));
        if (st == KTM_OK)
            st = ktm_device_alloc((void **)&slot[k].geom, sizeof(ktm_trace_geom) * (size_t)nb);
        if (st == KTM_OK) st = ktm_rt_event_create(&slot[k].ready);
        if (st == KTM_OK) st = ktm_rt_event_create(&slot[k].freed);
    }
    if (st == KTM_OK) st = ktm_device_alloc((void **)&b->counter, sizeof(uint64_t));
    uint64_t zero = 0;
    if (st == KTM_OK) st = ktm_copy_to_device(b->counter, &zero, sizeof zero);
    for (int64_t bi = 0; bi < nblocks && st == KTM_OK; ++bi) {
        const int64_t i0 = bi * blk;
        const int64_t n = ntr - i0 < blk ? ntr - i0 : blk;
        const int k = (int)(bi & 1);
        /* the slot's buffers are free once the migration of block bi - 2 ends */
        if (bi >= 2) st = ktm_rt_stream_wait(s_in, slot[k].freed);
        if (path) {
            if (st == KTM_OK && fseeko(f, SEGY_FILE_HDR + i0 * rec, SEEK_SET) != 0) st = KTM_ERR_IO;
            if (st == KTM_OK)
                st = ring_put(&ring, slot[k].raw, NULL, f, (size_t)(n * rec), (size_t)rec, 1, s_in);
            if (st == KTM_OK)
                st = ktm_ingest(plan, slot[k].raw, n, KTM_SEGY_IBM, KTM_PREP_REFERENCE,
                                KTM_TRACE_F32, slot[k].prepared, slot[k].geom, NULL, s_in);
        } else {
            const float *src_tr = traces + i0 * s.ns;
            const int32_t *src_co = coords + 4 * i0;
            for (int64_t q = 0; q < 4 * n && st == KTM_OK; ++q)
                if (src_co[q] < 0) st = KTM_ERR_ARG; /* htonl() - ox on a negative word is undefined */
            if (st == KTM_OK)
                st = ring_put(&ring, slot[k].tr, src_tr, NULL, sizeof(float) * (size_t)(n * s.ns), 0,
                              0, s_in);
            if (st == KTM_OK)
                st = ring_put(&ring, slot[k].co, src_co, NULL, sizeof(int32_t) * (size_t)(4 * n), 0,
                              0, s_in);
            if (st == KTM_OK)
                st = ktm_prep(plan, slot[k].tr, slot[k].co, n, KTM_PREP_REFERENCE, KTM_TRACE_F32,
                              slot[k].prepared, slot[k].geom, s_in);
        }
        if (st == KTM_OK) st = ktm_rt_event_record(slot[k].ready, s_in);
        if (st == KTM_OK) st = ktm_rt_stream_wait(s_mig, slot[k].ready);
        if (st == KTM_OK) st = ktm_rt_event_record(tev[2 * bi], s_mig);
        if (st == KTM_OK)
            st = ktm_migrate_block(plan, slot[k].prepared, KTM_TRACE_F32, KTM_PREP_REFERENCE,
                                   slot[k].geom, n, KTM_ACCUM_PLAIN, job, b->km, b->fold, NULL,
                                   b->counter, s_mig);
        if (st == KTM_OK) st = ktm_rt_event_record(tev[2 * bi + 1], s_mig);
        if (st == KTM_OK) st = ktm_rt_event_record(slot[k].freed, s_mig);
    }
    /* drain both streams before anything is freed, on success or not */
    if (s_in && ktm_rt_stream_sync(s_in) != KTM_OK && st == KTM_OK) st = KTM_ERR_DEVICE;
    if (s_mig && ktm_rt_stream_sync(s_mig) != KTM_OK && st == KTM_OK) st = KTM_ERR_DEVICE;
    double t_mig = 0.0;
    for (int64_t bi = 0; bi < nblocks && st == KTM_OK; ++bi) {
        float ms = 0.0f;
        st = ktm_rt_event_elapsed_ms(tev[2 * bi], tev[2 * bi + 1], &ms);
        t_mig += ms;
    }
    if (st == KTM_OK) st = ktm_copy_to_host(contrib, b->counter, sizeof(uint64_t));
    *mig_ms = t_mig;
    ring_free(&ring);
    for (int k = 0; k < 2; ++k) {
        ktm_device_free(slot[k].raw);
        ktm_device_free(slot[k].tr);
        ktm_device_free(slot[k].co);
        ktm_device_free(slot[k].prepared);
        ktm_device_free(slot[k].geom);
        ktm_rt_event_destroy(slot[k].ready);
        ktm_rt_event_destroy(slot[k].freed);
    }
    for (int64_t e = 0; tev && e < 2 * nblocks; ++e) ktm_rt_event_destroy(tev[e]);
    free(tev);
    ktm_rt_stream_destroy(s_in);
    ktm_rt_stream_destroy(s_mig);
    if (f) fclose(f);
    return st;
}

static int migrate_arrays(const ktm_config *cfg, int ns, float dt, const float *traces,
                          const int32_t *coords, int64_t ntr, const float *vel, int nsm, float dtm,
                          float *km, float *fold, float *image, ktm_stats *stats)
{
    if (!cfg || ntr < 0 || (ntr > 0 && (!traces || !coords)) || !vel) return KTM_ERR_ARG;
    const double t_start = now_ms();
    ktm_plan *plan = NULL;
    int st = ktm_plan_create(&plan, cfg, ns, dt, nsm, dtm, vel);
    if (st != KTM_OK) return st;
    const size_t nimg = (size_t)ktm_plan_image_size(plan);
    dev_bufs b;
    memset(&b, 0, sizeof b);
    st = ktm_device_alloc((void **)&b.km, sizeof(float) * nimg);
    if (st == KTM_OK) st = ktm_device_alloc((void **)&b.fold, sizeof(float) * nimg);
    if (st == KTM_OK) st = ktm_image_reset(plan, b.km, b.fold, NULL, NULL);
    uint64_t contrib = 0;
    double mig_ms = 0.0;
    if (st == KTM_OK)
        st = run_survey(plan, cfg, NULL, traces, coords, ntr, -1, &b, &contrib, &mig_ms);
    if (st == KTM_OK && image) {
        st = ktm_device_alloc((void **)&b.image, sizeof(float) * nimg);
        if (st == KTM_OK) st = ktm_poststack(plan, b.km, b.fold, b.image, NULL);
        if (st == KTM_OK) st = ktm_copy_to_host(image, b.image, sizeof(float) * nimg);
    }
    if (st == KTM_OK && km) st = ktm_copy_to_host(km, b.km, sizeof(float) * nimg);
    if (st == KTM_OK && fold) st = ktm_copy_to_host(fold, b.fold, sizeof(float) * nimg);
    free_bufs(&b);
    ktm_plan_destroy(plan);
    if (stats) {
        stats->contributions = contrib;
        stats->migrate_ms = mig_ms;
        stats->total_ms = now_ms() - t_start;
    }
    return st;
}

int ktm_migrate(const ktm_config *cfg, const ktm_traces *traces, const ktm_velocity *vel,
                ktm_image *out, ktm_stats *stats)
{
    if (!traces || !vel) return KTM_ERR_ARG;
    return migrate_arrays(cfg, traces->ns, traces->dt, traces->samples, traces->coords,
                          traces->ntr, vel->v, vel->nsm, vel->dtm, out ? out->km : NULL,
                          out ? out->fold : NULL, out ? out->image : NULL, stats);
}

/* ------------------------------------------------------------------------ */
/* kirchhoff(job_id): the reference signature                                */
/* ------------------------------------------------------------------------ */
static struct {
    int ready, status;
    ktm_config cfg;
    char input_data[4096], image_dir[4096];
    int ns, nt;
    float dt;
    ktm_plan *plan;
    ktm_stats stats;
} G;

int ktm_last_status(void) { return G.status; }

int ktm_last_stats(ktm_stats *out)
{
    if (!out) return KTM_ERR_ARG;
    *out = G.stats;
    return KTM_OK;
}

void ktm_globals_free(void)
{
    ktm_plan_destroy(G.plan);
    memset(&G, 0, sizeof G);
}

int ktm_globals_init(const ktm_config *cfg, const char *input_data, const char *input_vel,
                     const float *vel, int nsm, float dtm, const char *image_dir)
{
    ktm_globals_free();
    if (!cfg || !input_data || !image_dir || (!input_vel && !vel)) return G.status = KTM_ERR_ARG;
    if (strlen(input_data) >= sizeof G.input_data || strlen(image_dir) >= sizeof G.image_dir)
        return G.status = KTM_ERR_ARG;
    G.cfg = *cfg;
    strcpy(G.input_data, input_data);
    strcpy(G.image_dir, image_dir);
    /* count_size_file(), Kirchhoff.c:263 */
    int st = ktm_segy_count(input_data, &G.ns, &G.dt, &G.nt);
    if (st != KTM_OK) return G.status = st;
    /* mkdir image_dir (Kirchhoff.c:254-260); an existing directory is fine */
    if (mkdir(image_dir, 0777) != 0 && errno != EEXIST) return G.status = KTM_ERR_IO;
    /* build_vel_model(), Kirchhoff.c:328 */
    float *vbuf = NULL;
    if (input_vel) {
        st = ktm_segy_read_velocity(input_vel, cfg, NULL, &nsm, &dtm, NULL);
        if (st != KTM_OK) return G.status = st;
        vbuf = (float *)malloc(sizeof(float) * (size_t)cfg->nvel_ilines * cfg->nvel_xlines * nsm);
        if (!vbuf) return G.status = KTM_ERR_NOMEM;
        st = ktm_segy_read_velocity(input_vel, cfg, vbuf, &nsm, &dtm, NULL);
        if (st != KTM_OK) { free(vbuf); return G.status = st; }
        vel = vbuf;
    }
    st = ktm_plan_create(&G.plan, cfg, G.ns, G.dt, nsm, dtm, vel);
    free(vbuf);
    if (st != KTM_OK) return G.status = st;
    /* ebdic_header.bin / binary_header.bin, Kirchhoff.c:321-322, 729-814 */
    uint8_t hdr[3600];
    ktm_segy_file_header(cfg, G.ns, G.dt, 0, hdr);
    char name[4200];
    snprintf(name, sizeof name, "%s/ebdic_header.bin", image_dir);
    FILE *f = fopen(name, "wb");
    if (!f) return G.status = KTM_ERR_IO;
    int ok = fwrite(hdr, 1, 3200, f) == 3200;
    ok &= fclose(f) == 0;
    snprintf(name, sizeof name, "%s/binary_header.bin", image_dir);
    f = fopen(name, "wb");
    if (!f) return G.status = KTM_ERR_IO;
    ok &= fwrite(hdr + 3200, 1, 400, f) == 400;
    ok &= fclose(f) == 0;
    if (!ok) return G.status = KTM_ERR_IO;
    G.ready = 1;
    return G.status = KTM_OK;
}

static int kirchhoff_job(int job_id)
{
    memset(&G.stats, 0, sizeof G.stats);
    if (!G.ready) return KTM_ERR_ARG;
    const double t_start = now_ms();
    ktm_setup s;
    ktm_plan_setup(G.plan, &s);
    if (job_id < 0 || job_id >= s.no_job) return KTM_ERR_ARG;
    const size_t nimg = (size_t)ktm_plan_image_size(G.plan);
    dev_bufs b;
    memset(&b, 0, sizeof b);
    int st = ktm_device_alloc((void **)&b.km, sizeof(float) * nimg);
    if (st == KTM_OK) st = ktm_device_alloc((void **)&b.fold, sizeof(float) * nimg);
    if (st == KTM_OK) st = ktm_device_alloc((void **)&b.image, sizeof(float) * nimg);
    if (st == KTM_OK) st = ktm_image_reset(G.plan, b.km, b.fold, NULL, NULL);
    uint64_t contrib = 0;
    double ms = 0.0;
    if (st == KTM_OK)
        st = run_survey(G.plan, &G.cfg, G.input_data, NULL, NULL, G.nt, job_id, &b, &contrib, &ms);
    G.stats.contributions = contrib;
    G.stats.migrate_ms = ms;
    if (st == KTM_OK) st = ktm_poststack(G.plan, b.km, b.fold, b.image, NULL);
    /* image<job>.su: per (xc, yc, ih) a 240-byte header and ns big-endian
     * IEEE floats, in the order of Kirchhoff.c:1440-1598, encoded on the
     * device and written as one block */
    const size_t nrec = (size_t)s.nx * s.ny * s.nh;
    const size_t bytes = nrec * (SEGY_TRACE_HDR + 4 * (size_t)s.ns);
    void *drec = NULL;
    if (st == KTM_OK) st = ktm_device_alloc(&drec, bytes);
    if (st == KTM_OK) st = ktm_image_records(G.plan, b.image, job_id, drec, NULL);
    char name[4200];
    snprintf(name, sizeof name, "%s/image%d.su", G.image_dir, job_id);
    /* the records stream from the device into the file (a stream ordered
     * after the records kernel on the null stream) */
    if (st == KTM_OK) st = write_device_file(name, drec, bytes);
    ktm_device_free(drec);
    free_bufs(&b);
    G.stats.total_ms = now_ms() - t_start;
    return st;
}

void kirchhoff(int job_id) { G.status = kirchhoff_job(job_id); }

/* ------------------------------------------------------------------------ */
/* csp(job_id): the EOM.c signature                                          */
/* ------------------------------------------------------------------------ */
static struct {
    int ready, status;
    ktm_eom_config cfg;
    char input_data[4096], image_dir[4096];
    int ns, nt;
    float dt;
    ktm_eom_plan *plan;
} E;

int ktm_eom_last_status(void) { return E.status; }

void ktm_eom_globals_free(void)
{
    ktm_eom_plan_destroy(E.plan);
    memset(&E, 0, sizeof E);
}

static int write_whole(const char *path, const void *data, size_t n)
{
    FILE *f = fopen(path, "wb");
    if (!f) return KTM_ERR_IO;
    int ok = fwrite(data, 1, n, f) == n;
    ok &= fclose(f) == 0;
    return ok ? KTM_OK : KTM_ERR_IO;
}

int ktm_eom_globals_init(const ktm_eom_config *cfg, const char *input_data, const char *image_dir)
{
    ktm_eom_globals_free();
    if (!cfg || !input_data || !image_dir) return E.status = KTM_ERR_ARG;
    if (strlen(input_data) >= sizeof E.input_data || strlen(image_dir) >= sizeof E.image_dir)
        return E.status = KTM_ERR_ARG;
    E.cfg = *cfg;
    strcpy(E.input_data, input_data);
    strcpy(E.image_dir, image_dir);
    /* count_size_file(), EOM.c:276 and 398-444 */
    int st = ktm_segy_count(input_data, &E.ns, &E.dt, &E.nt);
    if (st != KTM_OK) return E.status = st;
    /* mkdir image_dir (EOM.c:330-337); an existing directory is fine */
    if (mkdir(image_dir, 0777) != 0 && errno != EEXIST) return E.status = KTM_ERR_IO;
    /* build_vel_model() with velocity_model 0 (EOM.c:869-878): velocityStart
     * + (velocityEnd - velocityStart) * (i / ns) in int arithmetic, where
     * i / ns is 0 for every sample (squared by the plan) */
    float *vel = (float *)malloc(sizeof(float) * (size_t)E.ns);
    if (!vel) return E.status = KTM_ERR_NOMEM;
    for (int i = 0; i < E.ns; ++i)
        vel[i] = (float)(cfg->velocity_start + (cfg->velocity_end - cfg->velocity_start) * (i / E.ns));
    /* the CSP grid (EOM.c:279-320) and the filters (create_lowpass_filter,
     * create_sqrho_filter, EOM.c:340-343) are set up by the plan */
    st = ktm_eom_plan_create(&E.plan, cfg, E.ns, E.dt, vel);
    free(vel);
    if (st != KTM_OK) return E.status = st;
    /* write_segy_ebdic_header(), EOM.c:350 and 450-537: the text of
     * EOM.c:480 (the rest of the 3200 bytes, uninitialised there, as 0) and
     * the binary header (the uninitialised short of EOM.c:503-505 as 0) */
    uint8_t hdr[SEGY_FILE_HDR];
    memset(hdr, 0, sizeof hdr);
    static const char text[] = "written by G. O'Brien\n";
    memcpy(hdr, text, sizeof text - 1);
    uint8_t *bin = hdr + 3200;
    ktm_put_be16(bin + 16, (uint16_t)(E.dt * 1000000));
    ktm_put_be16(bin + 20, (uint16_t)E.ns);
    ktm_put_be16(bin + 22, (uint16_t)E.ns);
    ktm_put_be16(bin + 24, 5);
    ktm_put_be16(bin + 26, 1);
    ktm_put_be16(bin + 28, 1);
    ktm_put_be16(bin + 30, 1);
    char name[4200];
    snprintf(name, sizeof name, "%s/ebdic_header.bin", image_dir);
    st = write_whole(name, hdr, 3200);
    snprintf(name, sizeof name, "%s/binary_header.bin", image_dir);
    if (st == KTM_OK) st = write_whole(name, hdr + 3200, 400);
    if (st != KTM_OK) return E.status = st;
    E.ready = 1;
    return E.status = KTM_OK;
}

static int eom_csp_job(int job_id)
{
    if (!E.ready) return KTM_ERR_ARG;
    ktm_eom_setup s;
    ktm_eom_plan_setup(E.plan, &s);
    if (job_id < 0 || job_id >= s.no_job) return KTM_ERR_ARG;
    const size_t ng = (size_t)ktm_eom_job_size(E.plan);
    const int64_t blk = E.cfg.ntblock > 0 ? E.cfg.ntblock : 20000;
    const int64_t nb = E.nt < blk ? E.nt : blk;
    const int64_t rec = SEGY_TRACE_HDR + 4 * (int64_t)s.ns;
    const size_t nrec_max = (size_t)s.nx * s.ny * s.nh;
    const int64_t nblocks = E.nt > 0 ? (E.nt + blk - 1) / blk : 0;
    float *d_csp = NULL, *d_fold = NULL, *d_out = NULL;
    void *d_rec = NULL;
    struct {
        void *raw;
        float *tr;
        int32_t *co;
        void *ready, *freed;               /* events: input done, binning done */
    } slot[2];
    memset(slot, 0, sizeof slot);
    void *s_in = NULL, *s_bin = NULL;
    stage_ring ring;
    memset(&ring, 0, sizeof ring);
    FILE *f = NULL;
    /* csp_data = 0, fold = 1 for the job's slab (EOM.c:1192-1211) */
    int st = ktm_device_alloc((void **)&d_csp, sizeof(float) * ng);
    if (st == KTM_OK) st = ktm_device_alloc((void **)&d_fold, sizeof(float) * ng);
    if (st == KTM_OK) st = ktm_eom_job_reset(E.plan, d_csp, d_fold, NULL);
    /* the trace loop in blocks of ntblock (EOM.c:1221-1316): raw records
     * read as stored, IBM-decoded and binned on the device; the input of
     * block i + 1 (read, copy, decode on one stream) overlaps the binning of
     * block i (another stream), the blocks binned in order */
    if (st == KTM_OK && nblocks > 0) {
        if (!(f = fopen(E.input_data, "rb"))) st = KTM_ERR_IO;
        if (st == KTM_OK)
            st = ring_init(&ring, (size_t)rec > KTM_STAGE_BYTES ? (size_t)rec : KTM_STAGE_BYTES);
        if (st == KTM_OK) st = ktm_rt_stream_create(&s_in);
        if (st == KTM_OK) st = ktm_rt_stream_create(&s_bin);
        for (int k = 0; k < 2 && k < nblocks && st == KTM_OK; ++k) {
            st = ktm_device_alloc(&slot[k].raw, (size_t)(nb * rec));
            if (st == KTM_OK)
                st = ktm_device_alloc((void **)&slot[k].tr, sizeof(float) * (size_t)(nb * s.ns));
            if (st == KTM_OK)
                st = ktm_device_alloc((void **)&slot[k].co, sizeof(int32_t) * (size_t)(4 * nb));
            if (st == KTM_OK) st = ktm_rt_event_create(&slot[k].ready);
            if (st == KTM_OK) st = ktm_rt_event_create(&slot[k].freed);
        }
    }
    for (int64_t bi = 0; bi < nblocks && st == KTM_OK; ++bi) {
        const int64_t i0 = bi * blk;
        const int64_t n = E.nt - i0 < blk ? E.nt - i0 : blk;
        const int k = (int)(bi & 1);
        if (bi >= 2) st = ktm_rt_stream_wait(s_in, slot[k].freed);
        if (st == KTM_OK && fseeko(f, SEGY_FILE_HDR + i0 * rec, SEEK_SET) != 0) st = KTM_ERR_IO;
        if (st == KTM_OK)
            st = ring_put(&ring, slot[k].raw, NULL, f, (size_t)(n * rec), (size_t)rec, 0, s_in);
        if (st == KTM_OK)
            st = ktm_segy_decode(slot[k].raw, n, s.ns, KTM_SEGY_IBM, slot[k].tr, slot[k].co, s_in);
        if (st == KTM_OK) st = ktm_rt_event_record(slot[k].ready, s_in);
        if (st == KTM_OK) st = ktm_rt_stream_wait(s_bin, slot[k].ready);
        if (st == KTM_OK)
            st = ktm_eom_job_block(E.plan, slot[k].tr, slot[k].co, n, job_id, d_csp, d_fold, NULL,
                                   s_bin);
        if (st == KTM_OK) st = ktm_rt_event_record(slot[k].freed, s_bin);
    }
    if (s_in && ktm_rt_stream_sync(s_in) != KTM_OK && st == KTM_OK) st = KTM_ERR_DEVICE;
    if (s_bin && ktm_rt_stream_sync(s_bin) != KTM_OK && st == KTM_OK) st = KTM_ERR_DEVICE;
    ring_free(&ring);
    for (int k = 0; k < 2; ++k) {
        ktm_device_free(slot[k].raw);
        ktm_device_free(slot[k].tr);
        ktm_device_free(slot[k].co);
        ktm_rt_event_destroy(slot[k].ready);
        ktm_rt_event_destroy(slot[k].freed);
    }
    ktm_rt_stream_destroy(s_in);
    ktm_rt_stream_destroy(s_bin);
    if (f) fclose(f);
    /* scale, low-pass and sqrt-rho filter (EOM.c:1386-1413), then the
     * records of image<job>.su (EOM.c:1346-1537), all on the device */
    if (st == KTM_OK) st = ktm_device_alloc((void **)&d_out, sizeof(float) * ng);
    if (st == KTM_OK) st = ktm_eom_job_post(E.plan, d_csp, d_fold, d_out, NULL);
    ktm_device_free(d_csp);
    ktm_device_free(d_fold);
    int64_t nrec = 0;
    if (st == KTM_OK) st = ktm_device_alloc(&d_rec, nrec_max * (size_t)rec);
    if (st == KTM_OK) st = ktm_eom_job_records(E.plan, d_out, job_id, d_rec, &nrec, NULL);
    if (st == KTM_OK) {
        char name[4200];
        snprintf(name, sizeof name, "%s/image%d.su", E.image_dir, job_id);
        st = write_device_file(name, d_rec, (size_t)nrec * (size_t)rec);
    }
    ktm_device_free(d_out);
    ktm_device_free(d_rec);
    return st;
}

void csp(int job_id) { E.status = eom_csp_job(job_id); }

/* ------------------------------------------------------------------------ */
/* combine_all_files(): Kirchhoff.c:821-874, EOM.c:543-590                    */
/* ------------------------------------------------------------------------ */
int ktm_combine_script(const char *image_dir, const char *output_data, int no_job)
{
    if (!image_dir || !output_data || no_job < 0) return KTM_ERR_ARG;
    char name[4200];
    if (snprintf(name, sizeof name, "%s/combine_script", image_dir) >= (int)sizeof name)
        return KTM_ERR_ARG;
    FILE *f = fopen(name, "wb");
    if (!f) return KTM_ERR_IO;
    /* the lines of Kirchhoff.c:860-867: both header files, then every job's
     * image in job order, then the parts removed (no final newline) */
    int ok = fprintf(f, "cat ebdic_header.bin binary_header.bin > %s \n", output_data) > 0;
    for (int i = 0; i < no_job && ok; ++i) ok = fprintf(f, "cat image%d.su >> %s\n", i, output_data) > 0;
    ok = ok && fprintf(f, "rm image*.su\n") > 0;
    ok = ok && fprintf(f, "rm ebdic_header.bin binary_header.bin ") > 0;
    ok &= fclose(f) == 0;
    if (!ok) return KTM_ERR_IO;
    /* chmod a+x (Kirchhoff.c:869-870) */
    struct stat sb;
    if (stat(name, &sb) != 0 || chmod(name, (sb.st_mode & 07777) | 0111) != 0) return KTM_ERR_IO;
    return KTM_OK;
}

static int append_file(FILE *out, const char *path, uint8_t *buf, size_t cap)
{
    FILE *in = fopen(path, "rb");
    if (!in) return KTM_ERR_IO;
    int st = KTM_OK;
    for (;;) {
        const size_t n = fread(buf, 1, cap, in);
        if (n && fwrite(buf, 1, n, out) != n) { st = KTM_ERR_IO; break; }
        if (n < cap) {
            if (ferror(in)) st = KTM_ERR_IO;
            break;
        }
    }
    fclose(in);
    return st;
}

int ktm_combine_files(const char *image_dir, const char *output_data, int no_job)
{
    if (!image_dir || !output_data || no_job < 0) return KTM_ERR_ARG;
    const size_t cap = (size_t)1 << 22;
    uint8_t *buf = (uint8_t *)malloc(cap);
    if (!buf) return KTM_ERR_NOMEM;
    FILE *out = fopen(output_data, "wb");
    if (!out) { free(buf); return KTM_ERR_IO; }
    char name[4200];
    int st = KTM_OK;
    snprintf(name, sizeof name, "%s/ebdic_header.bin", image_dir);
    st = append_file(out, name, buf, cap);
    snprintf(name, sizeof name, "%s/binary_header.bin", image_dir);
    if (st == KTM_OK) st = append_file(out, name, buf, cap);
    for (int i = 0; i < no_job && st == KTM_OK; ++i) {
        snprintf(name, sizeof name, "%s/image%d.su", image_dir, i);
        st = append_file(out, name, buf, cap);
    }
    if (fclose(out) != 0 && st == KTM_OK) st = KTM_ERR_IO;
    free(buf);
    return st;
}
