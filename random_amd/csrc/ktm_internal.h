/*
 * ktm_internal.h -- host tables shared by the C setup code and the HIP glue.
 */
#ifndef KTM_INTERNAL_H
#define KTM_INTERNAL_H

#include "ktm.h"

#ifdef __cplusplus
extern "C" {
#endif

/* Everything the kernels need that depends only on the configuration, built
 * on the host with the promotions and rounding of the reference's own
 * expressions (compiled with -ffp-contract=off). */
typedef struct ktm_tables {
    ktm_setup s;
    float   *anti_alias;  /* [ns]   create_aliasfilter, Kirchhoff.c:401-423          */
    float   *t2;          /* [ns]   0.25 t^2 dt^2 as float, Kirchhoff.c:1246-1247      */
    int32_t *t1;          /* [ns]   velocity time index, Kirchhoff.c:1376-1378         */
    float   *colx;        /* [nxp]  image inline coordinate IX, Kirchhoff.c:1346       */
    float   *coly;        /* [nyp]  image crossline coordinate IY, Kirchhoff.c:1347    */
    int32_t *velx;        /* [nxp]  velocity grid inline index, Kirchhoff.c:1357-1372  */
    int32_t *vely;        /* [nyp]  velocity grid xline index, Kirchhoff.c:1358-1372   */
    float   *slab_cx;     /* [no_job] slab centre IX, Kirchhoff.c:1332                */
    float   *slab_cy;     /* [no_job] slab centre IY, Kirchhoff.c:1333                */
    double  *slab_thr;    /* [no_job] aperture + add_on, Kirchhoff.c:1335-1338        */
    float   *slab_x2thr;  /* [no_job] the same gate on the squared distance (exact)  */
    float    ap_x2thr;    /* aperture gate of Kirchhoff.c:1354 on the squared distance */
    float   *w;           /* [ns+2] CLOSED-mode sample weights: x[t] = tr[min(t,ns-1)]*w[t] */
    double  *hconv;       /* [ns]   circulant post-stack kernel, real IDFT of anti_alias / ns */
} ktm_tables;

int  ktm_tables_build(const ktm_config *cfg, int ns, float dt, int nsm, float dtm,
                      ktm_tables *tb);
void ktm_tables_free(ktm_tables *tb);

/* Velocity index of one image column along one axis (0 = inline, 1 = xline),
 * -1 where the reference would read below the model. */
int  ktm_vel_index(const ktm_config *cfg, int col, int axis);

/* Host pipeline runtime for the C host code (ktm_device.hip): pinned host
 * memory, streams at the default priority that are ordered with work on the
 * null stream (hipStreamCreate), and timing events.  Handles are opaque. */
int  ktm_rt_pinned_alloc(void **p, size_t bytes);
void ktm_rt_pinned_free(void *p);
int  ktm_rt_stream_create(void **s);
void ktm_rt_stream_destroy(void *s);
int  ktm_rt_stream_sync(void *s);
int  ktm_rt_stream_wait(void *s, void *e);          /* s waits for e's work */
int  ktm_rt_event_create(void **e);
void ktm_rt_event_destroy(void *e);
int  ktm_rt_event_record(void *e, void *s);
int  ktm_rt_event_sync(void *e);                     /* the host waits for e */
int  ktm_rt_event_elapsed_ms(void *e0, void *e1, float *ms);
int  ktm_rt_copy_h2d_async(void *dst, const void *src, size_t bytes, void *s);
int  ktm_rt_copy_d2h_async(void *dst, const void *src, size_t bytes, void *s);

/* Shared by the C host code: big-endian helpers. */
static inline void ktm_put_be32(uint8_t *o, uint32_t u)
{
    o[0] = (uint8_t)(u >> 24); o[1] = (uint8_t)(u >> 16);
    o[2] = (uint8_t)(u >> 8);  o[3] = (uint8_t)u;
}
static inline void ktm_put_be16(uint8_t *o, uint16_t v)
{
    o[0] = (uint8_t)(v >> 8); o[1] = (uint8_t)v;
}
static inline uint32_t ktm_get_be32(const uint8_t *b)
{
    return ((uint32_t)b[0] << 24) | ((uint32_t)b[1] << 16) | ((uint32_t)b[2] << 8) |
           (uint32_t)b[3];
}
static inline uint16_t ktm_get_be16(const uint8_t *b)
{
    return (uint16_t)(((uint16_t)b[0] << 8) | (uint16_t)b[1]);
}

#ifdef __cplusplus
}
#endif
#endif
