/*
 * kirchhoff_main.c -- the main() of Kirchhoff.c (Kirchhoff.c:215-373) on this
 * library: count_size_file, create_aliasfilter, build_vel_model, mkdir and the
 * header files (ktm_globals_init), the round-robin job loop over
 * kirchhoff(job_id) (Kirchhoff.c:347-350) and combine_all_files
 * (Kirchhoff.c:355).  The reference fixes its configuration in compile-time
 * macros (Kirchhoff.c:17-65) and its paths in string macros (Kirchhoff.c:21,
 * 30-31); here the paths are arguments and any macro can be overridden as
 * name=value with the ktm_config field name (the rest keep the reference's
 * values).  rank/size replace MPI_Comm_rank/size: each process runs jobs
 * rank, rank + size, ... and rank 0 writes the combine step.
 *
 *   kirchhoff_main input.sgy velocity.sgy image_dir output.sgy
 *                  [rank=R size=S] [name=value ...]
 *
 * Exit status: 0, 1 on a library error (its status printed), 2 on bad usage.
 */
#include <stddef.h>
#include <stdio.h>
#include <stdlib.h>
#include <string.h>

#include "ktm.h"

enum { F_I32, F_F64 };
typedef struct field {
    const char *name;
    size_t off;
    int type;
} field;

#define FI(n) {#n, offsetof(ktm_config, n), F_I32}
#define FD(n) {#n, offsetof(ktm_config, n), F_F64}
static const field FIELDS[] = {
    FI(start_iline), FI(end_iline), FI(start_xline), FI(end_xline), FD(inc_iline),
    FD(inc_xline), FD(dh), FI(nh), FD(oh), FD(ox), FD(oy), FD(angle_e_to_n),
    FD(iline_distance), FD(xline_distance), FD(aperture), FD(anti_alias_cutoff),
    FI(nvel_ilines), FI(nvel_xlines), FI(dvel_ilines), FI(dvel_xlines), FI(ovel_ilines),
    FI(ovel_xlines), FI(ntblock), FI(jobs_x), FI(jobs_y), FI(image_origin_zero),
    FI(vel_lower_clamp), FD(scale_xy),
};

/* name=value into cfg; 0 on success */
static int set_field(ktm_config *cfg, const char *arg)
{
    const char *eq = strchr(arg, '=');
    if (!eq) return -1;
    const size_t n = (size_t)(eq - arg);
    for (size_t i = 0; i < sizeof FIELDS / sizeof FIELDS[0]; ++i) {
        if (strlen(FIELDS[i].name) != n || strncmp(FIELDS[i].name, arg, n) != 0) continue;
        char *end = NULL;
        char *p = (char *)cfg + FIELDS[i].off;
        if (FIELDS[i].type == F_I32) {
            const long v = strtol(eq + 1, &end, 10);
            if (*end) return -1;
            *(int32_t *)p = (int32_t)v;
        } else {
            const double v = strtod(eq + 1, &end);
            if (*end) return -1;
            *(double *)p = v;
        }
        return 0;
    }
    return -1;
}

int main(int argc, char **argv)
{
    if (argc < 5) {
        fprintf(stderr, "usage: %s input.sgy velocity.sgy image_dir output.sgy "
                        "[rank=R size=S] [name=value ...]\n", argv[0]);
        return 2;
    }
    const char *input_data = argv[1], *input_vel = argv[2], *image_dir = argv[3],
               *output_data = argv[4];
    ktm_config cfg;
    ktm_config_default(&cfg);
    int rank = 0, size = 1;
    for (int a = 5; a < argc; ++a) {
        if (!strncmp(argv[a], "rank=", 5)) rank = atoi(argv[a] + 5);
        else if (!strncmp(argv[a], "size=", 5)) size = atoi(argv[a] + 5);
        else if (set_field(&cfg, argv[a]) != 0) {
            fprintf(stderr, "unknown or malformed setting: %s\n", argv[a]);
            return 2;
        }
    }
    if (size < 1 || rank < 0 || rank >= size) {
        fprintf(stderr, "need 0 <= rank < size\n");
        return 2;
    }
    int st = ktm_globals_init(&cfg, input_data, input_vel, NULL, 0, 0.0f, image_dir);
    if (st != KTM_OK) {
        fprintf(stderr, "setup failed: status %d\n", st);
        return 1;
    }
    const int no_job = cfg.jobs_x * cfg.jobs_y;
    for (int job = rank; job < no_job; job += size) {
        kirchhoff(job);
        if ((st = ktm_last_status()) != KTM_OK) {
            fprintf(stderr, "kirchhoff(%d) failed: status %d\n", job, st);
            ktm_globals_free();
            return 1;
        }
    }
    if (rank == 0) {
        st = ktm_combine_script(image_dir, output_data, no_job);
        if (st == KTM_OK && size == 1) st = ktm_combine_files(image_dir, output_data, no_job);
        if (st != KTM_OK) {
            fprintf(stderr, "combine failed: status %d\n", st);
            ktm_globals_free();
            return 1;
        }
    }
    ktm_globals_free();
    return 0;
}
