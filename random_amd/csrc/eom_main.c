/*
 * eom_main.c -- the main() of EOM.c (EOM.c:238-390) on this library: the CSP
 * grid, filters, velocity model 0, mkdir and header files
 * (ktm_eom_globals_init), the round-robin job loop over csp(job_id)
 * (EOM.c:367-370) and combine_all_files (EOM.c:374).  Paths are
 * arguments and the macros of EOM.c:60-99 can be overridden as name=value
 * with the ktm_eom_config field name; rank/size stand in for MPI.
 *
 *   eom_main input.sgy image_dir output.sgy [rank=R size=S] [name=value ...]
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

#define FI(n) {#n, offsetof(ktm_eom_config, n), F_I32}
#define FD(n) {#n, offsetof(ktm_eom_config, n), F_F64}
static const field FIELDS[] = {
    FD(ox), FD(oy), FD(angle_e_to_n), FD(oxcsp), FD(oycsp), FD(excsp), FD(eycsp),
    FD(scale_xy), FD(dcsp_x), FD(dcsp_y), FD(dh), FD(aperture), FD(low_pass_cutoff),
    FI(nh), FI(oh), FI(jobs_x), FI(jobs_y), FI(ntblock), FD(velocity_start), FD(velocity_end),
};

static int set_field(ktm_eom_config *cfg, const char *arg)
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
    if (argc < 4) {
        fprintf(stderr, "usage: %s input.sgy image_dir output.sgy [rank=R size=S] "
                        "[name=value ...]\n", argv[0]);
        return 2;
    }
    const char *input_data = argv[1], *image_dir = argv[2], *output_data = argv[3];
    ktm_eom_config cfg;
    ktm_eom_config_default(&cfg);
    int rank = 0, size = 1;
    for (int a = 4; a < argc; ++a) {
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
    int st = ktm_eom_globals_init(&cfg, input_data, image_dir);
    if (st != KTM_OK) {
        fprintf(stderr, "setup failed: status %d\n", st);
        return 1;
    }
    const int no_job = cfg.jobs_x * cfg.jobs_y;
    for (int job = rank; job < no_job; job += size) {
        csp(job);
        if ((st = ktm_eom_last_status()) != KTM_OK) {
            fprintf(stderr, "csp(%d) failed: status %d\n", job, st);
            ktm_eom_globals_free();
            return 1;
        }
    }
    if (rank == 0) {
        st = ktm_combine_script(image_dir, output_data, no_job);
        if (st == KTM_OK && size == 1) st = ktm_combine_files(image_dir, output_data, no_job);
        if (st != KTM_OK) {
            fprintf(stderr, "combine failed: status %d\n", st);
            ktm_eom_globals_free();
            return 1;
        }
    }
    ktm_eom_globals_free();
    return 0;
}
