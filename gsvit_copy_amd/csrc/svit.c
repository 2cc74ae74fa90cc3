/*
 * svit.c -- command-line driver: `gsvit_amd file.par [--killer file]`.
 * Follows the reference main loop (src3d/svit.c:109-275) with the GPU branch
 * always taken: parse, plan, sv_gpu_new (via the planner's GPU request),
 * sv_pool_gpu_init, copyto, then per step the drop-in entry points in the
 * reference order, output on output steps, and the killer-file check
 * (src3d/svit.c:90-107).
 */
#include <stdio.h>
#include <stdlib.h>
#include <string.h>
#include "gsvit_internal.h"

static int
killcheck(const SvSet *set)
{
    FILE *fr;
    int val = 0;
    if (!set->sc.killer[0])
        return 0;
    fr = fopen(set->sc.killer, "r");
    if (!fr)
        return 0;
    if (fscanf(fr, "%d", &val) != 1)
        val = 0;
    fclose(fr);
    return val == 1;
}

int
main(int argc, char *argv[])
{
    SvSet set;
    SvPool *mp;
    int i, a;

    if (argc < 2) {
        printf("Usage: gsvit_amd parameter_file [--killer file]\n");
        return 1;
    }
    clear_settings(&set, 0);
    if (!parse_settings(argv[1], &set, 1)) {
        printf("Settings error!\n");
        return 1;
    }
    for (a = 2; a + 1 < argc; a++)
        if (!strcmp(argv[a], "--killer"))
            snprintf(set.sc.killer, sizeof(set.sc.killer), "%s", argv[a + 1]);
    if (set.sc.verbose > 0)
        printf("Parameter file loaded successfully\n");
    set.sc.usegpu = 1;

    mp = init_and_make_plan(&set);
    if (!mp)
        return 1;
    if (!sv_gpu_new(mp, &set) || set.plan.gpumode != SV_GPUMODE_FULL) {
        fprintf(stderr, "Error: no GPU available; this engine has no CPU path\n");
        return 1;
    }
    if (sv_pool_gpu_init(mp, &set) || sv_pool_gpu_copyto(mp, &set, SV_SYNC_ALL))
        return 1;

    for (i = 0; i < set.sc.nsteps; i++) {
        int rc = 0;
        if (set.sc.verbose > 0) {
            printf("______________ step %d ______________\n", set.sc.step_act);
            fflush(stdout);
        }
        rc |= sv_pool_gpu_boundary_copy(mp, &set);
        rc |= sv_pool_gpu_ystep_h(mp, &set);
        rc |= sv_pool_gpu_apply_source_hstep(mp, &set);
        rc |= sv_pool_gpu_boundary_hstep(mp, &set);
        rc |= sv_pool_gpu_ystep_e(mp, &set);
        rc |= sv_pool_gpu_apply_source_estep(mp, &set);
        rc |= sv_pool_gpu_boundary_estep(mp, &set);
        rc |= sv_pool_gpu_getsum(mp, &set);
        rc |= sv_pool_gpu_apply_farfield(mp, &set);
        rc |= sv_pool_gpu_apply_pfarfield(mp, &set);
        rc |= sv_pool_gpu_getpoints(mp, &set);
        rc |= sv_pool_gpu_getforce(mp, &set);
        if (rc) {
            fprintf(stderr, "GSvit GPU step failed at step %d\n", set.sc.step_act);
            return 2;
        }
        if (is_output(mp, &set) || i == (set.sc.nsteps - 1)) {
            if (sv_pool_gpu_copyfrom(mp, &set, SV_SYNC_EH))
                return 2;
            sv_pool_output(mp, &set);
        }
        if (killcheck(&set)) {
            if (set.sc.verbose)
                printf("GSvit forced to quit by external file\n");
            break;
        }
        set.sc.step_act++;
    }
    printf("GSvit successfully finished.\n");
    return 0;
}
