/*
 * semsks -- command-line drop-in for `sem2d_psvsh par=run.par`
 * (semsyn/src/sem2d_psvsh.c:82-519; sem2d.c / sem2dMP20140816.c with
 * solver=1 / solver=2).  Reads the same parameter keys and files, runs the
 * time loop on the GPU and writes seis.u / seis.w / seis.v when
 * plot_trace=1.  Exit statuses follow the reference's exit() calls.
 */
#include <stdio.h>
#include <stdlib.h>
#include "semsks.h"

static int exit_code(semsks_status st)
{
    switch (st) {
    case SEMSKS_OK: return 0;
    case SEMSKS_EIO: return 255;      /* exit(-1) */
    case SEMSKS_ENP: return 254;      /* exit(-2) */
    case SEMSKS_ERANGE: return 253;   /* exit(-3) */
    case SEMSKS_EPAR: return 100;
    case SEMSKS_ESTOP: return 101;
    default: return 1;
    }
}

int main(int argc, char **argv)
{
    semsks_model *m = NULL;
    semsks_status st = semsks_load_par(argc, argv, &m);
    if (st != SEMSKS_OK) {
        fprintf(stderr, "semsks: %s\n", semsks_last_error());
        return exit_code(st);
    }
    const semsks_problem *p = semsks_model_problem(m);
    int dims[8];
    semsks_model_dims(m, dims);
    printf("nx=%d nz=%d nex=%d nez=%d eh=%f ntrace=%d nst=%d nt_sem=%d materials=%d\n", dims[0], dims[1], p->nex,
           p->nez, p->eh, p->ntrace, p->nst, dims[4], dims[7]);
    /* check=1 (the default): the setup diagnostics, before the solve as in
       the reference (sem2d_psvsh.c:247-270, discret :804-825) */
    st = semsks_write_check(m);
    if (st != SEMSKS_OK) {
        fprintf(stderr, "semsks: %s\n", semsks_last_error());
        semsks_model_free(m);
        return exit_code(st);
    }
    size_t n = (size_t)p->nst * 3 * (p->ntrace > 0 ? p->ntrace : 1);
    float *tr = (float *)calloc(n, sizeof(float));
    st = semsks_run_problem(p, tr);
    /* an unstable run still writes its traces, NaN/Inf included, as the
       reference does (its fwrite in kirrecord_zhl, sem2d_psvsh.c:2116-2165,
       never checks the values): a later uvw2sac step then converts this
       event's output, not a stale seis.* of an earlier one */
    if ((st == SEMSKS_OK || st == SEMSKS_ENONFINITE) && dims[5]) {
        const semsks_status run_st = st;
        char why[512];
        snprintf(why, sizeof why, "%s", semsks_last_error());
        st = semsks_write_seis(semsks_model_outdir(m), p->nst, p->ntrace, tr);
        if (st == SEMSKS_OK && run_st != SEMSKS_OK) {
            fprintf(stderr, "semsks: %s (seis.* written)\n", why);
            st = run_st;
            free(tr);
            semsks_model_free(m);
            return exit_code(st);
        }
    }
    if (st != SEMSKS_OK) fprintf(stderr, "semsks: %s\n", semsks_last_error());
    else printf("SEM calculation finished!\n");
    free(tr);
    semsks_model_free(m);
    return exit_code(st);
}
