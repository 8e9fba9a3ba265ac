/*
 * semsks -- command-line drop-in for `sem2d_psvsh par=run.par`
 * (semsyn/src/sem2d_psvsh.c:82-519; sem2d.c / sem2dMP20140816.c with
 * solver=1 / solver=2).  Reads the same parameter keys and files, runs the
 * time loop on the GPU and writes seis.u / seis.w / seis.v when
 * plot_trace=1.  Exit statuses follow the reference's exit() calls.
 */
#include <stdio.h>
#include <stdlib.h>
#include <string.h>
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

/* The run with plot_snap=1: the engine stepped from snapshot level to
   snapshot level (startsnap + j incresnap, j = 0..25, the levels the run
   reaches, sem2d_psvsh.c:483-510), each written as out<j+1> before the
   rest of the run; the traces as semsks_run_problem gives them. */
static semsks_status run_with_snapshots(const semsks_model *m, const semsks_problem *p, int nx, int nz, float *tr)
{
    int k[3];
    semsks_model_snapshots(m, k);
    long lv[26];
    int idx[26], n = 0;
    for (int j = 0; j < 26; ++j) {
        const long L = (long)k[1] + (long)j * k[2];
        if (L < 0 || L > p->nst - 1) continue;
        int i = n++;
        for (; i > 0 && lv[i - 1] > L; --i) { lv[i] = lv[i - 1]; idx[i] = idx[i - 1]; }
        lv[i] = L;
        idx[i] = j;
    }
    semsks_engine_opts o;
    memset(&o, 0, sizeof o);
    o.max_slots = 1;
    o.use_graph = 1;
    semsks_engine *e = NULL;
    semsks_status st = semsks_engine_create(&o, &e);
    if (st != SEMSKS_OK) return st;
    int slot = -1;
    long level = 0;
    float *field = (float *)malloc(sizeof(float) * 3 * (size_t)nx * (size_t)nz);
    if (!field) {
        fprintf(stderr, "semsks: out of host memory for a %d x %d snapshot\n", nx, nz);
        semsks_engine_destroy(e);
        return SEMSKS_EINVAL;
    }
    st = semsks_admit(e, p, -1, &slot);
    for (int i = 0; i < n && st == SEMSKS_OK; ++i) {
        if (lv[i] > level) {
            st = semsks_advance(e, (int)(lv[i] - level), 1);
            level = lv[i];
        }
        char name[16];
        snprintf(name, sizeof name, "out%d", idx[i] + 1);
        if (st == SEMSKS_OK) st = semsks_get_snapshot(e, slot, field);
        if (st == SEMSKS_OK) st = semsks_write_snapshot(semsks_model_outdir(m), name, nx, nz, field);
    }
    if (st == SEMSKS_OK && level < p->nst - 1) st = semsks_advance(e, (int)(p->nst - 1 - level), 1);
    if (st == SEMSKS_OK) st = semsks_get_traces(e, slot, 0, p->nst, tr);
    free(field);
    semsks_engine_destroy(e);
    return st;
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
    int snap[3];
    semsks_model_snapshots(m, snap);
    st = snap[0] ? run_with_snapshots(m, p, dims[0], dims[1], tr) : semsks_run_problem(p, tr);
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
