/*
 * semsks_sac.c -- SAC output of the seismograms, the uvw2sac step
 * (semsyn/plot/uvw2sac.c) that follows the SEM solve in the reference's
 * workflow (applications/test/prem_epi83.pl:34).
 *
 * Reads seis.u / seis.v / seis.w (nst records of ntrace float32) and writes,
 * per receiver k, k.ce (u, east), k.cn (v, north) and k.cz (w, vertical) as
 * SAC files, plus pm.xy and log.txt, with the same parameter keys:
 *   decimation istep = (int)(2*sdt/sdt/kdt) (uvw2sac.c:92,207), amplitude
 *   x enlarge (default 4), the last sample dropped when nt is odd
 *   (uvw2sac.c:300-303), header fields of uvw2sac.c:313-347,377-379 and the
 *   component orientations of uvw2sac.c:462-467;
 *   Conv_Source=1: trapezoid source-time function (source.f:137-163) of unit
 *   area convolved in (source.f:61-106, the linear convolution scaled by
 *   dt), 1/sqrt(r) spreading, optional 2-D to point-source step
 *   (uvw2sac.c:474-503 with fa, source.f:166-173) and time derivative
 *   (source.f:175-185).
 * Extension keys: indir (where seis.* are, default "."), outdir (default ".").
 *
 * The SAC header is the standard 632-byte version-6 layout (70 floats,
 * 40 integers, 24 eight-byte strings with kevnm taking two), undefined
 * fields -12345, written in the host's byte order as the reference's
 * write_sac does.
 */
#include <math.h>
#include <stdint.h>
#include <stdio.h>
#include <stdlib.h>
#include <string.h>
#include "semsks.h"
#include "semsks_internal.h"

typedef struct {
    float f[70];
    int32_t i[40];
    char k[192];
} sachdr;

enum {  /* float words */
    H_DELTA = 0, H_B = 5, H_E = 6, H_O = 7, H_STLA = 31, H_STLO = 32, H_STEL = 33, H_EVLA = 35, H_EVLO = 36,
    H_EVDP = 38, H_GCARC = 53, H_CMPAZ = 57, H_CMPINC = 58
};
enum {  /* integer words, counted from the first integer */
    H_NZYEAR = 0, H_NZJDAY = 1, H_NZHOUR = 2, H_NZMIN = 3, H_NZSEC = 4, H_NZMSEC = 5, H_NVHDR = 6, H_NPTS = 9,
    H_IFTYPE = 15, H_LEVEN = 35
};

static void sac_null(sachdr *h)
{
    for (int j = 0; j < 70; ++j) h->f[j] = -12345.f;
    for (int j = 0; j < 40; ++j) h->i[j] = -12345;
    for (int j = 0; j < 24; ++j) {
        if (j == 2) continue;                          /* second half of kevnm */
        memcpy(h->k + 8 * j, j == 1 ? "-12345          " : "-12345  ", j == 1 ? 16 : 8);
    }
}

static void sac_kstnm(sachdr *h, const char *name)
{
    char b[8];
    memset(b, ' ', 8);
    size_t n = strlen(name);
    memcpy(b, name, n < 8 ? n : 8);
    memcpy(h->k, b, 8);
}

static semsks_status write_sac(const char *fn, const sachdr *h, const float *y, int n)
{
    FILE *f = fopen(fn, "wb");
    if (!f) return semsks_set_error(SEMSKS_EIO, "cannot write %s", fn);
    int ok = fwrite(h->f, 4, 70, f) == 70 && fwrite(h->i, 4, 40, f) == 40 && fwrite(h->k, 1, 192, f) == 192 &&
             (n <= 0 || fwrite(y, 4, (size_t)n, f) == (size_t)n);
    fclose(f);
    return ok ? SEMSKS_OK : semsks_set_error(SEMSKS_EIO, "short write on %s", fn);
}

/* stime (source.f:137-163): trapezoid with rise n1, flat n2, fall n3 samples,
   scaled to unit area under the trapezoid rule */
static int stime(float dt1, float dt2, float dt3, float dt, float *f, int cap)
{
    const int n1 = (int)(dt1 / dt + 0.1f), n2 = (int)(dt2 / dt + 0.1f), n3 = (int)(dt3 / dt + 0.1f);
    const int n12 = n1 + n2 + 1, n = n1 + n2 + n3 + 1;
    if (n > cap) return -1;
    float sum = 0.f;
    f[0] = 0.f;
    for (int i = 2; i <= n; ++i) {
        float v = 1.f;
        if (i < n1 + 1) v = (float)(i - 1) * (1.f / (float)n1);
        if (i > n12) v = 1.f - (float)(i - n12) * (1.f / (float)n3);
        f[i - 1] = v;
        sum += (f[i - 1] + f[i - 2]) * 0.5f;
    }
    sum *= dt;
    for (int i = 0; i < n; ++i) f[i] /= sum;
    return n;
}

/* convt (source.f:61-106): z = dt * (x * y), linear convolution (the
   reference's FFT is padded to twice the longer input, so nothing wraps),
   first max(nx, ny) samples */
static int convt(const float *x, int nx, const float *y, int ny, float *z, float dt)
{
    const int nz = nx > ny ? nx : ny;
    double *acc = (double *)calloc((size_t)nz, sizeof(double));
    if (!acc) return -1;
    for (int i = 0; i < nx; ++i)
        for (int j = 0; j < ny && i + j < nz; ++j) acc[i + j] += (double)x[i] * (double)y[j];
    for (int i = 0; i < nz; ++i) z[i] = (float)(acc[i] * dt);
    free(acc);
    return nz;
}

/* sstep + convs (uvw2sac.c:474-503) with fa (source.f:166-173): the 2-D
   line-source trace convolved with 1/sqrt(t) by the reference's Simpson
   sum, every nfad-th output of the first half kept.  Reads the reference
   makes before the start of the trace count as 0. */
static void sstep(int nfa, int nfad, float dp, float *rp, float *pp)
{
    float *f = (float *)malloc(sizeof(float) * (size_t)(nfa > 1 ? nfa : 1));
    if (!f) return;
    for (int j = 2; j <= nfa; ++j) f[j - 1] = 1.f / sqrtf((float)(j - 1) * dp);
    f[0] = (11.f - 4.f * powf(2.f, 0.5f)) / sqrtf(2.f * dp);
    const int nff = nfa / 2;
    for (int j = 0; j < nfa; ++j) pp[j] = rp[j];
    int l = 0;
    for (int j = 0; j < nff; j += nfad) {
        const int nn = j;
        const int ndo = nn > (nfa - 1) / 2 ? (nfa - 1) / 2 : nn;
        int ip = 1, np = 2 * nn - 1;
#define PP(q) ((q) >= 0 && (q) < nfa ? pp[q] : 0.f)
        float even = f[ip < nfa ? ip : 0] * PP(np), odd = 0.f;
        if (ndo > 0) {
            for (int i = 1; i < ndo; ++i) {
                ++ip; --np;
                odd += f[i] * PP(np);
                ++ip; --np;
                even += f[i] * PP(np);
            }
        }
        const float ends = f[0] * PP(2 * nn) + (ip + 1 < nfa ? f[ip + 1] : 0.f) * PP(np - 1);
#undef PP
        rp[l++] = dp * (ends + 4.f * even + 2.f * odd) / 3.f;
    }
    free(f);
}

semsks_status semsks_uvw2sac(int argc, char **argv)
{
    semsks_pars *P = semsks_pars_open(argc, argv);
    if (!P) return semsks_set_error(SEMSKS_EINVAL, "out of host memory");
    int flat = 1, nex = 0, nst = 0, hpn = 0, nfil = 1, kdt = 1, xtrace = 700, ztrace = 1, idzplot = 1, idxplot = 1;
    int get_rec = 0, fix_rec = 0, ite = 0, key = 0, idev = 0;
    float eh = 0, dp = 0, dt = 0, xmin = 0, tstart = 0, zbeg = 0, xbeg = 0, zs = 0, enlarge = 4.f, baz = 270.f;
    float amom = 0, dt1 = 0, dt2 = 0, dt3 = 0, rhos = 0;
    char raymodel[512] = "", rec_file[512] = "", pos_file[512] = "", indir[512] = ".", outdir[512] = ".";
    semsks_par_i(P, "flat", &flat);
    semsks_mstpar_s(P, "raymodel", raymodel, sizeof raymodel);
    semsks_mstpar_i(P, "nex", &nex);
    semsks_mstpar_i(P, "nst", &nst);
    semsks_mstpar_i(P, "hpn", &hpn);
    semsks_mstpar_f(P, "eh", &eh);
    semsks_mstpar_f(P, "dp", &dp);
    semsks_mstpar_f(P, "sdt", &dt);
    semsks_par_i(P, "nfil", &nfil);
    semsks_par_i(P, "kdt", &kdt);
    semsks_par_f(P, "xmin", &xmin);
    semsks_par_f(P, "tstart", &tstart);
    semsks_par_i(P, "xtrace", &xtrace);
    semsks_par_i(P, "ztrace", &ztrace);
    semsks_mstpar_f(P, "zbeg", &zbeg);
    semsks_mstpar_i(P, "idzplot", &idzplot);
    semsks_mstpar_f(P, "xbeg", &xbeg);
    semsks_mstpar_i(P, "idxplot", &idxplot);
    semsks_mstpar_f(P, "zs", &zs);
    semsks_par_i(P, "get_rec", &get_rec);
    semsks_par_f(P, "enlarge", &enlarge);
    if (get_rec == 1) semsks_mstpar_s(P, "rec_file", rec_file, sizeof rec_file);
    semsks_par_i(P, "fix_rec", &fix_rec);
    if (fix_rec == 1) {
        semsks_mstpar_s(P, "pos_file", pos_file, sizeof pos_file);
        semsks_mstpar_f(P, "baz", &baz);
    }
    semsks_mstpar_i(P, "Conv_Source", &ite);
    if (ite == 1) {
        semsks_mstpar_f(P, "Moment", &amom);
        semsks_mstpar_f(P, "Dt1", &dt1);
        semsks_mstpar_f(P, "Dt2", &dt2);
        semsks_mstpar_f(P, "Dt3", &dt3);
        semsks_mstpar_f(P, "beta", &rhos);
        semsks_par_i(P, "point_source", &key);
        semsks_par_i(P, "finaldev", &idev);
    }
    semsks_par_s(P, "indir", indir, sizeof indir);
    semsks_par_s(P, "outdir", outdir, sizeof outdir);
    semsks_status st = semsks_pars_close(P);
    if (st != SEMSKS_OK) return st;
    if (hpn < 1 || kdt < 1 || !(dt > 0)) return semsks_set_error(SEMSKS_EINVAL, "bad hpn/kdt/sdt");

    int nez = 0;
    st = semsks_raymodel_mesh(raymodel, flat, nfil, eh, &nez, &eh);
    if (st != SEMSKS_OK) return st;
    const float h = eh / hpn;
    const int nz = nez * hpn + 1;
    const int ixbeg = (int)(xbeg / h);
    int izbeg = (int)(zbeg / h) + nz + 1;
    if (izbeg > nz - 1) izbeg = nz - 1;

    int *ix1 = NULL, *iz1 = NULL;
    float *longs = NULL, *lats = NULL;
    char (*names)[64] = NULL;
    float *outu = NULL, *outv = NULL, *outw = NULL, *hold = NULL, *us = NULL, *vs = NULL, *ws = NULL, *stf = NULL;
    FILE *fu = NULL, *fv = NULL, *fw = NULL, *fpm = NULL;
    char path[1200];
    if (get_rec == 1) {
        ztrace = 1;
        FILE *r = fopen(rec_file, "r");
        if (!r) return semsks_set_error(SEMSKS_EIO, "Error Opening file %s with access rt", rec_file);
        if (fscanf(r, "%d", &xtrace) != 1 || xtrace < 1) { fclose(r); return semsks_set_error(SEMSKS_EIO, "bad %s", rec_file); }
        ix1 = (int *)calloc((size_t)xtrace, sizeof(int));
        iz1 = (int *)calloc((size_t)xtrace, sizeof(int));
        for (int t = 0; t < xtrace; ++t) {
            float rx = 0;
            if (fscanf(r, "%f", &rx) != 1) break;
            ix1[t] = (int)(rx / h);
            iz1[t] = izbeg;
        }
        fclose(r);
    } else {
        if (xtrace < 1 || ztrace < 1) return semsks_set_error(SEMSKS_EINVAL, "xtrace/ztrace must be >= 1");
        ix1 = (int *)calloc((size_t)xtrace * ztrace, sizeof(int));
        iz1 = (int *)calloc((size_t)xtrace * ztrace, sizeof(int));
        for (int t = 0; t < xtrace; ++t)
            for (int j = 0; j < ztrace; ++j) {
                ix1[t + j * xtrace] = ixbeg + idxplot * t;
                iz1[t + j * xtrace] = izbeg - idzplot * j;
            }
    }
    if (fix_rec == 1) {
        FILE *r = fopen(pos_file, "r");
        if (!r) { st = semsks_set_error(SEMSKS_EIO, "Error Opening file %s with access rt", pos_file); goto done; }
        int ptrace = 0;
        if (fscanf(r, "%d", &ptrace) != 1) ptrace = 0;
        if (ptrace != xtrace) { fclose(r); st = semsks_set_error(SEMSKS_EINVAL, "inconsistent position file and receiver file"); goto done; }
        longs = (float *)calloc((size_t)ptrace, sizeof(float));
        lats = (float *)calloc((size_t)ptrace, sizeof(float));
        names = calloc((size_t)ptrace, sizeof *names);
        for (int j = 0; j < ptrace; ++j)
            if (fscanf(r, "%f %f %63s", &longs[j], &lats[j], names[j]) != 3) break;
        fclose(r);
    }

    const int istep0 = (int)(dt * 2 / dt / kdt);
    const int istep = istep0 < 1 ? 1 : istep0;
    const int nt = nst / kdt;
    const int ntrace = xtrace * ztrace;
    const float dis2deg = 111.2f;
    {
        const float t_e = tstart + nt * dt * kdt;
        float x_b, x_e;
        if (get_rec == 1) {
            x_b = (xmin + ix1[0] * h) / dis2deg;
            x_e = (xmin + ix1[xtrace - 1] * h) / dis2deg;
        } else {
            x_b = (xmin + xbeg) / dis2deg;
            x_e = (xmin + xbeg + idxplot * (xtrace - 1) * h) / dis2deg;
        }
        snprintf(path, sizeof path, "%s/log.txt", outdir);
        FILE *lf = fopen(path, "w");
        if (!lf) { st = semsks_set_error(SEMSKS_EIO, "cannot write %s", path); goto done; }
        fprintf(lf, "%-9.3f %-9.3f\n", tstart, t_e);
        fprintf(lf, "%-9.3f %-9.3f\n", x_b, x_e);
        fclose(lf);
    }
    snprintf(path, sizeof path, "%s/seis.u", indir); fu = fopen(path, "rb");
    snprintf(path, sizeof path, "%s/seis.v", indir); fv = fopen(path, "rb");
    snprintf(path, sizeof path, "%s/seis.w", indir); fw = fopen(path, "rb");
    if (!fu || !fv || !fw) { st = semsks_set_error(SEMSKS_EIO, "cannot open seis.u/.v/.w in %s", indir); goto done; }
    const int len = nt / istep + 1;
    outu = (float *)calloc((size_t)len * ntrace, sizeof(float));
    outv = (float *)calloc((size_t)len * ntrace, sizeof(float));
    outw = (float *)calloc((size_t)len * ntrace, sizeof(float));
    hold = (float *)calloc((size_t)3 * ntrace, sizeof(float));
    if (!outu || !outv || !outw || !hold) { st = semsks_set_error(SEMSKS_EINVAL, "out of host memory"); goto done; }
    int num = 0;
    for (int it = 0; it < nt; ++it) {
        /* a short file leaves the record as the previous one, like fread */
        if (fread(hold, 4, (size_t)ntrace, fu) != (size_t)ntrace ||
            fread(hold + ntrace, 4, (size_t)ntrace, fv) != (size_t)ntrace ||
            fread(hold + 2 * ntrace, 4, (size_t)ntrace, fw) != (size_t)ntrace) { /* keep going */ }
        if (it % istep) continue;
        memcpy(outu + (size_t)num * ntrace, hold, 4 * (size_t)ntrace);
        memcpy(outv + (size_t)num * ntrace, hold + ntrace, 4 * (size_t)ntrace);
        memcpy(outw + (size_t)num * ntrace, hold + 2 * ntrace, 4 * (size_t)ntrace);
        ++num;
    }
    if (num > nt / istep) num--;

    sachdr H;
    sac_null(&H);
    H.f[H_DELTA] = istep * dt * kdt;
    H.i[H_NZYEAR] = 2006; H.i[H_NZJDAY] = 0; H.i[H_NZHOUR] = 0; H.i[H_NZMIN] = 0; H.i[H_NZSEC] = 0; H.i[H_NZMSEC] = 0;
    H.f[H_EVLA] = 0.f;
    H.f[H_EVLO] = 0.f;
    if (fix_rec == 1) {
        const float g0 = get_rec == 1 ? (xmin + ix1[0] * h) / dis2deg : (xmin + xbeg + idxplot * h) / dis2deg;
        if (baz == 270) H.f[H_EVLO] = longs[0] - g0;
        if (baz == 90) H.f[H_EVLO] = longs[0] + g0;
    }
    H.f[H_EVDP] = zs;
    H.f[H_STLA] = 0.f;
    H.f[H_STEL] = 0.f;
    H.f[H_O] = -tstart;
    H.f[H_B] = 0.f;
    H.f[H_E] = num * H.f[H_DELTA];
    H.i[H_NPTS] = num;
    H.i[H_IFTYPE] = 1;
    H.i[H_LEVEN] = 1;
    H.i[H_NVHDR] = 6;
    const int cap = num > 0 ? 2 * num + 20000 : 20000;
    us = (float *)calloc((size_t)cap, sizeof(float));
    vs = (float *)calloc((size_t)cap, sizeof(float));
    ws = (float *)calloc((size_t)cap, sizeof(float));
    stf = (float *)calloc((size_t)cap, sizeof(float));
    if (!us || !vs || !ws || !stf) { st = semsks_set_error(SEMSKS_EINVAL, "out of host memory"); goto done; }
    amom = (float)((amom * (1.e-20)) / (4. * 3.1415927 * rhos));
    snprintf(path, sizeof path, "%s/pm.xy", outdir);
    fpm = fopen(path, "w");
    if (!fpm) { st = semsks_set_error(SEMSKS_EIO, "cannot write %s", path); goto done; }
    const float ndt = H.f[H_DELTA];
    const float ne_pole = (fix_rec == 1 && baz == 90) ? -1.f : 1.f;
    for (int t = 0; t < xtrace && st == SEMSKS_OK; ++t) {
        const float gcarc = get_rec == 1 ? (xmin + ix1[t] * h) / dis2deg : (xmin + xbeg + idxplot * t * h) / dis2deg;
        for (int j = 0; j < ztrace && st == SEMSKS_OK; ++j) {
            char name[64];
            if (fix_rec == 0) H.f[H_GCARC] = gcarc;
            H.f[H_STLO] = gcarc;
            if (fix_rec == 1) {
                H.f[H_STLO] = longs[t];
                snprintf(name, sizeof name, "%s", names[t]);
            } else {
                snprintf(name, sizeof name, "%d", t);
            }
            sac_kstnm(&H, name);
            fprintf(fpm, ">> %-8.2f %-8.2f %-8.2f %d\n", gcarc, H.f[H_B], H.f[H_DELTA], num);
            for (int k = 0; k < num; ++k) {
                const size_t off = (size_t)k * ntrace + j * xtrace + t;
                us[k] = outu[off] * enlarge;
                vs[k] = outv[off] * enlarge;
                ws[k] = outw[off] * enlarge;
                fprintf(fpm, "%-11.4e %-11.4e\n", us[k], ws[k]);
            }
            int npts = num;
            if (ite == 1) {
                float sdt = ndt;
                const float xr = gcarc * dis2deg, zr = 0.f, xs = 0.f;
                float *comp[3] = {us, vs, ws};
                for (int c = 0; c < 3; ++c) {
                    float *b = comp[c];
                    int nnn = num, found = 0;
                    for (int k = num - 1; k >= 0; --k) {
                        if (fabsf(b[k]) <= 1.e-30f && !found) nnn = k;
                        else found = 1;
                    }
                    if (key == 1) {
                        sstep(nnn, 1, sdt, b, stf);
                        nnn = nnn / 2;
                        sdt = ndt * 2;
                        H.f[H_DELTA] = sdt;
                    } else {
                        amom = 1.f;
                    }
                    const int nfa = stime(dt1, dt2, dt3, sdt, stf, cap);
                    if (nfa < 0) { st = semsks_set_error(SEMSKS_EINVAL, "source time function longer than %d samples", cap); break; }
                    int nnum = nnn;
                    if (key == 1 && idev == 1 && nnum > 0) {   /* diff (source.f:175-185) */
                        float y1 = b[0];
                        b[0] = 0.f;
                        for (int q = 1; q < nnum; ++q) { const float y2 = b[q]; b[q] = (b[q] - y1) / sdt; y1 = y2; }
                    }
                    const float roo = sqrtf((xr - xs) * (xr - xs) + (zr - zs) * (zr - zs));
                    const float cc3 = 1.4142f / sqrtf(roo);
                    for (int q = 0; q < nnum; ++q) b[q] *= cc3;
                    for (int q = 0; q < nnum; ++q) b[q] = b[q] * amom;
                    nnum = convt(b, nnum, stf, nfa, b, sdt);
                    if (nnum < 0) { st = semsks_set_error(SEMSKS_EINVAL, "out of host memory"); break; }
                    if (c < 2) for (int q = 0; q < nnum; ++q) b[q] *= ne_pole;
                    npts = nnum;
                }
                if (st != SEMSKS_OK) break;
                H.i[H_NPTS] = npts;
            }
            struct { const char *ext; float az, inc; const float *y; } out[3] = {
                {".ce", 90.f, 90.f, us}, {".cn", 0.f, 90.f, vs}, {".cz", 0.f, 0.f, ws}};
            for (int c = 0; c < 3 && st == SEMSKS_OK; ++c) {
                H.f[H_CMPAZ] = out[c].az;
                H.f[H_CMPINC] = out[c].inc;
                snprintf(path, sizeof path, "%s/%s%s", outdir, name, out[c].ext);
                st = write_sac(path, &H, out[c].y, H.i[H_NPTS]);
            }
        }
    }
done:
    if (fu) fclose(fu);
    if (fv) fclose(fv);
    if (fw) fclose(fw);
    if (fpm) fclose(fpm);
    free(ix1); free(iz1); free(longs); free(lats); free(names);
    free(outu); free(outv); free(outw); free(hold); free(us); free(vs); free(ws); free(stf);
    return st;
}
