/*
 * semsks_io.c -- host-side inputs of the SEM solve in the reference's formats.
 *
 *   parameters : getpar semantics (semsyn/src/getpar/src/getpar.c:137-260,
 *                439-595, 625-712): name=value arguments, par=file included
 *                recursively (depth <= 8), '#' comments, quotes, $(name)
 *                substitution, backward lookup so the last definition wins,
 *                mstpar for mandatory keys (status 100), STOP (status 101).
 *   raymodel   : readmodel.c:43-109 (+ flatten, readmodel.c:150-165)
 *   semmodel   : sem2d_psvsh.c:700-731
 *   mesh sizes : sem2d_psvsh.c:591-611 (nez and eh from the GRT thickness)
 *   material   : sem2d_psvsh.c:618-797 (1-D prefill, then semmodel layers
 *                per node column with lip(), ratio layers compounding)
 *   receivers  : sem2d_psvsh.c:330-370
 *   GRT files  : greenfile/shgreenfile, time-major float32 records
 *                (demult.c:63-77)
 *
 * Material is stored de-duplicated: one record {rho, 15 constants} per
 * distinct value set and a 16-bit record id per node.  The per-node values
 * are the reference's float values bit for bit.
 */
#include <stdio.h>
#include <stdlib.h>
#include <string.h>
#include <stdarg.h>
#include <math.h>
#include "semsks.h"
#include "semsks_internal.h"

/* ---------------------------------------------------------------------- */
/* error text                                                              */
/* ---------------------------------------------------------------------- */
static __thread char g_err[1024];

const char *semsks_last_error(void) { return g_err; }

semsks_status semsks_set_error(semsks_status st, const char *fmt, ...)
{
    va_list ap;
    va_start(ap, fmt);
    vsnprintf(g_err, sizeof g_err, fmt, ap);
    va_end(ap);
    return st;
}

/* ---------------------------------------------------------------------- */
/* getpar                                                                  */
/* ---------------------------------------------------------------------- */
typedef struct { char **name, **val; int n, cap; semsks_status st; } parlist;

static void pl_add(parlist *L, const char *name, const char *val)
{
    if (L->n == L->cap) {
        L->cap = L->cap ? L->cap * 2 : 64;
        L->name = (char **)realloc(L->name, L->cap * sizeof(char *));
        L->val = (char **)realloc(L->val, L->cap * sizeof(char *));
    }
    L->name[L->n] = strdup(name);
    L->val[L->n] = strdup(val);
    L->n++;
}

static const char *pl_get(const parlist *L, const char *name)
{
    for (int i = L->n - 1; i >= 0; --i)
        if (!strcmp(L->name[i], name)) return L->val[i];
    return NULL;
}

/* scans one value starting at s; quoted values may hold blanks, an escaped
   quote stays literal; $(key) expands to the latest definition of key */
static const char *scan_value(const parlist *L, const char *s, char *out, size_t cap, int blank_ends)
{
    size_t o = 0;
    char q = 0;
    if (*s == '"' || *s == '\'') q = *s++;
    for (; *s && *s != '\n'; ++s) {
        if (q && *s == q) {
            if (o && out[o - 1] == '\\') { out[o - 1] = q; continue; }
            ++s;
            break;
        }
        if (!q && blank_ends && (*s == ' ' || *s == '\t')) break;
        if (s[0] == '$' && s[1] == '(') {
            const char *e = strchr(s + 2, ')');
            if (e) {
                char key[256];
                size_t kl = (size_t)(e - s - 2) < sizeof key ? (size_t)(e - s - 2) : sizeof key - 1;
                memcpy(key, s + 2, kl);
                key[kl] = 0;
                const char *v = pl_get(L, key);
                while (v && *v && o + 1 < cap) out[o++] = *v++;
                s = e;
                continue;
            }
        }
        if (o + 1 < cap) out[o++] = *s;
    }
    out[o] = 0;
    return s;
}

static void pl_file(parlist *L, const char *fn, int depth)
{
    if (depth > 8) { L->st = semsks_set_error(SEMSKS_EPAR, "%d (too many) recursive par files", depth); return; }
    if (!*fn) return;
    FILE *f = fopen(fn, "r");
    if (!f) { L->st = semsks_set_error(SEMSKS_EPAR, "cannot open par file %s", fn); return; }
    char line[16384];
    size_t used = 0;
    while (fgets(line + used, (int)(sizeof line - used), f)) {
        size_t len = strlen(line);
        if (len >= 2 && line[len - 1] == '\n' && line[len - 2] == '\\') { used = len - 2; continue; }
        used = 0;
        const char *s = line;
        for (;;) {
            char name[256], val[8192];
            size_t nl = 0;
            while (*s == ' ' || *s == '\t') ++s;
            if (!*s || *s == '\n' || *s == '#') break;
            while (*s && *s != '=' && *s != ' ' && *s != '\t' && *s != '\n') {
                if (nl + 1 < sizeof name) name[nl++] = *s;
                ++s;
            }
            name[nl] = 0;
            if (*s == '=') ++s;
            s = scan_value(L, s, val, sizeof val, 1);
            pl_add(L, name, val);
            if (!strcmp(name, "par")) pl_file(L, val, depth + 1);
        }
    }
    fclose(f);
}

static void pl_free(parlist *L)
{
    for (int i = 0; i < L->n; ++i) { free(L->name[i]); free(L->val[i]); }
    free(L->name);
    free(L->val);
}

typedef struct { parlist L; const char *missing; } pars;

static int get_i(pars *P, const char *k, int *v) { const char *s = pl_get(&P->L, k); if (!s) return 0; *v = atoi(s); return 1; }
static int get_f(pars *P, const char *k, float *v) { const char *s = pl_get(&P->L, k); if (!s) return 0; *v = (float)atof(s); return 1; }
static int get_s(pars *P, const char *k, char *v, size_t cap) { const char *s = pl_get(&P->L, k); if (!s) return 0; snprintf(v, cap, "%s", s); return 1; }
static void req_i(pars *P, const char *k, int *v) { if (!get_i(P, k, v) && !P->missing) P->missing = k; }
static void req_f(pars *P, const char *k, float *v) { if (!get_f(P, k, v) && !P->missing) P->missing = k; }
static void req_s(pars *P, const char *k, char *v, size_t cap) { if (!get_s(P, k, v, cap) && !P->missing) P->missing = k; }

/* ---------------------------------------------------------------------- */
/* material de-duplication                                                 */
/* ---------------------------------------------------------------------- */
typedef struct {
    float *rec;            /* [n][16] */
    int n, cap;
    int *hash;             /* open addressing, -1 empty */
    int hcap;
} mattab;

static uint64_t rec_hash(const float *r)
{
    uint64_t h = 1469598103934665603ull;
    const unsigned char *b = (const unsigned char *)r;
    for (int i = 0; i < SEMSKS_MAT_STRIDE * 4; ++i) { h ^= b[i]; h *= 1099511628211ull; }
    return h;
}

static void mt_rehash(mattab *T)
{
    free(T->hash);
    T->hcap = T->hcap ? T->hcap * 2 : 1024;
    T->hash = (int *)malloc(T->hcap * sizeof(int));
    for (int i = 0; i < T->hcap; ++i) T->hash[i] = -1;
    for (int k = 0; k < T->n; ++k) {
        size_t h = rec_hash(T->rec + (size_t)k * SEMSKS_MAT_STRIDE) & (T->hcap - 1);
        while (T->hash[h] >= 0) h = (h + 1) & (T->hcap - 1);
        T->hash[h] = k;
    }
}

static int mt_id(mattab *T, const float *r)
{
    if (2 * (T->n + 1) > T->hcap) mt_rehash(T);
    size_t h = rec_hash(r) & (T->hcap - 1);
    while (T->hash[h] >= 0) {
        if (!memcmp(T->rec + (size_t)T->hash[h] * SEMSKS_MAT_STRIDE, r, SEMSKS_MAT_STRIDE * 4)) return T->hash[h];
        h = (h + 1) & (T->hcap - 1);
    }
    if (T->n >= 65535) return -1;
    if (T->n == T->cap) {
        T->cap = T->cap ? T->cap * 2 : 64;
        T->rec = (float *)realloc(T->rec, (size_t)T->cap * SEMSKS_MAT_STRIDE * sizeof(float));
    }
    memcpy(T->rec + (size_t)T->n * SEMSKS_MAT_STRIDE, r, SEMSKS_MAT_STRIDE * 4);
    T->hash[h] = T->n;
    return T->n++;
}

/* monoclinic (+ tilted-axis) constants in Voigt naming, as cmatrix3d */
typedef struct {
    float c11, c12, c13, c16, c22, c23, c26, c33, c36, c44, c45, c55, c66, rho;
    float c14, c15, c34, c35, c46, c56;
} consts;

/* the record the time loop uses: rho and the 5x5 block of the Voigt matrix
   over indices {xx, zz, yz, xz, xy} -- exactly the c[i][j][k][l] entries
   Stress_ij reads (sem2d_psvsh.c:1363-1376) after Cijkl_3D */
static void to_record(const consts *a, float *r)
{
    r[0] = a->rho;
    r[1] = a->c11; r[2] = a->c13; r[3] = a->c14; r[4] = a->c15; r[5] = a->c16;
    r[6] = a->c33; r[7] = a->c34; r[8] = a->c35; r[9] = a->c36;
    r[10] = a->c44; r[11] = a->c45; r[12] = a->c46;
    r[13] = a->c55; r[14] = a->c56;
    r[15] = a->c66;
}

/* ---------------------------------------------------------------------- */
/* the model                                                               */
/* ---------------------------------------------------------------------- */
struct semsks_model {
    semsks_problem prob;
    int nx, nz, nt_sem, plot_trace;
    int check, output_vd;           /* check=1 (the default): the setup diagnostics files */
    int plot_snap, startsnap, incresnap;
    float xmin;                     /* epicentral distance of the mesh's x = 0, km */
    char outdir[1024];
    float *mat_table;
    uint16_t *mat_id;
    int *rec_ix, *rec_iz;
    float *psv, *sh;
};

void semsks_model_free(semsks_model *m)
{
    if (!m) return;
    free(m->mat_table); free(m->mat_id); free(m->rec_ix); free(m->rec_iz);
    free(m->psv); free(m->sh);
    free(m);
}

const semsks_problem *semsks_model_problem(const semsks_model *m) { return &m->prob; }
const char *semsks_model_outdir(const semsks_model *m) { return m->outdir; }

void semsks_model_dims(const semsks_model *m, int d[8])
{
    d[0] = m->nx; d[1] = m->nz; d[2] = 2 * (m->nx + m->nz); d[3] = m->nx + m->nz;
    d[4] = m->nt_sem; d[5] = m->plot_trace; d[6] = m->prob.ntrace; d[7] = m->prob.n_mat;
}

typedef struct {
    int np, ratio;
    consts a;
    float *an, *bn;
} semlayer;

/* lip (sem2d_psvsh.c:831-846) */
static int lip_height(const semlayer *L, float x, float *dep)
{
    if (x < L->an[0] || x > L->an[L->np - 1]) return -1;
    float d = 0.f;
    for (int i = 0; i < L->np - 1; ++i)
        if (x >= L->an[i] && x < L->an[i + 1]) {
            d = (L->bn[i + 1] - L->bn[i]) / (L->an[i + 1] - L->an[i]);
            d = L->bn[i] + d * (x - L->an[i]);
        }
    *dep = d;
    return 0;
}

static semsks_status read_raymodel(const char *fn, int flat, int nfil, float eh_in, int *nez, float *eh,
                                   int *nrec_out, int *jo_out, float **cc, float **ss, float **dd, float **tth,
                                   float *thickness)
{
    FILE *f = fopen(fn, "r");
    if (!f) return semsks_set_error(SEMSKS_EIO, "Error Opening file %s with access r", fn);
    int jo = 0, nb = 0, lfinal = 0;
    if (fscanf(f, "%d %d\n", &jo, &nb) != 2 || jo <= 0) { fclose(f); return semsks_set_error(SEMSKS_EIO, "bad raymodel %s", fn); }
    float *c = calloc(jo + 1, 4), *s = calloc(jo + 1, 4), *d = calloc(jo + 1, 4), *t = calloc(jo + 1, 4);
    for (int j = 0; j < jo; ++j)
        if (fscanf(f, "%f %f %f %f\n", &c[j], &s[j], &d[j], &t[j]) != 4) { fclose(f); free(c); free(s); free(d); free(t); return semsks_set_error(SEMSKS_EIO, "bad raymodel layer %d in %s", j, fn); }
    if (flat == 0) {
        /* earth flattening: all factors from the unscaled thicknesses,
           then scale (readmodel.c:69-84, flatten 150-165) */
        double *q = calloc(jo + 1, sizeof(double));
        for (int j = 0; j < jo; ++j) {
            double th = 0.0;
            for (int i = 0; i < j; ++i) th += t[i];
            th += 0.5 * t[j];
            q[j] = 6371.0 / (6371.0 - th);
        }
        for (int j = 0; j < jo; ++j) { c[j] *= q[j]; s[j] *= q[j]; d[j] *= q[j]; t[j] *= q[j]; }
        free(q);
    }
    int nen0 = 0, na1 = -1;
    if (fscanf(f, "%d\n", &lfinal) == 1 && lfinal > 0 && fscanf(f, "%d", &nen0) == 1) {
        for (int j = 0; j < nen0; ++j) {
            int v;
            if (fscanf(f, " %d", &v) != 1) break;
            if (j == nen0 - 2) na1 = v;
        }
    }
    fclose(f);
    if (na1 <= 0 || na1 > jo + 1) { free(c); free(s); free(d); free(t); return semsks_set_error(SEMSKS_EIO, "raymodel %s: cannot find nrec in the ray list", fn); }
    int nrec = na1;
    float thk = 0.0f;
    if (nfil > nrec) { for (int i = nrec; i < nfil; ++i) thk += t[i]; }
    else { for (int i = nfil - 1; i < nrec - 1; ++i) thk += t[i]; }
    *nez = (int)((thk - 1.e-3) / eh_in);
    if (*nez <= 0) { free(c); free(s); free(d); free(t); return semsks_set_error(SEMSKS_EINVAL, "SEM thickness %g gives nez=%d", thk, *nez); }
    *eh = (float)((thk - 1.e-3) / *nez);
    *nrec_out = nrec; *jo_out = jo; *cc = c; *ss = s; *dd = d; *tth = t; *thickness = thk;
    return SEMSKS_OK;
}

static semsks_status assign_material(semsks_model *M, const char *semmodel, const char *semdip, int readpars,
                                     int nfil, int nrec, int jo, const float *cc, const float *ss, const float *dd,
                                     const float *tth)
{
    const int nx = M->nx, nz = M->nz;
    const float h = M->prob.eh / M->prob.hpn;
    mattab T = {0};
    uint16_t *id = (uint16_t *)calloc((size_t)nx * nz, sizeof(uint16_t));
    semsks_status st = SEMSKS_OK;
    float rec[SEMSKS_MAT_STRIDE];
    M->mat_id = id;
    {   /* a zero record for nodes the model never reaches */
        memset(rec, 0, sizeof rec);
        mt_id(&T, rec);
    }
    /* 1-D prefill, bottom up (sem2d_psvsh.c:618-674) */
    int nlay1 = (nfil > nrec) ? nrec : nrec - 2;
    int nlay2 = (nfil > nrec) ? nfil + 200 : nfil - 200;
    if (nlay2 < 0) nlay2 = 0;
    int nlaynum = abs(nlay2 - nlay1) + 1, inlay = (nfil > nrec) ? 1 : -1, iz1 = 0;
    float thk2 = 0.0f;
    for (int k = 0; k < nlaynum; ++k) {
        int i = nlay1 + k * inlay;
        if (i < 0 || i >= jo) break;
        thk2 += tth[i];
        int iz2 = (int)((thk2 + 1.e-3) / h) + 1;
        if (iz2 > nz - 1) iz2 = nz - 1;
        if (iz2 < iz1) break;
        float ivp = cc[i] * cc[i], ivs = ss[i] * ss[i], iden = dd[i];
        float t11 = ivp * iden, t44 = ivs * iden, t13 = (float)(t11 - 2.0 * t44);
        consts a;
        memset(&a, 0, sizeof a);
        a.c11 = t11; a.c12 = t13; a.c13 = t13; a.c22 = t11; a.c23 = t13; a.c33 = t11;
        a.c44 = t44; a.c55 = t44; a.c66 = t44; a.rho = iden;
        to_record(&a, rec);
        int r = mt_id(&T, rec);
        for (int j = iz1; j <= iz2; ++j)
            for (int x = 0; x < nx; ++x) id[(size_t)j * nx + x] = (uint16_t)r;
        iz1 = iz2 + 1;
        if (iz2 == nz - 1) break;
    }
    if (readpars == 1) {
        FILE *f = fopen(semmodel, "r");
        if (!f) { st = semsks_set_error(SEMSKS_EIO, "Error Opening file %s with access r", semmodel); goto done; }
        int nll = 0, first = 0;
        if (fscanf(f, "%d\n", &nll) != 1) { fclose(f); st = semsks_set_error(SEMSKS_EIO, "bad semmodel %s", semmodel); goto done; }
        if (nll < 0) { first = 1; nll = -nll; }
        semlayer *L = (semlayer *)calloc(nll + 1, sizeof(semlayer));
        for (int j = 0; j < nll && st == SEMSKS_OK; ++j) {
            consts *a = &L[j].a;
            int np = 0;
            if (fscanf(f, "%d %f %f %f %f %f %f %f", &np, &a->c11, &a->c12, &a->c13, &a->c16, &a->c22, &a->c23, &a->c26) != 8 ||
                fscanf(f, "%f %f %f %f %f %f %f\n", &a->c33, &a->c36, &a->c44, &a->c45, &a->c55, &a->c66, &a->rho) != 7) {
                st = semsks_set_error(SEMSKS_EIO, "bad semmodel layer %d in %s", j, semmodel);
                break;
            }
            if (np < 0) { L[j].ratio = 1; np = -np; }
            if (np > 1000) { st = semsks_set_error(SEMSKS_ENP, "an bn allocation error, np[j] > 1000"); break; }
            if (np < 1) { st = semsks_set_error(SEMSKS_EIO, "semmodel layer %d has no control points", j); break; }
            L[j].np = np;
            L[j].an = (float *)calloc(np, sizeof(float));
            L[j].bn = (float *)calloc(np, sizeof(float));
            for (int i = 0; i < np; ++i) if (fscanf(f, "%f ", &L[j].an[i]) != 1) break;
            if (fscanf(f, "\n") < 0) {}
            for (int i = 0; i < np; ++i) if (fscanf(f, "%f ", &L[j].bn[i]) != 1) break;
            if (fscanf(f, "\n") < 0) {}
        }
        fclose(f);
        if (st == SEMSKS_OK && semdip && *semdip) {
            FILE *g = fopen(semdip, "r");
            if (!g) st = semsks_set_error(SEMSKS_EIO, "Error Opening file %s with access r", semdip);
            for (int j = 0; g && j < nll; ++j) {
                consts *a = &L[j].a;
                if (fscanf(g, "%f %f %f %f %f %f", &a->c14, &a->c15, &a->c34, &a->c35, &a->c46, &a->c56) != 6) {
                    st = semsks_set_error(SEMSKS_EIO, "semdip %s: need 6 constants for layer %d", semdip, j);
                    break;
                }
            }
            if (g) fclose(g);
        }
        /* per column, bottom up (sem2d_psvsh.c:733-797) */
        consts ap;
        memset(&ap, 0, sizeof ap);
        for (int i = 0; i < nx && st == SEMSKS_OK; ++i) {
            float x = i * h, dep = 0.f;
            int z1 = 0;
            if (first == 1) {
                if (lip_height(&L[0], x, &dep)) { st = semsks_set_error(SEMSKS_ERANGE, "Model out of range, x=%9.3f layer 0", x); break; }
                z1 = (int)((dep + 1.e-3) / h);
            }
            if (z1 < 0) z1 = 0;
            for (int j = first; j < nll; ++j) {
                if (lip_height(&L[j], x, &dep)) { st = semsks_set_error(SEMSKS_ERANGE, "Model out of range, x=%9.3f layer %d", x, j); break; }
                int z2 = (int)((dep + 1.e-3) / h);
                if (z2 > nz - 1) z2 = nz - 1;
                if (L[j].ratio) {
                    const consts *m = &L[j].a;
                    ap.c11 *= m->c11; ap.c12 *= m->c12; ap.c13 *= m->c13; ap.c16 *= m->c16;
                    ap.c22 *= m->c22; ap.c23 *= m->c23; ap.c26 *= m->c26; ap.c33 *= m->c33; ap.c36 *= m->c36;
                    ap.c44 *= m->c44; ap.c45 *= m->c45; ap.c55 *= m->c55; ap.c66 *= m->c66; ap.rho *= m->rho;
                } else {
                    ap = L[j].a;
                }
                if (z1 <= z2) {
                    to_record(&ap, rec);
                    int r = mt_id(&T, rec);
                    if (r < 0) { st = semsks_set_error(SEMSKS_EINVAL, "more than 65535 distinct material records"); break; }
                    for (int z = z1; z <= z2; ++z) id[(size_t)z * nx + i] = (uint16_t)r;
                }
                z1 = z2 + 1;
            }
        }
        for (int j = 0; j < nll; ++j) { free(L[j].an); free(L[j].bn); }
        free(L);
    }
done:
    M->mat_table = T.rec;
    M->prob.n_mat = T.n;
    free(T.hash);
    return st;
}

static semsks_status read_records(const char *fn, int nt, int width, float **out, int *nread)
{
    FILE *f = fopen(fn, "rb");
    if (!f) return semsks_set_error(SEMSKS_EIO, "Error Opening file %s with access rb", fn);
    float *r = (float *)calloc((size_t)nt * width + 1, sizeof(float));
    if (!r) { fclose(f); return semsks_set_error(SEMSKS_ENOHBM, "host memory for %s", fn); }
    size_t got = fread(r, sizeof(float) * (size_t)width, (size_t)nt, f);
    fclose(f);
    *out = r;
    *nread = (int)got;
    return SEMSKS_OK;
}

/* setpar (getpar.c:137-260): name=value arguments, par= files */
static void pars_args(pars *P, int argc, char **argv)
{
    for (int i = 1; i < argc; ++i) {
        const char *a = argv[i];
        char name[256], val[8192];
        size_t nl = 0;
        while (*a == ' ' || *a == '\t') ++a;
        while (*a && *a != '=') { if (nl + 1 < sizeof name) name[nl++] = *a; ++a; }
        name[nl] = 0;
        if (*a == '=') ++a;
        scan_value(&P->L, a, val, sizeof val, 0);
        if (name[0] == '-') pl_add(&P->L, "SWITCH", name + 1);
        else pl_add(&P->L, name, val);
        if (!strcmp(name, "par")) pl_file(&P->L, val, 1);
        if (!strcmp(name, "ENDPAR")) break;
    }
}

/* the same parameter machinery for the library's other readers */
struct semsks_pars { pars P; };

semsks_pars *semsks_pars_open(int argc, char **argv)
{
    semsks_pars *q = (semsks_pars *)calloc(1, sizeof *q);
    if (q) pars_args(&q->P, argc, argv);
    return q;
}
int semsks_par_i(semsks_pars *q, const char *k, int *v) { return get_i(&q->P, k, v); }
int semsks_par_f(semsks_pars *q, const char *k, float *v) { return get_f(&q->P, k, v); }
int semsks_par_s(semsks_pars *q, const char *k, char *v, size_t cap) { return get_s(&q->P, k, v, cap); }
void semsks_mstpar_i(semsks_pars *q, const char *k, int *v) { req_i(&q->P, k, v); }
void semsks_mstpar_f(semsks_pars *q, const char *k, float *v) { req_f(&q->P, k, v); }
void semsks_mstpar_s(semsks_pars *q, const char *k, char *v, size_t cap) { req_s(&q->P, k, v, cap); }

semsks_status semsks_pars_close(semsks_pars *q)
{
    semsks_status st = SEMSKS_OK;
    if (q->P.L.st) st = q->P.L.st;
    else if (q->P.missing) st = semsks_set_error(SEMSKS_EPAR, "must specify value for '%s'", q->P.missing);
    else if (pl_get(&q->P.L, "STOP")) st = semsks_set_error(SEMSKS_ESTOP, "stop due to STOP in input");
    pl_free(&q->P.L);
    free(q);
    return st;
}

semsks_status semsks_raymodel_mesh(const char *fn, int flat, int nfil, float eh_in, int *nez, float *eh)
{
    int nrec = 0, jo = 0;
    float *cc = NULL, *ss = NULL, *dd = NULL, *tth = NULL, thk = 0;
    semsks_status st = read_raymodel(fn, flat, nfil, eh_in, nez, eh, &nrec, &jo, &cc, &ss, &dd, &tth, &thk);
    free(cc); free(ss); free(dd); free(tth);
    return st;
}

semsks_status semsks_load_par(int argc, char **argv, semsks_model **out)
{
    pars P;
    memset(&P, 0, sizeof P);
    *out = NULL;
    pars_args(&P, argc, argv);
    semsks_model *M = (semsks_model *)calloc(1, sizeof *M);
    semsks_problem *p = &M->prob;
    int flat = 1, nfil = 1, readpars = 1, get_rec = 0, xtrace = 20, idxplot = 20, msource = 0, nex = 0;
    int hpn = 5, neh = 1, absorb = 1, nt = 0, nst = 0, solver = 0, kahan = 1;
    float dp = 0, eh = 0, xmin = 0, sdt = 0, xbeg = 0, zbeg = 0;
    char greenfile[1024] = "", shgreenfile[1024] = "", raymodel[1024] = "", semmodel[1024] = "";
    char rec_file[1024] = "", semdip[1024] = "";
    semsks_status st = P.L.st;

    get_i(&P, "flat", &flat);
    get_i(&P, "solver", &solver);
    req_s(&P, "greenfile", greenfile, sizeof greenfile);
    if (solver == SEMSKS_SOLVER_PSVSH) req_s(&P, "shgreenfile", shgreenfile, sizeof shgreenfile);
    req_s(&P, "raymodel", raymodel, sizeof raymodel);
    req_s(&P, "semmodel", semmodel, sizeof semmodel);
    req_i(&P, "nt", &nt);
    req_f(&P, "dp", &dp);
    req_i(&P, "nex", &nex);
    req_f(&P, "eh", &eh);
    req_f(&P, "xmin", &xmin);
    req_i(&P, "nst", &nst);
    req_f(&P, "sdt", &sdt);
    get_i(&P, "nfil", &nfil);
    get_i(&P, "hpn", &hpn);
    get_i(&P, "absorb", &absorb);
    get_i(&P, "readpars", &readpars);
    get_i(&P, "neh", &neh);
    get_i(&P, "xtrace", &xtrace);
    req_f(&P, "xbeg", &xbeg);
    get_f(&P, "zbeg", &zbeg);
    get_i(&P, "get_rec", &get_rec);
    get_i(&P, "plot_trace", &M->plot_trace);
    M->check = 1;                   /* sem2d_psvsh.c:47,179 */
    get_i(&P, "check", &M->check);
    get_i(&P, "output_vd", &M->output_vd);
    M->startsnap = 100;             /* sem2d_psvsh.c:149-150,202-204 */
    M->incresnap = 50;
    get_i(&P, "startsnap", &M->startsnap);
    get_i(&P, "incresnap", &M->incresnap);
    get_i(&P, "plot_snap", &M->plot_snap);
    if (get_rec == 1) req_s(&P, "rec_file", rec_file, sizeof rec_file);
    req_i(&P, "idxplot", &idxplot);
    req_i(&P, "msource", &msource);
    get_s(&P, "semdip", semdip, sizeof semdip);
    get_i(&P, "kahan", &kahan);
    snprintf(M->outdir, sizeof M->outdir, ".");
    get_s(&P, "outdir", M->outdir, sizeof M->outdir);
    int atten = 0;
    float ass = 0.93f;
    get_i(&P, "atten", &atten);
    get_f(&P, "ass", &ass);
    int stop = pl_get(&P.L, "STOP") != NULL;
    if (st == SEMSKS_OK && P.missing)
        st = semsks_set_error(SEMSKS_EPAR, "must specify value for '%s'", P.missing);
    if (st == SEMSKS_OK && stop) st = semsks_set_error(SEMSKS_ESTOP, "stop due to STOP in input");
    pl_free(&P.L);
    if (st == SEMSKS_OK && msource != 0) st = semsks_set_error(SEMSKS_EINVAL, "msource=1 (moment source) is not part of the GRT-boundary path");
    if (st == SEMSKS_OK && (hpn < 3 || hpn > 6)) st = semsks_set_error(SEMSKS_EINVAL, "hpn=%d outside 3..6", hpn);
    if (st == SEMSKS_OK && (solver < 0 || solver > 2)) st = semsks_set_error(SEMSKS_EINVAL, "solver=%d outside 0..2", solver);
    if (st == SEMSKS_OK && (nex <= neh || nst < 1 || !(sdt > 0) || !(dp > 0))) st = semsks_set_error(SEMSKS_EINVAL, "bad sizes nex=%d nst=%d sdt=%g dp=%g", nex, nst, sdt, dp);
    if (st != SEMSKS_OK) { semsks_model_free(M); return st; }

    int nez = 0, nrec = 0, jo = 0;
    float *cc = NULL, *ss = NULL, *dd = NULL, *tth = NULL, thk = 0;
    st = read_raymodel(raymodel, flat, nfil, eh, &nez, &eh, &nrec, &jo, &cc, &ss, &dd, &tth, &thk);
    if (st != SEMSKS_OK) { semsks_model_free(M); return st; }
    M->xmin = xmin;
    p->nex = nex; p->nez = nez; p->hpn = hpn; p->neh = neh; p->eh = eh; p->sdt = sdt; p->nst = nst;
    p->solver = solver; p->absorb = absorb; p->flags = kahan ? 0 : SEMSKS_F_NO_KAHAN;
    p->atten = atten; p->ass = ass;
    M->nx = nex * hpn + 1;
    M->nz = nez * hpn + 1;
    if (nez <= neh) st = semsks_set_error(SEMSKS_EINVAL, "nez=%d must exceed neh=%d", nez, neh);
    if (st == SEMSKS_OK) st = assign_material(M, semmodel, semdip, readpars, nfil, nrec, jo, cc, ss, dd, tth);
    free(cc); free(ss); free(dd); free(tth);
    if (st != SEMSKS_OK) { semsks_model_free(M); return st; }
    p->mat_table = M->mat_table;
    p->mat_id = M->mat_id;

    /* receivers (sem2d_psvsh.c:330-370) */
    {
        const float h = eh / hpn;
        int ixbeg = (int)(xbeg / h - 1), izbeg = (int)(zbeg / h);
        izbeg += M->nz - 1;
        if (zbeg == 0.) izbeg = M->nz - 1;
        if (izbeg > M->nz - 1) izbeg = M->nz - 1;
        if (get_rec == 0) {
            p->ntrace = xtrace;
            M->rec_ix = (int *)calloc(xtrace + 1, sizeof(int));
            M->rec_iz = (int *)calloc(xtrace + 1, sizeof(int));
            for (int i = 0; i < xtrace; ++i) { M->rec_ix[i] = ixbeg + idxplot * i; M->rec_iz[i] = izbeg; }
        } else {
            FILE *f = fopen(rec_file, "r");
            int n = 0;
            if (!f) { semsks_model_free(M); return semsks_set_error(SEMSKS_EIO, "Error Opening file %s with access rt", rec_file); }
            if (fscanf(f, "%d\n", &n) != 1 || n < 0) n = 0;
            p->ntrace = n;
            M->rec_ix = (int *)calloc(n + 1, sizeof(int));
            M->rec_iz = (int *)calloc(n + 1, sizeof(int));
            for (int i = 0; i < n; ++i) {
                float rx = 0.f;
                if (fscanf(f, "%f", &rx) != 1) { p->ntrace = i; break; }
                M->rec_ix[i] = (int)(rx / h);
                M->rec_iz[i] = izbeg;
            }
            fclose(f);
        }
        for (int i = 0; i < p->ntrace; ++i)
            if (M->rec_ix[i] < 0 || M->rec_ix[i] >= M->nx || M->rec_iz[i] < 0 || M->rec_iz[i] >= M->nz) {
                int bad = i;
                semsks_model_free(M);
                return semsks_set_error(SEMSKS_EINVAL, "receiver %d outside the mesh", bad);
            }
        p->rec_ix = M->rec_ix;
        p->rec_iz = M->rec_iz;
    }

    /* GRT records (sem2d_psvsh.c:372-396) */
    {
        int mt = 2 * (M->nx + M->nz), sht = M->nx + M->nz, got = 0;
        p->nt = nt;
        p->dp = dp;
        M->nt_sem = (int)((float)((nt - 3) * dp / sdt));
        st = read_records(greenfile, nt, mt, &M->psv, &got);
        if (st == SEMSKS_OK && solver == SEMSKS_SOLVER_PSVSH) st = read_records(shgreenfile, nt, sht, &M->sh, &got);
        if (st != SEMSKS_OK) { semsks_model_free(M); return st; }
        p->psv = M->psv;
        p->sh = M->sh;
    }
    *out = M;
    return SEMSKS_OK;
}

semsks_status semsks_write_seis(const char *outdir, int nst, int ntrace, const float *tr)
{
    static const char *names[3] = {"seis.u", "seis.w", "seis.v"};
    for (int k = 0; k < 3; ++k) {
        char path[2048];
        snprintf(path, sizeof path, "%s/%s", outdir && *outdir ? outdir : ".", names[k]);
        FILE *f = fopen(path, "wb");
        if (!f) return semsks_set_error(SEMSKS_EIO, "Error Opening file %s with access wb", path);
        for (int it = 0; it < nst; ++it)
            if (fwrite(tr + ((size_t)it * 3 + k) * ntrace, sizeof(float), ntrace, f) != (size_t)ntrace) {
                fclose(f);
                return semsks_set_error(SEMSKS_EIO, "short write on %s", path);
            }
        fclose(f);
    }
    return SEMSKS_OK;
}

/* ---------------------------------------------------------------------- */
/* check=1 diagnostics: grid.xz, legendre.xz, vel.xy                       */
/* ---------------------------------------------------------------------- */
/* GLL nodes in float as the reference computes them (lagrange.c:323-366:
   closed forms in double, stored to float; for degrees 5 and 6 the
   quadratic's coefficients are floats, so c*c-4*a*e is a float product) */
static void gll_float(int n, float *x)
{
    x[0] = -1.f;
    x[n] = 1.f;
    if (n == 3) { x[1] = (float)-sqrt(1 / 5.); x[2] = (float)sqrt(1 / 5.); }
    if (n == 4) { x[1] = (float)-sqrt(15. / 35.); x[2] = 0.f; x[3] = (float)sqrt(15. / 35.); }
    if (n == 5 || n == 6) {
        const float a = n == 5 ? 315.f : 1386.f, c = n == 5 ? -210.f : -1260.f, e = n == 5 ? 15.f : 210.f;
        const float disc = c * c - 4 * a * e;
        const float r1 = (float)sqrt((-c + sqrt(disc)) / 2 / a), r2 = (float)sqrt((-c - sqrt(disc)) / 2 / a);
        x[1] = -r1; x[2] = -r2;
        if (n == 5) { x[3] = r2; x[4] = r1; }
        else { x[3] = 0.f; x[4] = r2; x[5] = r1; }
    }
}

/* Lagrange basis i of the nodes x at s, in float (lagrange.c:20-32) */
static float lagrange_float(int n, int i, const float *x, float s)
{
    float l = 1.f;
    for (int j = 0; j <= n; ++j)
        if (j != i) l = l * (s - x[j]) / (x[i] - x[j]);
    return l;
}

static FILE *open_out(const semsks_model *m, const char *name, char *path, size_t cap)
{
    snprintf(path, cap, "%s/%s", m->outdir[0] ? m->outdir : ".", name);
    return fopen(path, "w");
}

semsks_status semsks_write_check(const semsks_model *m)
{
    if (!m) return semsks_set_error(SEMSKS_EINVAL, "bad arguments");
    if (m->check != 1) return SEMSKS_OK;
    const int nx = m->nx, nz = m->nz, hpn = m->prob.hpn;
    const float h = m->prob.eh / hpn;
    const float *T = m->mat_table;
    const uint16_t *id = m->mat_id;
    char path[2048];
    /* grid.xz (sem2d_psvsh.c:249-260): rho on every second node column and
       row, then, for each column, the top row's rho printed with the
       coordinates of the last row written */
    FILE *f = open_out(m, "grid.xz", path, sizeof path);
    if (!f) return semsks_set_error(SEMSKS_EIO, "Error Opening file %s with access wt", path);
    setvbuf(f, NULL, _IOFBF, 1 << 20);
    for (int i = 0; i < nx; i += 2) {
        float gx = 0.f, gz = 0.f;
        for (int j = 0; j < nz; j += 2) {
            gx = i * h;
            gz = (nz - 1 - j) * h;
            fprintf(f, "%-8.2f %-9.2f %9.4f\n", gx, gz, T[(size_t)id[(size_t)j * nx + i] * SEMSKS_MAT_STRIDE]);
        }
        fprintf(f, "%-8.2f %-9.2f %9.4f\n", gx, gz, T[(size_t)id[(size_t)(nz - 1) * nx + i] * SEMSKS_MAT_STRIDE]);
    }
    if (fclose(f) != 0) return semsks_set_error(SEMSKS_EIO, "short write on %s", path);
    /* legendre.xz (:262-270): each basis function on [-1, 1] in steps of
       0.05 (a float sum, as the reference's loop variable) */
    f = open_out(m, "legendre.xz", path, sizeof path);
    if (!f) return semsks_set_error(SEMSKS_EIO, "Error Opening file %s with access wt", path);
    float x[8];
    gll_float(hpn, x);
    for (int i = 0; i <= hpn; ++i) {
        fprintf(f, "> %-9.5f\n", x[i]);
        for (float lx = -1; lx <= 1.0; lx = (float)(lx + 0.05))
            fprintf(f, "%-8.2f %-9.3f\n", lx, lagrange_float(hpn, i, x, lx));
    }
    if (fclose(f) != 0) return semsks_set_error(SEMSKS_EIO, "short write on %s", path);
    /* vel.xy (:804-825): per node, x in degrees of epicentral distance,
       depth, and vp = sqrt(c11/rho) or, with output_vd, the shear-wave
       splitting (v44 - v55)/(v44 + v55)/2 */
    f = open_out(m, "vel.xy", path, sizeof path);
    if (!f) return semsks_set_error(SEMSKS_EIO, "Error Opening file %s with access w", path);
    setvbuf(f, NULL, _IOFBF, 1 << 20);
    const double deg = 6371 * 3.14159265 / 180.;
    for (int i = 0; i < nx; ++i)
        for (int j = 0; j < nz; ++j) {
            const float *r = T + (size_t)id[(size_t)j * nx + i] * SEMSKS_MAT_STRIDE;
            float v;
            if (m->output_vd == 0) {
                v = (float)sqrt(r[1] / r[0]);
            } else {
                const float v44 = (float)sqrt(r[10] / r[0]), v55 = (float)sqrt(r[13] / r[0]);
                v = (float)((v44 - v55) / (v44 + v55) / 2.);
            }
            fprintf(f, "%-8.3f %-8.2f %-8.2f\n", (m->xmin + i * h) / deg, (nz - 1 - j) * h, v);
        }
    if (fclose(f) != 0) return semsks_set_error(SEMSKS_EIO, "short write on %s", path);
    return SEMSKS_OK;
}

void semsks_model_snapshots(const semsks_model *m, int k[3])
{
    k[0] = m->plot_snap; k[1] = m->startsnap; k[2] = m->incresnap;
}

semsks_status semsks_write_snapshot(const char *outdir, const char *name, int nx, int nz, const float *u)
{
    if (!name || !u || nx < 1 || nz < 1) return semsks_set_error(SEMSKS_EINVAL, "bad arguments");
    char path[2048];
    snprintf(path, sizeof path, "%s/%s", outdir && *outdir ? outdir : ".", name);
    FILE *f = fopen(path, "wt");
    if (!f) return semsks_set_error(SEMSKS_EIO, "Error Opening file %s with access wt", path);
    setvbuf(f, NULL, _IOFBF, 1 << 20);
    const size_t plane = (size_t)nx * nz;
    for (int ix = 0; ix < nx; ix += 2)
        for (int iz = 0; iz < nz; iz += 2) {
            const size_t q = (size_t)iz * nx + ix;
            fprintf(f, "%d %d %e %e\n", ix, iz, u[q], u[plane + q]);
        }
    if (fclose(f) != 0) return semsks_set_error(SEMSKS_EIO, "short write on %s", path);
    return SEMSKS_OK;
}

