/*
 * sem_oracle.c -- CPU restatement of the s-sks 2-D SEM time loop.
 *
 * TEST INFRASTRUCTURE ONLY.  The tests compare the HIP product path against
 * this file, and the bench times it as the CPU baseline.  Nothing in the
 * product library links, loads or calls it.
 *
 * The file is compiled twice:
 *   -DREAL=float  : "faithful" build.  Every expression keeps the operand types
 *                   of the reference, including its double-promoted sums
 *                   (2.*p, 2.*Mass*u, sdt/2., 1.e-3), so that the arithmetic is
 *                   the reference's own float arithmetic.
 *   -DREAL=double : FP64 oracle.  Same algorithm, same inputs (parameters are
 *                   parsed as float exactly like the reference), same discrete
 *                   structure (mesh sizes, layer cuts, receiver indices, record
 *                   schedule), with every value-carrying operation in double.
 * Both must be compiled with -ffp-contract=off and without -ffast-math.
 *
 * Reference files restated (paths relative to the reference root):
 *   semsyn/src/sem2d_psvsh.c  main loop 418-512, discret 534-828, lip 831-846,
 *                             Time_diff_hybrid 1004-1029, Mass_Matrix 1042-1071,
 *                             Absorb_Matrix 1093-1162, Stiff_Matrix 1173-1256,
 *                             Grid_Stiff 1274-1313, F_ik 1323-1336,
 *                             Stress_ij 1345-1379, dUj_di_2D 1444-1482,
 *                             Cijkl_3D 1532-1572, inv_index 1584-1597,
 *                             atten_wave2 1780-1823, Absorbs 1834-1889,
 *                             GRT_Interface 1953-2005, GRT_Interface3 2045-2074,
 *                             GRT_Interface4 2076-2105, kirrecord_zhl 2116-2165
 *   semsyn/src/sem2d.c        single-stream P-SV variant (main 401-489)
 *   semsyn/src/sem2dMP20140816.c  GRT_Interface3 variant (433, 2001-2037)
 *   semsyn/src/interpl.c      interpl 16-54, interp_const 59-91
 *   semsyn/src/lagrange.c     Lag_N_i 20-33, dLag_N_i 42-67, Legendre 102-122,
 *                             GLLWeight 256-278, GLL 323-366
 *   semsyn/src/readmodel.c    open_file 9-21, read_model 43-109, flatten 150-165
 *   semsyn/src/getpar/src/getpar.c  setpar/getpar/mstpar/endpar semantics
 *
 * How each part is pinned: checked against something it does not restate
 * (tests/test_oracle_pins.py unless another file is named):
 *   GLL, GLLWeight, dLag_N_i  closed forms (hpn 3-5), quadrature exactness,
 *                             derivative-matrix identities
 *   interpl, interp_const     bit-exact against the reference's own compiled
 *                             interpl.c (tests/golden/ref_interpl*.npz)
 *   the +-10 guard            its sequential rule traced by hand
 *   read_model, flatten       bit-exact against the reference's own compiled
 *                             readmodel.c (tests/golden/ref_readmodel.npz)
 *   discret mesh sizes, receiver indices   the float-arithmetic cases
 *   semmodel fill, Cijkl_3D, inv_index     Christoffel matrices of a
 *                             closed-form transversely isotropic tensor
 *                             (tests/test_material_pins.py)
 *   layer cuts (lip)          the interface polylines' geometry
 *                             (tests/test_material_pins.py)
 *   atten_taper, atten_wave2  closed-form factors and strip membership
 *                             (tests/test_material_pins.py)
 *   Stiff_Matrix and callees  independent B-matrix assembly, null space,
 *                             patch test
 *   Mass_Matrix               sum of M = integral of rho
 *   Time_diff_hybrid          discrete energy conserved without absorption
 *   Absorb_Matrix, Absorbs    a normally incident plane pulse leaves through
 *                             the Lysmer edge and returns from a free one
 *   GRT_Interface3/4          closed-form amplitude radiated by the kick
 *   stiff_fast (colour-parallel sweep)   within 1e-12 of Stiff_Matrix's order
 *   REAL=float vs REAL=double tracking bound on a 300-step run
 *   getpar semantics, status codes, greenfile layout   tests/test_formats.py
 *   kirrecord_zhl             a vertically incident SV, SH or P pulse reaches
 *                             the surface receivers in its own channel
 *                             (u, -w, v) with the greenfile's sign and twice
 *                             the radiated amplitude
 */
#include <stdio.h>
#include <stdlib.h>
#include <string.h>
#include <math.h>
#include <stdint.h>
#include <time.h>
#ifdef _OPENMP
#include <omp.h>
#endif

#ifndef REAL
#define REAL float
#endif
typedef REAL real;

#define ORC_EXPORT __attribute__((visibility("default")))

/* ------------------------------------------------------------------------ */
/* status codes (the reference calls exit() with these values)              */
/* ------------------------------------------------------------------------ */
#define ORC_OK          0
#define ORC_EIO        -1   /* open_file failure: exit(-1)  readmodel.c:15-19 */
#define ORC_ENP        -2   /* np>1000: exit(-2)            sem2d_psvsh.c:725-728 */
#define ORC_ERANGE     -3   /* lip out of range: exit(-3)   sem2d_psvsh.c:835-839 */
#define ORC_EPAR      100   /* missing mstpar: exit(100)    getpar.c:65,715-725 */
#define ORC_ESTOP     101   /* STOP in input: exit(101)     getpar.c:66,378-390 */
#define ORC_EINVAL    -22

static char orc_errmsg[512];
ORC_EXPORT const char *orc_last_error(void) { return orc_errmsg; }
static int orc_fail(int code, const char *msg, const char *arg)
{
    snprintf(orc_errmsg, sizeof orc_errmsg, "%s%s%s", msg, arg ? ": " : "", arg ? arg : "");
    return code;
}

/* ------------------------------------------------------------------------ */
/* getpar restatement: name=value arguments, par=file recursion (depth 8),   */
/* '#' comments, quotes, $(name) substitution, backward (last-wins) lookup   */
/* ------------------------------------------------------------------------ */
typedef struct { char *name; char *value; } orc_kv;
typedef struct { orc_kv *kv; int n, cap; int stop; int err; char errarg[256]; } orc_pars;

static void pars_add(orc_pars *p, const char *name, const char *value)
{
    if (p->n == p->cap) {
        p->cap = p->cap ? 2 * p->cap : 32;
        p->kv = (orc_kv *)realloc(p->kv, (size_t)p->cap * sizeof(orc_kv));
    }
    p->kv[p->n].name = strdup(name);
    p->kv[p->n].value = strdup(value);
    p->n++;
}

static const char *pars_find(const orc_pars *p, const char *name)
{
    int i;
    for (i = p->n - 1; i >= 0; i--)              /* last definition wins */
        if (strcmp(p->kv[i].name, name) == 0) return p->kv[i].value;
    return NULL;
}

/* copies a value, expanding $(name) from what is defined so far */
static const char *pars_value(const orc_pars *p, const char *s, char *out, size_t cap, int stop_at_space)
{
    size_t o = 0;
    char quote = 0;
    if (*s == '"' || *s == '\'') quote = *s++;
    while (*s && *s != '\n') {
        if (quote) {
            if (*s == quote) {
                if (o > 0 && out[o - 1] == '\\') { out[o - 1] = quote; s++; continue; }
                s++;
                break;
            }
        } else if (stop_at_space && (*s == ' ' || *s == '\t')) {
            break;
        }
        if (*s == '$' && s[1] == '(') {
            const char *e = strchr(s + 2, ')');
            if (e) {
                char key[256];
                size_t kl = (size_t)(e - (s + 2));
                if (kl >= sizeof key) kl = sizeof key - 1;
                memcpy(key, s + 2, kl);
                key[kl] = 0;
                {
                    const char *v = pars_find(p, key);
                    if (v) while (*v && o + 1 < cap) out[o++] = *v++;
                }
                s = e + 1;
                continue;
            }
        }
        if (o + 1 < cap) out[o++] = *s;
        s++;
    }
    out[o] = 0;
    return s;
}

static void pars_file(orc_pars *p, const char *fname, int level)
{
    FILE *f;
    char line[8192];
    if (level > 8) { p->err = ORC_EPAR; snprintf(p->errarg, sizeof p->errarg, "too many recursive par files"); return; }
    if (!*fname) return;
    f = fopen(fname, "r");
    if (!f) { p->err = ORC_EPAR; snprintf(p->errarg, sizeof p->errarg, "cannot open par file %s", fname); return; }
    while (fgets(line, sizeof line, f)) {
        size_t L = strlen(line);
        /* backslash-newline continues the line */
        while (L >= 2 && line[L - 1] == '\n' && line[L - 2] == '\\') {
            line[L - 2] = 0;
            L -= 2;
            if (!fgets(line + L, (int)(sizeof line - L), f)) break;
            L = strlen(line);
        }
        {
            const char *s = line;
            for (;;) {
                char name[256], value[5120];
                size_t nl = 0;
                while (*s == ' ' || *s == '\t') s++;
                if (*s == 0 || *s == '\n' || *s == '#') break;
                while (*s && *s != '=' && *s != ' ' && *s != '\t' && *s != '\n') {
                    if (nl + 1 < sizeof name) name[nl++] = *s;
                    s++;
                }
                name[nl] = 0;
                if (*s == '=') s++;
                s = pars_value(p, s, value, sizeof value, 1);
                pars_add(p, name, value);
                if (strcmp(name, "par") == 0) pars_file(p, value, level + 1);
            }
        }
    }
    fclose(f);
}

static void pars_init(orc_pars *p, int argc, char **argv)
{
    int i;
    memset(p, 0, sizeof *p);
    for (i = 1; i < argc; i++) {
        const char *a = argv[i];
        char name[256], value[5120];
        size_t nl = 0;
        while (*a == ' ' || *a == '\t') a++;
        while (*a && *a != '=') { if (nl + 1 < sizeof name) name[nl++] = *a; a++; }
        name[nl] = 0;
        if (*a == '=') a++;
        pars_value(p, a, value, sizeof value, 0);
        if (name[0] == '-') pars_add(p, "SWITCH", name + 1);
        else pars_add(p, name, value);
        if (strcmp(name, "par") == 0) pars_file(p, value, 1);
        if (strcmp(name, "ENDPAR") == 0) break;
    }
    {
        const char *st = pars_find(p, "STOP");
        if (st) p->stop = 1;
    }
}

static void pars_free(orc_pars *p)
{
    int i;
    for (i = 0; i < p->n; i++) { free(p->kv[i].name); free(p->kv[i].value); }
    free(p->kv);
}

static int gp_int(orc_pars *p, const char *n, int *v) { const char *s = pars_find(p, n); if (!s) return 0; *v = atoi(s); return 1; }
static int gp_flt(orc_pars *p, const char *n, float *v) { const char *s = pars_find(p, n); if (!s) return 0; *v = (float)atof(s); return 1; }
static int gp_str(orc_pars *p, const char *n, char *v, size_t cap) { const char *s = pars_find(p, n); if (!s) return 0; snprintf(v, cap, "%s", s); return 1; }
static void mst_missing(orc_pars *p, const char *n) { if (!p->err) { p->err = ORC_EPAR; snprintf(p->errarg, sizeof p->errarg, "must specify value for '%s'", n); } }
static void mst_int(orc_pars *p, const char *n, int *v) { if (!gp_int(p, n, v)) mst_missing(p, n); }
static void mst_flt(orc_pars *p, const char *n, float *v) { if (!gp_flt(p, n, v)) mst_missing(p, n); }
static void mst_str(orc_pars *p, const char *n, char *v, size_t cap) { if (!gp_str(p, n, v, cap)) mst_missing(p, n); }

/* ------------------------------------------------------------------------ */
/* material record per GLL node: the full c[3][3][3][3] tensor plus rho,     */
/* as the reference's struct cmatrix (82 values per node)                    */
/* ------------------------------------------------------------------------ */
typedef struct { real c[3][3][3][3]; real rho; } orc_cmat;
typedef struct { real u[3]; } orc_vec;

/* 13 monoclinic constants (+6 dipping extensions) and rho, in Voigt naming */
typedef struct {
    float c11, c12, c13, c16, c22, c23, c26, c33, c36, c44, c45, c55, c66, rho;
    float c14, c15, c34, c35, c46, c56;   /* zero in reference-format models */
} orc_apoint;

/* Voigt index of (i,j): 00->0 11->1 22->2 12->3 02->4 01->5
   (inv_index, sem2d_psvsh.c:1584-1597) */
static int voigt(int i, int j)
{
    static const int t[3][3] = { {0, 5, 4}, {5, 1, 3}, {4, 3, 2} };
    return t[i][j];
}

/* Cijkl_3D (sem2d_psvsh.c:1532-1572), extended with the six constants a
   dipping symmetry axis adds.  With those six zero it is the reference map. */
static float cijkl(int i, int j, int k, int l, const orc_apoint *a)
{
    int m = voigt(i, j), n = voigt(k, l), s, t;
    if (m > n) { s = m; m = n; n = s; }
    t = m * 6 + n;
    switch (t) {
    case 0:  return a->c11;  case 1:  return a->c12;  case 2:  return a->c13;
    case 3:  return a->c14;  case 4:  return a->c15;  case 5:  return a->c16;
    case 7:  return a->c22;  case 8:  return a->c23;  case 11: return a->c26;
    case 14: return a->c33;  case 15: return a->c34;  case 16: return a->c35;
    case 17: return a->c36;  case 21: return a->c44;  case 22: return a->c45;
    case 23: return a->c46;  case 28: return a->c55;  case 29: return a->c56;
    case 35: return a->c66;
    default: return 0.f;   /* c24 c25 stay zero: they never enter 2-D stress */
    }
}

static void fill_tensor(orc_cmat *dst, const orc_apoint *a)
{
    int i, j, k, l;
    for (i = 0; i < 3; i++) for (j = 0; j < 3; j++) for (k = 0; k < 3; k++) for (l = 0; l < 3; l++)
        dst->c[i][j][k][l] = (real)cijkl(i, j, k, l, a);
}

/* ------------------------------------------------------------------------ */
/* GLL / Lagrange tables  (lagrange.c)                                      */
/* ------------------------------------------------------------------------ */
static real legendre(int n, real x)
{
    int i;
    real pn = 0, p1, p2;
    if (n == 0) return 1;
    if (n == 1) return x;
    p2 = 1; p1 = x;
    for (i = 2; i <= n; i++) {
        pn = ((2 * i - 1) * x * p1 - (i - 1) * p2) / i;
        p2 = p1;
        p1 = pn;
    }
    return pn;
}

ORC_EXPORT int orc_gll(int np, real *gllp, real *gllw)
{
    int i;
    if (np < 3 || np > 6) return orc_fail(ORC_EINVAL, "GLL degree outside 3..6", NULL);
    if (np == 3) { gllp[0] = -1.; gllp[1] = (real)-sqrt(1 / 5.); gllp[2] = (real)sqrt(1 / 5.); gllp[3] = 1.; }
    if (np == 4) { gllp[0] = -1.; gllp[1] = (real)-sqrt(15. / 35.); gllp[2] = 0.; gllp[3] = (real)sqrt(15. / 35.); gllp[4] = 1.; }
    if (np == 5 || np == 6) {
        real a = np == 5 ? 315. : 1386., c = np == 5 ? -210. : -1260., e = np == 5 ? 15. : 210.;
        real disc = c * c - 4 * a * e;
        real r1 = (real)sqrt((-c + sqrt(disc)) / 2 / a), r2 = (real)sqrt((-c - sqrt(disc)) / 2 / a);
        gllp[0] = -1.;
        gllp[1] = -r1; gllp[2] = -r2;
        if (np == 5) { gllp[3] = r2; gllp[4] = r1; gllp[5] = 1.; }
        else { gllp[3] = 0.; gllp[4] = r2; gllp[5] = r1; gllp[6] = 1.; }
    }
    for (i = 0; i <= np; i++) {
        real w;
        if (i == 0 || i == np) w = (real)(2. / (np * (np + 1)));
        else {
            real t = legendre(np, gllp[i]);
            w = (real)(2. / (np * (np + 1) * t * t));
        }
        gllw[i] = w;
    }
    return ORC_OK;
}

/* l'_alpha(seta) by the product rule (dLag_N_i, lagrange.c:42-67) */
static real dlag(int n, int alpha, const real *g, real seta)
{
    int i, k;
    real value = 0, down = 1.0;
    for (i = 0; i <= n; i++) if (i != alpha) down = down * (g[alpha] - g[i]);
    for (i = 0; i <= n; i++) {
        real L = 1;
        if (i == alpha) continue;
        for (k = 0; k <= n; k++) if (k != alpha && k != i) L = L * (seta - g[k]);
        value = value + L;
    }
    return value / down;
}

/* dl[j*(n+1)+i] = l'_i(gllp[j])   (sem2d_psvsh.c:235-246) */
ORC_EXPORT void orc_dlag(int n, const real *gllp, real *dl)
{
    int i, j;
    for (i = 0; i <= n; i++)
        for (j = 0; j <= n; j++) dl[j * (n + 1) + i] = dlag(n, i, gllp, gllp[j]);
}

/* ------------------------------------------------------------------------ */
/* solver context                                                            */
/* ------------------------------------------------------------------------ */
typedef struct {
    /* parameters (sem2d_psvsh.c:154-205) */
    int flat, nt, nex, nst, nfil, hpn, absorb, readpars, neh, atten, check, output_vd;
    int xtrace, ztrace, idxplot, idzplot, get_rec, plot_trace, msource;
    float dp, eh, xmin, sdt, ass, xbeg, zbeg;
    char greenfile[1024], shgreenfile[1024], raymodel[1024], semmodel[1024], rec_file[1024];
    char semdip[1024];    /* optional: six tilted-axis constants per semmodel layer */
    int variant;          /* 0 psvsh/Interface4, 1 P-SV/Interface4, 2 P-SV/Interface3 */
    /* mesh */
    int nez, nx, nz, nel;
    float h;
    real *gllp, *gllw, *dl, *dl_jei, *dl_gllw;
    real je, jei, jb;
    orc_cmat *mesh;
    /* fast FP64 path (orc_run_fast): the distinct material records and a
       record id per node, built on first use */
    orc_cmat *rtab;
    int *rid, nrtab;
    real *mass, *mdt2;
    orc_vec *ab;
    /* receivers */
    int ntrace;
    int *ix1, *iz1;
    /* boundary */
    int nt_sem;
    float rdt;
    float *taper;
} orc_ctx;

static void *xcalloc(size_t n, size_t s) { void *p = calloc(n ? n : 1, s); if (!p) { fprintf(stderr, "oracle: out of memory\n"); abort(); } return p; }

/* read_model (readmodel.c:43-109): layers and the ray list.  Only nrec and
   the thicknesses are used by the solver. */
typedef struct { int jo, nb, lfinal; float *cc, *ss, *dd, *tth; int *nen, *na; } orc_raymodel;

static int read_raymodel(const char *fn, int flat, orc_raymodel *m)
{
    FILE *f = fopen(fn, "r");
    int j, n, k = 0;
    if (!f) return orc_fail(ORC_EIO, "Error opening file", fn);
    memset(m, 0, sizeof *m);
    if (fscanf(f, "%d %d\n", &m->jo, &m->nb) != 2) { fclose(f); return orc_fail(ORC_EIO, "bad raymodel header", fn); }
    m->cc = (float *)xcalloc((size_t)m->jo + 1, sizeof(float));
    m->ss = (float *)xcalloc((size_t)m->jo + 1, sizeof(float));
    m->dd = (float *)xcalloc((size_t)m->jo + 1, sizeof(float));
    m->tth = (float *)xcalloc((size_t)m->jo + 1, sizeof(float));
    for (j = 0; j < m->jo; j++)
        if (fscanf(f, "%f %f %f %f\n", &m->cc[j], &m->ss[j], &m->dd[j], &m->tth[j]) != 4) { fclose(f); return orc_fail(ORC_EIO, "bad raymodel layer", fn); }
    if (flat == 0) {   /* earth flattening (readmodel.c:69-84,150-165) */
        double *q = (double *)xcalloc((size_t)m->jo + 1, sizeof(double));
        for (j = 0; j < m->jo; j++) {
            double th = 0.0; int i;
            for (i = 0; i < j; i++) th += m->tth[i];
            th += 0.5 * m->tth[j];
            q[j] = 6371.0 / (6371.0 - th);
        }
        for (j = 0; j < m->jo; j++) { m->cc[j] *= q[j]; m->ss[j] *= q[j]; m->dd[j] *= q[j]; m->tth[j] *= q[j]; }
        free(q);
    }
    if (fscanf(f, "%d\n", &m->lfinal) != 1) { fclose(f); return orc_fail(ORC_EIO, "bad raymodel ray count", fn); }
    m->nen = (int *)xcalloc(m->lfinal + 1, sizeof(int));
    m->na = (int *)xcalloc(100000, sizeof(int));
    for (n = 0; n < m->lfinal; n++) {
        int dummy;
        if (fscanf(f, "%d", &m->nen[n]) != 1) break;
        if (k + m->nen[n] > 100000) break;
        for (j = 0; j < m->nen[n]; j++) if (fscanf(f, " %d", &m->na[k + j]) != 1) break;
        if (fscanf(f, "%d", &dummy) != 1) break;
        for (j = 0; j < m->nen[n]; j++) if (fscanf(f, " %d", &dummy) != 1) break;
        k += m->nen[n];   /* ray n's layer list starts at the sum of earlier nen */
    }
    fclose(f);
    return ORC_OK;
}

static void free_raymodel(orc_raymodel *m) { free(m->cc); free(m->ss); free(m->dd); free(m->tth); free(m->nen); free(m->na); }

typedef struct { int np; int ratio; orc_apoint a; float *an, *bn; } orc_semlayer;

/* lip (sem2d_psvsh.c:831-846): piecewise-linear interface height at x */
static int lip(const orc_semlayer *L, float x, float *dep)
{
    int i;
    float d = 0.f;
    if (x < L->an[0] || x > L->an[L->np - 1]) return orc_fail(ORC_ERANGE, "Model out of range", NULL);
    for (i = 0; i < L->np - 1; i++)
        if (x >= L->an[i] && x < L->an[i + 1]) {
            d = (L->bn[i + 1] - L->bn[i]) / (L->an[i + 1] - L->an[i]);
            d = L->bn[i] + d * (x - L->an[i]);
        }
    *dep = d;
    return ORC_OK;
}

/* discret (sem2d_psvsh.c:534-828): mesh sizes and the per-node material */
static int discret(orc_ctx *c)
{
    orc_raymodel rm;
    int rc, i, j, k, nrec, iz1, iz2;
    float thickness = 0.0f, thk2 = 0.0f;
    int nlay1, nlay2, nlaynum, inlay;
    rc = read_raymodel(c->raymodel, c->flat, &rm);
    if (rc) return rc;
    nrec = rm.na[rm.nen[0] - 2];
    if (c->nfil > nrec) { for (i = nrec; i < c->nfil; i++) thickness += rm.tth[i]; }
    else { for (i = c->nfil - 1; i < nrec - 1; i++) thickness += rm.tth[i]; }
    c->nez = (int)((thickness - 1.e-3) / c->eh);
    c->eh = (float)((thickness - 1.e-3) / c->nez);
    c->nel = c->nex * c->nez;
    c->nx = c->nex * c->hpn + 1;
    c->nz = c->nez * c->hpn + 1;
    c->h = c->eh / c->hpn;
    c->mesh = (orc_cmat *)xcalloc((size_t)c->nx * c->nz, sizeof(orc_cmat));

    /* step 1: 1-D raymodel prefill, bottom up */
    iz1 = 0;
    nlay1 = (c->nfil > nrec) ? nrec : nrec - 2;
    nlay2 = (c->nfil > nrec) ? c->nfil + 200 : c->nfil - 200;
    if (nlay2 < 0) nlay2 = 0;
    nlaynum = nlay2 - nlay1;
    if (nlaynum < 0) nlaynum = -nlaynum;
    nlaynum += 1;
    inlay = (c->nfil > nrec) ? 1 : -1;
    for (k = 0; k < nlaynum; k++) {
        orc_apoint a;
        orc_cmat t;
        float ivp, ivs, iden, t11, t44, t13;
        i = nlay1 + k * inlay;
        thk2 += rm.tth[i];
        iz2 = (int)((thk2 + 1.e-3) / c->h) + 1;
        if (iz2 > c->nz - 1) iz2 = c->nz - 1;
        if (iz2 < iz1) break;
        ivp = rm.cc[i] * rm.cc[i]; ivs = rm.ss[i] * rm.ss[i]; iden = rm.dd[i];
        t11 = ivp * iden; t44 = ivs * iden;
        t13 = (float)(t11 - 2.0 * t44);
        memset(&a, 0, sizeof a);
        a.c11 = t11; a.c12 = t13; a.c13 = t13; a.c22 = t11; a.c23 = t13;
        a.c33 = t11; a.c44 = t44; a.c55 = t44; a.c66 = t44;
        fill_tensor(&t, &a);
        t.rho = iden;
        for (j = iz1; j <= iz2; j++) {
            int x;
            for (x = 0; x < c->nx; x++) c->mesh[(size_t)j * c->nx + x] = t;
        }
        iz1 = iz2 + 1;
        if (iz2 == c->nz - 1) break;
    }
    free_raymodel(&rm);

    /* step 2: semmodel layers, per column, bottom up */
    if (c->readpars == 1) {
        FILE *f = fopen(c->semmodel, "r");
        int nll, firstlayer = 0;
        orc_semlayer *L;
        orc_apoint ap;
        if (!f) return orc_fail(ORC_EIO, "Error opening file", c->semmodel);
        if (fscanf(f, "%d\n", &nll) != 1) { fclose(f); return orc_fail(ORC_EIO, "bad semmodel", c->semmodel); }
        if (nll < 0) { firstlayer = 1; nll = -nll; }
        L = (orc_semlayer *)xcalloc((size_t)nll + 2, sizeof(orc_semlayer));
        for (j = 0; j < nll; j++) {
            orc_apoint *a = &L[j].a;
            int np;
            float v;
            if (fscanf(f, "%d %f %f %f %f %f %f %f", &np, &a->c11, &a->c12, &a->c13, &a->c16, &a->c22, &a->c23, &a->c26) != 8
                || fscanf(f, "%f %f %f %f %f %f %f\n", &a->c33, &a->c36, &a->c44, &a->c45, &a->c55, &a->c66, &a->rho) != 7) {
                fclose(f); return orc_fail(ORC_EIO, "bad semmodel layer", c->semmodel);
            }
            if (np < 0) { L[j].ratio = 1; np = -np; }
            if (np > 1000) { fclose(f); return orc_fail(ORC_ENP, "an bn allocation error, np > 1000", NULL); }
            L[j].np = np;
            L[j].an = (float *)xcalloc(1001, sizeof(float));
            L[j].bn = (float *)xcalloc(1001, sizeof(float));
            for (i = 0; i < np; i++) { if (fscanf(f, "%f ", &v) != 1) break; L[j].an[i] = v; }
            fscanf(f, "\n");
            for (i = 0; i < np; i++) { if (fscanf(f, "%f ", &v) != 1) break; L[j].bn[i] = v; }
            fscanf(f, "\n");
        }
        fclose(f);
        /* extension: c14 c15 c34 c35 c46 c56 per layer for a tilted symmetry
           axis, which the 13-constant semmodel cannot hold */
        if (c->semdip[0]) {
            FILE *g = fopen(c->semdip, "r");
            if (!g) return orc_fail(ORC_EIO, "Error opening file", c->semdip);
            for (j = 0; j < nll; j++) {
                orc_apoint *a = &L[j].a;
                if (fscanf(g, "%f %f %f %f %f %f", &a->c14, &a->c15, &a->c34, &a->c35, &a->c46, &a->c56) != 6) {
                    fclose(g);
                    return orc_fail(ORC_EIO, "bad semdip", c->semdip);
                }
            }
            fclose(g);
        }
        memset(&ap, 0, sizeof ap);   /* the reference leaves this uninitialised */
        for (i = 0; i < c->nx; i++) {
            float x = i * c->h, dep;
            iz1 = 0;
            if (firstlayer == 1) {
                rc = lip(&L[0], x, &dep);
                if (rc) return rc;
                iz1 = (int)((dep + 1.e-3) / c->h);
            }
            if (iz1 < 0) iz1 = 0;
            for (j = firstlayer; j < nll; j++) {
                orc_cmat t;
                int iz;
                rc = lip(&L[j], x, &dep);
                if (rc) return rc;
                iz2 = (int)((dep + 1.e-3) / c->h);
                if (iz2 > c->nz - 1) iz2 = c->nz - 1;
                if (L[j].ratio == 1) {
                    const orc_apoint *m = &L[j].a;
                    ap.c11 *= m->c11; ap.c12 *= m->c12; ap.c13 *= m->c13; ap.c16 *= m->c16;
                    ap.c22 *= m->c22; ap.c23 *= m->c23; ap.c26 *= m->c26; ap.c33 *= m->c33; ap.c36 *= m->c36;
                    ap.c44 *= m->c44; ap.c45 *= m->c45; ap.c55 *= m->c55; ap.c66 *= m->c66;
                    ap.rho *= m->rho;
                } else {
                    ap = L[j].a;
                }
                fill_tensor(&t, &ap);
                t.rho = ap.rho;
                for (iz = iz1; iz <= iz2; iz++) c->mesh[(size_t)iz * c->nx + i] = t;
                iz1 = iz2 + 1;
            }
        }
        for (j = 0; j < nll; j++) { free(L[j].an); free(L[j].bn); }
        free(L);
    }
    return ORC_OK;
}

/* optional override: per-layer dipping constants (c14 c15 c34 c35 c46 c56)
   for models that the 13-constant semmodel cannot hold.  Applied to every
   node whose (c11, rho) pair matches the given layer record. */
ORC_EXPORT int orc_set_node_tensor(orc_ctx *c, long node, const float *c21, float rho)
{
    /* c21: upper triangle of the 6x6 Voigt matrix, row-major (21 values) */
    orc_apoint a;
    int m, n, q = 0;
    float C[6][6];
    if (node < 0 || node >= (long)c->nx * c->nz) return ORC_EINVAL;
    for (m = 0; m < 6; m++) for (n = m; n < 6; n++) { C[m][n] = C[n][m] = c21[q++]; }
    memset(&a, 0, sizeof a);
    a.c11 = C[0][0]; a.c12 = C[0][1]; a.c13 = C[0][2]; a.c14 = C[0][3]; a.c15 = C[0][4]; a.c16 = C[0][5];
    a.c22 = C[1][1]; a.c23 = C[1][2]; a.c26 = C[1][5];
    a.c33 = C[2][2]; a.c34 = C[2][3]; a.c35 = C[2][4]; a.c36 = C[2][5];
    a.c44 = C[3][3]; a.c45 = C[3][4]; a.c46 = C[3][5];
    a.c55 = C[4][4]; a.c56 = C[4][5]; a.c66 = C[5][5];
    a.rho = rho;
    fill_tensor(&c->mesh[node], &a);
    c->mesh[node].rho = rho;
    return ORC_OK;
}

/* Mass_Matrix (sem2d_psvsh.c:1042-1071), lumped, element loop ex,ez,m,n */
static void mass_matrix(orc_ctx *c)
{
    int i, j, m, n, np1 = c->hpn + 1;
    size_t N = (size_t)c->nx * c->nz;
    for (i = 0; i < (int)N; i++) c->mass[i] = 0.;
    for (i = 0; i < c->nex; i++)
        for (j = 0; j < c->nez; j++)
            for (m = 0; m < np1; m++) {
                real rx = c->gllw[m];
                for (n = 0; n < np1; n++) {
                    real rz = c->gllw[n];
                    size_t idx = (size_t)(j * c->hpn + n) * c->nx + i * c->hpn + m;
                    c->mass[idx] = c->mass[idx] + rx * rz * c->je * c->mesh[idx].rho;
                }
            }
}

/* Absorb_Matrix (sem2d_psvsh.c:1093-1162): Lysmer coefficients * sdt/2 */
static void absorb_add(orc_vec *ab, const orc_cmat *m, real w, real jb, float sdt, int normal)
{
    real vp = (real)sqrt(m->c[0][0][0][0] / m->rho);
    real vs = (real)sqrt(m->c[1][2][1][2] / m->rho);
    int k;
    for (k = 0; k < 3; k++) {
        real v = (k == normal) ? vp : vs;
        ab->u[k] = (real)(ab->u[k] + w * m->rho * v * jb * sdt / 2.);
    }
}

static void absorb_matrix(orc_ctx *c)
{
    int ex, ez, a, np1 = c->hpn + 1;
    for (ez = 0; ez < c->nez; ez++)               /* left, normal x */
        for (a = 0; a < np1; a++) {
            size_t idx = (size_t)(ez * c->hpn + a) * c->nx;
            absorb_add(&c->ab[idx], &c->mesh[idx], c->gllw[a], c->jb, c->sdt, 0);
        }
    for (ex = 0; ex < c->nex; ex++)               /* bottom, normal z */
        for (a = 0; a < np1; a++) {
            size_t idx = (size_t)ex * c->hpn + a;
            absorb_add(&c->ab[idx], &c->mesh[idx], c->gllw[a], c->jb, c->sdt, 2);
        }
    for (ez = 0; ez < c->nez; ez++)               /* right, normal x */
        for (a = 0; a < np1; a++) {
            size_t idx = (size_t)(ez * c->hpn + a) * c->nx + c->nx - 1;
            absorb_add(&c->ab[idx], &c->mesh[idx], c->gllw[a], c->jb, c->sdt, 0);
        }
}

/* ------------------------------------------------------------------------ */
/* element force (Stiff_Matrix and callees)                                  */
/* ------------------------------------------------------------------------ */
typedef struct { real F0[3], F1[3], F2[3]; } orc_fik;
#define SEMSKS_ORC_MAXNP 7

/* Stress_ij (sem2d_psvsh.c:1345-1379); the material record is passed by value */
static real stress_ij(int i, int j, real ukl[3][3], orc_cmat s)
{
    real st;
    if (i == 1 && j == 1) return 0.;
    st = s.c[i][j][0][0] * ukl[0][0];
    st += s.c[i][j][0][1] * (ukl[0][1] + ukl[1][0]);
    st += s.c[i][j][0][2] * (ukl[2][0] + ukl[0][2]);
    st += s.c[i][j][1][1] * ukl[1][1];
    st += s.c[i][j][1][2] * (ukl[1][2] + ukl[2][1]);
    st += s.c[i][j][2][2] * ukl[2][2];
    return st;
}

/* computes Stiffs = K u over all elements, accumulating shared nodes in the
   reference's element order (ex outer, ez inner) */
static void stiff_matrix(const orc_ctx *c, orc_vec *S, const orc_vec *p)
{
    int ex, ez, al, be, i, th, n = c->hpn, np1 = n + 1, nx = c->nx;
    size_t N = (size_t)c->nx * c->nz, q;
    orc_fik *F = (orc_fik *)xcalloc((size_t)np1 * np1, sizeof(orc_fik));
#pragma omp parallel for
    for (q = 0; q < N; q++) { S[q].u[0] = 0.; S[q].u[1] = 0.; S[q].u[2] = 0.; }
    for (ex = 0; ex < c->nex; ex++)
        for (ez = 0; ez < c->nez; ez++) {
            for (al = 0; al < np1; al++)
                for (be = 0; be < np1; be++) {
                    size_t idx = (size_t)(ez * n + be) * nx + ex * n + al;
                    orc_cmat mesh = c->mesh[idx];
                    real ukl[3][3], T[3][3];
                    memset(ukl, 0, sizeof ukl);
                    /* dUj_di_2D (sem2d_psvsh.c:1444-1482) */
                    for (th = 0; th < np1; th++) {
                        size_t k = (size_t)(ez * n + be) * nx + ex * n + th;
                        real d = c->dl_jei[al * np1 + th];
                        ukl[0][0] = ukl[0][0] + p[k].u[0] * d;
                        ukl[1][0] = ukl[1][0] + p[k].u[1] * d;
                        ukl[2][0] = ukl[2][0] + p[k].u[2] * d;
                    }
                    for (th = 0; th < np1; th++) {
                        size_t k = (size_t)(ez * n + th) * nx + ex * n + al;
                        real d = c->dl_jei[be * np1 + th];
                        ukl[0][2] = ukl[0][2] + p[k].u[0] * d;
                        ukl[1][2] = ukl[1][2] + p[k].u[1] * d;
                        ukl[2][2] = ukl[2][2] + p[k].u[2] * d;
                    }
                    for (i = 0; i < 3; i++) {
                        T[i][0] = stress_ij(i, 0, ukl, mesh);
                        T[i][1] = 0.;
                        T[i][2] = stress_ij(i, 2, ukl, mesh);
                    }
                    for (i = 0; i < 3; i++) {          /* F_ik (:1323-1336) */
                        F[be * np1 + al].F0[i] = T[i][0] * c->jei;
                        F[be * np1 + al].F2[i] = T[i][2] * c->jei;
                    }
                }
            for (al = 0; al < np1; al++)               /* Grid_Stiff (:1274-1313) */
                for (be = 0; be < np1; be++) {
                    size_t idx = (size_t)(ez * n + be) * nx + ex * n + al;
                    for (i = 0; i < 3; i++) {
                        real it0 = 0., it2 = 0.;
                        int a;
                        for (a = 0; a < np1; a++) it0 = it0 + F[be * np1 + a].F0[i] * c->dl_gllw[a * np1 + al];
                        for (a = 0; a < np1; a++) it2 = it2 + F[a * np1 + al].F2[i] * c->dl_gllw[a * np1 + be];
                        S[idx].u[i] += (it0 * c->gllw[be] + it2 * c->gllw[al]) * c->je;
                    }
                }
        }
    free(F);
}

/* ------------------------------------------------------------------------ */
/* K u for the fast FP64 path: the same element arithmetic as stiff_matrix,   */
/* the elements swept in the four colours (ex%2, ez%2) -- no two elements of   */
/* a colour share a node, so a colour's elements run in parallel without       */
/* atomics -- and the material read through a per-node record id.  Only the   */
/* order in which a shared node sums its element contributions differs from    */
/* stiff_matrix (colour order instead of ex-outer, ez-inner).                  */
/* ------------------------------------------------------------------------ */
static real stress_p(int i, int j, real ukl[3][3], const orc_cmat *s)
{
    real st;
    if (i == 1 && j == 1) return 0.;
    st = s->c[i][j][0][0] * ukl[0][0];
    st += s->c[i][j][0][1] * (ukl[0][1] + ukl[1][0]);
    st += s->c[i][j][0][2] * (ukl[2][0] + ukl[0][2]);
    st += s->c[i][j][1][1] * ukl[1][1];
    st += s->c[i][j][1][2] * (ukl[1][2] + ukl[2][1]);
    st += s->c[i][j][2][2] * ukl[2][2];
    return st;
}

static void build_record_table(orc_ctx *c)
{
    size_t N = (size_t)c->nx * c->nz, q;
    int cap = 16, k;
    if (c->rid) return;
    c->rid = (int *)xcalloc(N, sizeof(int));
    c->rtab = (orc_cmat *)xcalloc(cap, sizeof(orc_cmat));
    c->nrtab = 0;
    for (q = 0; q < N; q++) {
        int found = -1;
        if (q > 0 && memcmp(&c->rtab[c->rid[q - 1]], &c->mesh[q], sizeof(orc_cmat)) == 0) found = c->rid[q - 1];
        for (k = 0; found < 0 && k < c->nrtab; k++)
            if (memcmp(&c->rtab[k], &c->mesh[q], sizeof(orc_cmat)) == 0) found = k;
        if (found < 0) {
            if (c->nrtab == cap) { cap *= 2; c->rtab = (orc_cmat *)realloc(c->rtab, cap * sizeof(orc_cmat)); }
            c->rtab[c->nrtab] = c->mesh[q];
            found = c->nrtab++;
        }
        c->rid[q] = found;
    }
}

static void stiff_fast(const orc_ctx *c, orc_vec *S, const orc_vec *p)
{
    int n = c->hpn, np1 = n + 1, nx = c->nx, cx, cz;
    long N = (long)c->nx * c->nz, q;
#pragma omp parallel for schedule(static)
    for (q = 0; q < N; q++) { S[q].u[0] = 0.; S[q].u[1] = 0.; S[q].u[2] = 0.; }
    for (cx = 0; cx < 2; cx++)
        for (cz = 0; cz < 2; cz++) {
            int nexc = (c->nex - cx + 1) / 2, nezc = (c->nez - cz + 1) / 2;
            long e;
#pragma omp parallel for schedule(dynamic, 64)
            for (e = 0; e < (long)nexc * nezc; e++) {
                int ex = cx + 2 * (int)(e / nezc), ez = cz + 2 * (int)(e % nezc), al, be, i, th;
                orc_fik F[SEMSKS_ORC_MAXNP * SEMSKS_ORC_MAXNP];
                for (al = 0; al < np1; al++)
                    for (be = 0; be < np1; be++) {
                        size_t idx = (size_t)(ez * n + be) * nx + ex * n + al;
                        const orc_cmat *mesh = &c->rtab[c->rid[idx]];
                        real ukl[3][3], T[3][3];
                        memset(ukl, 0, sizeof ukl);
                        for (th = 0; th < np1; th++) {
                            size_t k = (size_t)(ez * n + be) * nx + ex * n + th;
                            real d = c->dl_jei[al * np1 + th];
                            ukl[0][0] = ukl[0][0] + p[k].u[0] * d;
                            ukl[1][0] = ukl[1][0] + p[k].u[1] * d;
                            ukl[2][0] = ukl[2][0] + p[k].u[2] * d;
                        }
                        for (th = 0; th < np1; th++) {
                            size_t k = (size_t)(ez * n + th) * nx + ex * n + al;
                            real d = c->dl_jei[be * np1 + th];
                            ukl[0][2] = ukl[0][2] + p[k].u[0] * d;
                            ukl[1][2] = ukl[1][2] + p[k].u[1] * d;
                            ukl[2][2] = ukl[2][2] + p[k].u[2] * d;
                        }
                        for (i = 0; i < 3; i++) {
                            T[i][0] = stress_p(i, 0, ukl, mesh);
                            T[i][2] = stress_p(i, 2, ukl, mesh);
                            F[be * np1 + al].F0[i] = T[i][0] * c->jei;
                            F[be * np1 + al].F2[i] = T[i][2] * c->jei;
                        }
                    }
                for (al = 0; al < np1; al++)
                    for (be = 0; be < np1; be++) {
                        size_t idx = (size_t)(ez * n + be) * nx + ex * n + al;
                        for (i = 0; i < 3; i++) {
                            real it0 = 0., it2 = 0.;
                            int a;
                            for (a = 0; a < np1; a++) it0 = it0 + F[be * np1 + a].F0[i] * c->dl_gllw[a * np1 + al];
                            for (a = 0; a < np1; a++) it2 = it2 + F[a * np1 + al].F2[i] * c->dl_gllw[a * np1 + be];
                            S[idx].u[i] += (it0 * c->gllw[be] + it2 * c->gllw[al]) * c->je;
                        }
                    }
            }
        }
}

/* ------------------------------------------------------------------------ */
/* explicit update + absorbing edges                                          */
/* ------------------------------------------------------------------------ */
/* Time_diff_hybrid (sem2d_psvsh.c:1004-1029) */
static void time_diff(const orc_ctx *c, orc_vec *p1, const orc_vec *p, const orc_vec *p0, const orc_vec *G)
{
    int k;
    long i, N = (long)c->nx * c->nz;
    for (k = 0; k < 3; k++) {
#pragma omp parallel for
        for (i = 0; i < N; i++)
            p1[i].u[k] = (real)((-c->mdt2[i] * G[i].u[k]) + 2. * p[i].u[k] - p0[i].u[k]);
    }
}

static void absorb_node(const orc_ctx *c, long gn, orc_vec *p1, const orc_vec *p, const orc_vec *p0, const orc_vec *G, real dt2)
{
    int k;
    for (k = 0; k < 3; k++) {
        real coe = c->mass[gn] + c->ab[gn].u[k];
        real coe1 = c->mass[gn] - c->ab[gn].u[k];
        p1[gn].u[k] = (real)((-dt2 * G[gn].u[k] + 2. * c->mass[gn] * p[gn].u[k] - coe1 * p0[gn].u[k]) / coe);
    }
}

/* Absorbs (sem2d_psvsh.c:1834-1889) */
static void absorbs(const orc_ctx *c, orc_vec *p1, const orc_vec *p, const orc_vec *p0, const orc_vec *G)
{
    long i, nx = c->nx, nz = c->nz;
    real dt2 = (real)c->sdt * (real)c->sdt;
    for (i = 0; i < nz; i++) absorb_node(c, i * nx, p1, p, p0, G, dt2);
    for (i = 0; i < nx; i++) absorb_node(c, i, p1, p, p0, G, dt2);
    for (i = 0; i < nz; i++) absorb_node(c, i * nx + nx - 1, p1, p, p0, G, dt2);
    /* the top-row strip ix = 1..3*hpn; the reference's loop does not stop at
       the row end, which only matters (as an overrun) for nx <= 3*hpn+1,
       i.e. meshes of fewer than 3 element columns: clipped here as in the
       product */
    for (i = 1; i < 3 * c->hpn + 1 && i < nx; i++) absorb_node(c, (nz - 1) * nx + i, p1, p, p0, G, dt2);
}

/* atten_wave2 (sem2d_psvsh.c:1780-1823) with atten_taper (:1725-1738) */
static void atten_wave2(const orc_ctx *c, orc_vec *p)
{
    int i, ix, iz, nh = c->neh * c->hpn, nx = c->nx, nz = c->nz, at = c->atten;
    for (i = 0; i < 3; i++)
        for (ix = 0; ix < nx; ix++)
            for (iz = 0; iz < nz; iz++) {
                long o = (long)iz * nx + ix;
                if (at > 0 && iz < nh) p[o].u[i] *= c->taper[nh - iz - 1];
                if (at > 1 && ix < nh && iz >= nh) p[o].u[i] *= c->taper[nh - ix - 1];
                if (at > 2 && ix >= nx - nh && iz >= nh) p[o].u[i] *= c->taper[ix - nx + nh];
                if (at > 3 && iz >= nz - nh && ix < nx - nh && ix >= nh) p[o].u[i] *= c->taper[iz - nz + nh];
            }
}

/* ------------------------------------------------------------------------ */
/* GRT boundary: interpolation (interpl.c) and injection                     */
/* ------------------------------------------------------------------------ */
typedef struct { FILE *f; int n; float **sy; } orc_stream;   /* sy[i][slot] */

/* interp_const + interpl (interpl.c:16-91).  The control flow (t, record
   schedule, slot rotation, fractional offset) stays in float as in the
   reference; the weights and the weighted sum are computed in real. */
static void interpl(orc_stream *s, int it, float rdt, int *ns, real *out)
{
    float nst0 = (float)it, t = nst0 * rdt, d1;
    int ns0 = (int)(t + 0.00001), j, i;
    real w[4], dty1, dty2, dty3;
    if (ns0 > (int)((nst0 - 1) * rdt + 0.00001) || nst0 == 0.) {
        int r = ns[0];
        for (j = 0; j < 3; j++) ns[j] = ns[j + 1];
        ns[3] = r;
        {
            float *g = (float *)xcalloc((size_t)s->n, sizeof(float));
            size_t got = fread(g, sizeof(float), (size_t)s->n, s->f);
            (void)got;
            for (i = 0; i < s->n; i++) s->sy[i][ns[3]] = g[i];
            free(g);
        }
    }
    d1 = t - (int)(t + 0.00001);
    dty1 = d1;
    dty2 = dty1 * dty1;
    dty3 = dty1 * dty2;
    w[0] = (real)(-0.5 * dty3 + dty2 - 0.5 * dty1);
    w[1] = (real)(3.0 * dty3 - 2.5 * dty2 + 1.0);
    w[2] = (real)(-3.0 * dty3 + 2.0 * dty2 + 0.5 * dty1);
    w[3] = (real)(0.5 * (dty3 - dty2));
    for (i = 0; i < s->n; i++) {
        real acc = 0.0;
        for (j = 0; j < 4; j++) acc += w[j] * (real)s->sy[i][ns[j]];
        out[i] = acc;
    }
}

/* non-finite / |g|>=10 repair (sem2d_psvsh.c:433-443).  The reference reads
   g[-1] and g[n] at the ends (out of bounds); here they read as 0. */
static void guard(real *g, int n)
{
    int i;
    for (i = 0; i < n; i++)
        if (!(g[i] < 10 && g[i] > -10)) {
            real l = i > 0 ? g[i - 1] : 0, r = i + 1 < n ? g[i + 1] : 0;
            g[i] = (l + r) / 2;
        }
}

/* GRT_Interface at it==0 (sem2d_psvsh.c:1953-2005) */
static void grt_interface0(const orc_ctx *c, orc_vec *p, const real *g, const real *gv)
{
    int i, m, nbl = c->neh * c->hpn, nx = c->nx, nz = c->nz, n = c->hpn;
    int sh = (c->variant == 0);
    for (i = 0; i < nx; i++) { long q = (long)nbl * nx + i; p[q].u[0] = 0; if (sh) p[q].u[1] = 0; p[q].u[2] = 0; }
    for (i = 0; i < nz; i++) { long q = (long)i * nx + nbl; p[q].u[0] = 0; if (sh) p[q].u[1] = 0; p[q].u[2] = 0; }
    for (i = c->neh; i < c->nex; i++)
        for (m = 0; m <= n; m++) {
            long q = (long)nbl * nx + i * n + m;
            p[q].u[0] = g[i * n + m];
            if (sh) p[q].u[1] = gv[i * n + m];
            p[q].u[2] = -g[nx + i * n + m];
        }
    for (i = c->neh; i < c->nez; i++)
        for (m = 0; m <= n; m++) {
            int idz = i * n + m;
            long q = (long)idz * nx + nbl;
            p[q].u[0] = g[2 * nx + idz];
            if (sh) p[q].u[1] = gv[nx + idz];
            p[q].u[2] = -g[2 * nx + nz + idz];
        }
}

/* GRT_Interface4 (sem2d_psvsh.c:2076-2105) and GRT_Interface3
   (sem2dMP20140816.c:2001-2037) */
static void grt_interface(const orc_ctx *c, orc_vec *p, const orc_vec *p0,
                          const real *g, const real *g0, const real *gv, const real *gv0)
{
    int i, nbl = c->neh * c->hpn, nx = c->nx, nz = c->nz;
    int sh = (c->variant == 0);
    for (i = nbl; i < nx; i++) {
        long q = (long)nbl * nx + i;
        if (c->variant == 2) {
            p[q].u[0] += -p0[q].u[0] + g[i];
            p[q].u[2] += -p0[q].u[2] - g[i + nx];
        } else {
            p[q].u[0] += -g0[i] + g[i];
            if (sh) p[q].u[1] += -gv0[i] + gv[i];
            p[q].u[2] += g0[i + nx] - g[i + nx];
        }
    }
    for (i = nbl + 1; i < nz; i++) {
        long q = (long)i * nx + nbl;
        if (c->variant == 2) {
            p[q].u[0] += -p0[q].u[0] + g[2 * nx + i];
            p[q].u[2] += -p0[q].u[2] - g[2 * nx + nz + i];
        } else {
            p[q].u[0] += -g0[2 * nx + i] + g[2 * nx + i];
            if (sh) p[q].u[1] += -gv0[nx + i] + gv[nx + i];
            p[q].u[2] += g0[2 * nx + nz + i] - g[2 * nx + nz + i];
        }
    }
}

/* ------------------------------------------------------------------------ */
/* setup                                                                     */
/* ------------------------------------------------------------------------ */
ORC_EXPORT void orc_destroy(orc_ctx *c)
{
    if (!c) return;
    free(c->gllp); free(c->gllw); free(c->dl); free(c->dl_jei); free(c->dl_gllw);
    free(c->mesh); free(c->mass); free(c->mdt2); free(c->ab);
    free(c->ix1); free(c->iz1); free(c->taper); free(c->rid); free(c->rtab);
    free(c);
}

/* parses the parameters exactly as sem2d_psvsh.c:154-205 and runs the
   one-off setup (Parts 1-5 and 6.0 of the reference main) */
ORC_EXPORT int orc_create(int argc, char **argv, orc_ctx **out)
{
    orc_pars P;
    orc_ctx *c = (orc_ctx *)xcalloc(1, sizeof(orc_ctx));
    int rc, i, j, np1, variant = 0;
    *out = NULL;
    c->flat = 1; c->nfil = 1; c->hpn = 5; c->absorb = 1; c->readpars = 1; c->neh = 1;
    c->atten = 0; c->ass = 0.93f; c->check = 1; c->xtrace = 20; c->ztrace = 1; c->idxplot = 20; c->idzplot = 1;
    pars_init(&P, argc, argv);
    gp_int(&P, "flat", &c->flat);
    gp_int(&P, "solver", &variant);     /* 0 sem2d_psvsh, 1 sem2d, 2 sem2dMP (extension key) */
    c->variant = variant;
    mst_str(&P, "greenfile", c->greenfile, sizeof c->greenfile);
    if (variant == 0) mst_str(&P, "shgreenfile", c->shgreenfile, sizeof c->shgreenfile);
    mst_str(&P, "raymodel", c->raymodel, sizeof c->raymodel);
    mst_str(&P, "semmodel", c->semmodel, sizeof c->semmodel);
    mst_int(&P, "nt", &c->nt);
    mst_flt(&P, "dp", &c->dp);
    mst_int(&P, "nex", &c->nex);
    mst_flt(&P, "eh", &c->eh);
    mst_flt(&P, "xmin", &c->xmin);
    mst_int(&P, "nst", &c->nst);
    mst_flt(&P, "sdt", &c->sdt);
    gp_int(&P, "nfil", &c->nfil);
    gp_int(&P, "hpn", &c->hpn);
    gp_int(&P, "absorb", &c->absorb);
    gp_int(&P, "readpars", &c->readpars);
    gp_int(&P, "neh", &c->neh);
    gp_int(&P, "atten", &c->atten);
    gp_flt(&P, "ass", &c->ass);
    gp_int(&P, "check", &c->check);
    gp_int(&P, "output_vd", &c->output_vd);
    gp_int(&P, "xtrace", &c->xtrace);
    mst_flt(&P, "xbeg", &c->xbeg);
    gp_flt(&P, "zbeg", &c->zbeg);
    gp_int(&P, "get_rec", &c->get_rec);
    gp_int(&P, "plot_trace", &c->plot_trace);
    if (c->get_rec == 1) mst_str(&P, "rec_file", c->rec_file, sizeof c->rec_file);
    mst_int(&P, "idxplot", &c->idxplot);
    mst_int(&P, "msource", &c->msource);
    gp_str(&P, "semdip", c->semdip, sizeof c->semdip);
    rc = P.err;
    if (rc) orc_fail(rc, P.errarg, NULL);
    else if (P.stop) rc = orc_fail(ORC_ESTOP, "stop due to STOP in input", NULL);
    pars_free(&P);
    if (rc) { orc_destroy(c); return rc; }
    if (c->msource != 0) { orc_destroy(c); return orc_fail(ORC_EINVAL, "only msource=0 (GRT boundary) is restated", NULL); }
    if (c->hpn < 3 || c->hpn > 6) { orc_destroy(c); return orc_fail(ORC_EINVAL, "hpn outside 3..6", NULL); }

    np1 = c->hpn + 1;
    c->gllp = (real *)xcalloc(np1, sizeof(real));
    c->gllw = (real *)xcalloc(np1, sizeof(real));
    c->dl = (real *)xcalloc((size_t)np1 * np1, sizeof(real));
    c->dl_jei = (real *)xcalloc((size_t)np1 * np1, sizeof(real));
    c->dl_gllw = (real *)xcalloc((size_t)np1 * np1, sizeof(real));
    orc_gll(c->hpn, c->gllp, c->gllw);
    orc_dlag(c->hpn, c->gllp, c->dl);
    if (c->atten) {
        int nh = c->neh * c->hpn;
        double al = 1.0 / nh * log10((double)c->ass) / log10(2.73);
        c->taper = (float *)xcalloc(nh, sizeof(float));
        for (i = 0; i < nh; i++) c->taper[i] = (float)exp((double)(al * i));
    }
    rc = discret(c);
    if (rc) { orc_destroy(c); return rc; }
    /* uniform square elements: Je = eh^2/4, Je^-1 = 2/eh, Jb = eh/2
       (sem2d_psvsh.c:279-292; All_Je only confirms uniformity) */
    c->je = (real)((real)c->eh * (real)c->eh / 4.);
    c->jei = (real)(2. / c->eh);
    c->jb = (real)(c->eh / 2.);
    for (i = 0; i < np1; i++)
        for (j = 0; j < np1; j++) {
            c->dl_jei[j * np1 + i] = c->dl[j * np1 + i] * c->jei;
            c->dl_gllw[j * np1 + i] = c->dl[j * np1 + i] * c->gllw[j];
        }
    {
        size_t N = (size_t)c->nx * c->nz;
        c->mass = (real *)xcalloc(N + 1, sizeof(real));
        c->mdt2 = (real *)xcalloc(N + 1, sizeof(real));
        c->ab = (orc_vec *)xcalloc(N + 1, sizeof(orc_vec));
        mass_matrix(c);
        for (i = 0; i < (long)N; i++) c->mdt2[i] = (real)c->sdt * (real)c->sdt / c->mass[i];
        if (c->absorb > 0) absorb_matrix(c);
    }
    /* receivers (sem2d_psvsh.c:330-370) */
    {
        int ixbeg = (int)(c->xbeg / c->h - 1), izbeg = (int)(c->zbeg / c->h);
        izbeg += c->nz - 1;
        if (c->zbeg == 0.) izbeg = c->nz - 1;
        if (izbeg > c->nz - 1) izbeg = c->nz - 1;
        if (c->get_rec == 0) {
            c->ntrace = c->xtrace * c->ztrace;
            c->ix1 = (int *)xcalloc(c->ntrace + 1, sizeof(int));
            c->iz1 = (int *)xcalloc(c->ntrace + 1, sizeof(int));
            for (i = 0; i < c->xtrace; i++)
                for (j = 0; j < c->ztrace; j++) {
                    c->ix1[i + j * c->xtrace] = ixbeg + c->idxplot * i;
                    c->iz1[i + j * c->xtrace] = izbeg - c->idzplot * j;
                }
        } else {
            FILE *f = fopen(c->rec_file, "r");
            int n;
            if (!f) { orc_destroy(c); return orc_fail(ORC_EIO, "Error opening file", c->rec_file); }
            if (fscanf(f, "%d\n", &n) != 1) n = 0;
            c->xtrace = n; c->ztrace = 1; c->ntrace = n;
            c->ix1 = (int *)xcalloc(n + 1, sizeof(int));
            c->iz1 = (int *)xcalloc(n + 1, sizeof(int));
            for (i = 0; i < n; i++) {
                float rx = 0.f;
                if (fscanf(f, "%f", &rx) != 1) break;
                c->ix1[i] = (int)(rx / c->h);
                c->iz1[i] = izbeg;
            }
            fclose(f);
        }
    }
    c->rdt = c->sdt / c->dp;
    c->nt_sem = (int)((float)((c->nt - 3) * c->dp / c->sdt));
    *out = c;
    return ORC_OK;
}

/* ------------------------------------------------------------------------ */
/* accessors for tests                                                       */
/* ------------------------------------------------------------------------ */
ORC_EXPORT void orc_dims(const orc_ctx *c, int *d)
{
    d[0] = c->nex; d[1] = c->nez; d[2] = c->nx; d[3] = c->nz; d[4] = c->hpn;
    d[5] = c->ntrace; d[6] = c->nt_sem; d[7] = c->nst; d[8] = c->neh; d[9] = c->variant;
}
ORC_EXPORT void orc_scalars(const orc_ctx *c, double *s)
{
    s[0] = c->eh; s[1] = c->h; s[2] = c->je; s[3] = c->jei; s[4] = c->jb; s[5] = c->sdt; s[6] = c->rdt; s[7] = c->dp;
}
ORC_EXPORT void orc_tables(const orc_ctx *c, real *gllp, real *gllw, real *dl)
{
    int n = c->hpn + 1;
    memcpy(gllp, c->gllp, n * sizeof(real));
    memcpy(gllw, c->gllw, n * sizeof(real));
    memcpy(dl, c->dl, (size_t)n * n * sizeof(real));
}
/* per node: rho and the 15 distinct constants entering 2-D stress, in the
   Voigt order (0,0)(0,2)(0,3)(0,4)(0,5)(2,2)(2,3)(2,4)(2,5)(3,3)(3,4)(3,5)(4,4)(4,5)(5,5) */
ORC_EXPORT void orc_material(const orc_ctx *c, real *out16)
{
    static const int pairs[15][4] = {
        {0,0,0,0},{0,0,2,2},{0,0,1,2},{0,0,0,2},{0,0,0,1},
        {2,2,2,2},{2,2,1,2},{2,2,0,2},{2,2,0,1},
        {1,2,1,2},{1,2,0,2},{1,2,0,1},
        {0,2,0,2},{0,2,0,1},
        {0,1,0,1} };
    size_t N = (size_t)c->nx * c->nz, q;
    int k;
    for (q = 0; q < N; q++) {
        const orc_cmat *m = &c->mesh[q];
        out16[q * 16] = m->rho;
        for (k = 0; k < 15; k++) out16[q * 16 + 1 + k] = m->c[pairs[k][0]][pairs[k][1]][pairs[k][2]][pairs[k][3]];
    }
}
ORC_EXPORT void orc_mass(const orc_ctx *c, real *mass, real *mdt2, real *ab)
{
    size_t N = (size_t)c->nx * c->nz;
    memcpy(mass, c->mass, N * sizeof(real));
    memcpy(mdt2, c->mdt2, N * sizeof(real));
    memcpy(ab, c->ab, N * 3 * sizeof(real));
}
ORC_EXPORT void orc_receivers(const orc_ctx *c, int *ix, int *iz)
{
    memcpy(ix, c->ix1, c->ntrace * sizeof(int));
    memcpy(iz, c->iz1, c->ntrace * sizeof(int));
}
/* K u for an arbitrary field (AoS, 3 components per node) */
ORC_EXPORT void orc_stiff(const orc_ctx *c, const real *u, real *f)
{
    stiff_matrix(c, (orc_vec *)f, (const orc_vec *)u);
}
/* the edge sponge atten_wave2 applied once to a field (AoS), in place; a
   no-op when atten = 0 */
ORC_EXPORT void orc_sponge(const orc_ctx *c, real *u)
{
    if (c->atten > 0) atten_wave2(c, (orc_vec *)u);
}
/* the sponge factors atten_taper writes (neh*hpn values; none when atten = 0) */
ORC_EXPORT int orc_taper(const orc_ctx *c, float *out)
{
    int i, nh = c->neh * c->hpn;
    if (!c->atten) return 0;
    for (i = 0; i < nh; i++) out[i] = c->taper[i];
    return nh;
}
/* one explicit update from (u, u_prev, K u) including the absorbing rows */
ORC_EXPORT void orc_update(const orc_ctx *c, const real *u, const real *u0, const real *G, real *u1)
{
    time_diff(c, (orc_vec *)u1, (const orc_vec *)u, (const orc_vec *)u0, (const orc_vec *)G);
    absorbs(c, (orc_vec *)u1, (const orc_vec *)u, (const orc_vec *)u0, (const orc_vec *)G);
}

/* the interpolated, guarded boundary vectors for every step it < min(nsteps,
   nt_sem), exactly as the time loop forms them.  g: [nsteps][2(nx+nz)],
   gv: [nsteps][nx+nz] (left zero for the P-SV solvers) */
ORC_EXPORT int orc_boundary_series(const orc_ctx *c, int nsteps, real *g, real *gv)
{
    orc_stream s[2];
    int ns[4] = {0, 1, 2, 3}, it, i, k, nstream = (c->variant == 0) ? 2 : 1;
    int mt = 2 * (c->nx + c->nz), sht = c->nx + c->nz;
    const char *fn[2];
    fn[0] = c->greenfile; fn[1] = c->shgreenfile;
    for (k = 0; k < nstream; k++) {
        s[k].n = k ? sht : mt;
        s[k].f = fopen(fn[k], "rb");
        if (!s[k].f) { if (k) fclose(s[0].f); return orc_fail(ORC_EIO, "Error opening file", fn[k]); }
        s[k].sy = (float **)xcalloc(s[k].n, sizeof(float *));
        for (i = 0; i < s[k].n; i++) s[k].sy[i] = (float *)xcalloc(4, sizeof(float));
    }
    for (it = 0; it < nsteps; it++) {
        real *gi = g + (size_t)it * mt, *gvi = gv + (size_t)it * sht;
        if (it < c->nt_sem) {
            interpl(&s[0], it, c->rdt, ns, gi);
            if (nstream == 2) interpl(&s[1], it, c->rdt, ns, gvi);
            else memset(gvi, 0, sht * sizeof(real));
            guard(gi, mt);
            if (nstream == 2) guard(gvi, sht);
        } else {
            memcpy(gi, gi - mt, mt * sizeof(real));
            memcpy(gvi, gvi - sht, sht * sizeof(real));
        }
    }
    for (k = 0; k < nstream; k++) {
        for (i = 0; i < s[k].n; i++) free(s[k].sy[i]);
        free(s[k].sy);
        fclose(s[k].f);
    }
    return ORC_OK;
}

/* the raw interpolation (no guard) of one or two in-memory record streams,
   stepping the slot ring exactly as the solvers do: one stream
   (sem2d.c:410-413) or two sharing ns (sem2d_psvsh.c:430-431) */
ORC_EXPORT int orc_interp_series(const float *rec0, int w0, const float *rec1, int w1, int nrec, int nsteps,
                                 float rdt, real *out0, real *out1)
{
    orc_stream s[2];
    int ns[4] = {0, 1, 2, 3}, it, i, k, nstream = rec1 ? 2 : 1;
    const float *rec[2];
    int w[2];
    rec[0] = rec0; rec[1] = rec1; w[0] = w0; w[1] = w1;
    for (k = 0; k < nstream; k++) {
        s[k].n = w[k];
        s[k].f = fmemopen((void *)rec[k], (size_t)nrec * w[k] * sizeof(float), "rb");
        if (!s[k].f) return ORC_EIO;
        s[k].sy = (float **)xcalloc(w[k], sizeof(float *));
        for (i = 0; i < w[k]; i++) s[k].sy[i] = (float *)xcalloc(4, sizeof(float));
    }
    for (it = 0; it < nsteps; it++) {
        interpl(&s[0], it, rdt, ns, out0 + (size_t)it * w0);
        if (nstream == 2) interpl(&s[1], it, rdt, ns, out1 + (size_t)it * w1);
    }
    for (k = 0; k < nstream; k++) {
        for (i = 0; i < w[k]; i++) free(s[k].sy[i]);
        free(s[k].sy);
        fclose(s[k].f);
    }
    return ORC_OK;
}

/* the guard on one vector (sem2d_psvsh.c:433-443) */
ORC_EXPORT void orc_guard(real *g, int n) { guard(g, n); }

/* read_model restated: returns jo and nrec, fills the (possibly flattened)
   layer arrays (capacity cap) */
ORC_EXPORT int orc_read_raymodel(const char *path, int flat, int cap, float *cc, float *ss, float *dd, float *tth,
                                 int *jo, int *nrec)
{
    orc_raymodel m;
    int rc = read_raymodel(path, flat, &m), j;
    if (rc) return rc;
    *jo = m.jo;
    *nrec = m.na[m.nen[0] - 2];
    for (j = 0; j < m.jo && j < cap; j++) { cc[j] = m.cc[j]; ss[j] = m.ss[j]; dd[j] = m.dd[j]; tth[j] = m.tth[j]; }
    free_raymodel(&m);
    return ORC_OK;
}

/* ------------------------------------------------------------------------ */
/* the time loop (sem2d_psvsh.c:418-512)                                     */
/* traces: [nsteps][3][ntrace] in the order u, w(sign-flipped), v, i.e. the  */
/* record written to seis.u, seis.w, seis.v by kirrecord_zhl                 */
/* ------------------------------------------------------------------------ */
/* optional per-step wall clock of orc_run (the timed CPU baseline): when
   set, step it of the next runs stores CLOCK_MONOTONIC seconds at its end
   into t[it] for it < n */
static double *orc_clock_t;
static int orc_clock_n;
ORC_EXPORT void orc_set_step_clock(double *t, int n) { orc_clock_t = t; orc_clock_n = t ? n : 0; }

static double orc_now(void)
{
    struct timespec ts;
    clock_gettime(CLOCK_MONOTONIC, &ts);
    return (double)ts.tv_sec + 1e-9 * (double)ts.tv_nsec;
}

static int run_loop(orc_ctx *c, int nsteps, float *traces, real *final_u, int fast);

ORC_EXPORT int orc_run(orc_ctx *c, int nsteps, float *traces, real *final_u)
{
    return run_loop(c, nsteps, traces, final_u, 0);
}

/* orc_run with the colour-parallel element force (stiff_fast): the FP64
   oracle at sizes the reference-order sweep cannot reach in reasonable time */
ORC_EXPORT int orc_run_fast(orc_ctx *c, int nsteps, float *traces, real *final_u)
{
    build_record_table(c);
    return run_loop(c, nsteps, traces, final_u, 1);
}

static int run_loop(orc_ctx *c, int nsteps, float *traces, real *final_u, int fast)
{
    size_t N = (size_t)c->nx * c->nz;
    orc_vec *uvw0 = (orc_vec *)xcalloc(N + 1, sizeof(orc_vec));
    orc_vec *uvw = (orc_vec *)xcalloc(N + 1, sizeof(orc_vec));
    orc_vec *uvw1 = (orc_vec *)xcalloc(N + 1, sizeof(orc_vec));
    orc_vec *GS = (orc_vec *)xcalloc(N + 1, sizeof(orc_vec));
    int mt = 2 * (c->nx + c->nz), sht = c->nx + c->nz, it, i, k;
    int nstream = (c->variant == 0) ? 2 : 1;
    real *g = (real *)xcalloc(mt, sizeof(real)), *g0 = (real *)xcalloc(mt, sizeof(real));
    real *gv = (real *)xcalloc(sht, sizeof(real)), *gv0 = (real *)xcalloc(sht, sizeof(real));
    orc_stream s[2];
    int ns[4] = {0, 1, 2, 3};
    const char *fn[2];
    fn[0] = c->greenfile; fn[1] = c->shgreenfile;
    for (k = 0; k < nstream; k++) {
        s[k].n = k ? sht : mt;
        s[k].f = fopen(fn[k], "rb");
        if (!s[k].f) return orc_fail(ORC_EIO, "Error opening file", fn[k]);
        s[k].sy = (float **)xcalloc(s[k].n, sizeof(float *));
        for (i = 0; i < s[k].n; i++) s[k].sy[i] = (float *)xcalloc(4, sizeof(float));
    }
    for (it = 0; it < nsteps; it++) {
        if (it > 0) { memcpy(g0, g, mt * sizeof(real)); memcpy(gv0, gv, sht * sizeof(real)); }
        if (it < c->nt_sem) {
            interpl(&s[0], it, c->rdt, ns, g);
            if (nstream == 2) interpl(&s[1], it, c->rdt, ns, gv);
            guard(g, mt);
            if (nstream == 2) guard(gv, sht);
        }
        if (it == 0) {
            memset(uvw, 0, (N + 1) * sizeof(orc_vec));
            memset(uvw0, 0, (N + 1) * sizeof(orc_vec));
            grt_interface0(c, uvw, g, gv);
        } else {
            orc_vec *t;
            if (it < c->nt_sem) grt_interface(c, uvw, uvw0, g, g0, gv, gv0);
            if (fast) stiff_fast(c, GS, uvw);
            else stiff_matrix(c, GS, uvw);
            time_diff(c, uvw1, uvw, uvw0, GS);
            /* Time_diff_hybrid calls Absorbs unconditionally; with absorb=0
               the reference reads an unallocated Ab_matrix, here A = 0 */
            absorbs(c, uvw1, uvw, uvw0, GS);
            if (c->atten > 0) atten_wave2(c, uvw1);
            /* Copy_Displacement(uvw,uvw0); Copy_Displacement(uvw1,uvw) */
            t = uvw0; uvw0 = uvw; uvw = uvw1; uvw1 = t;
        }
        if (traces) {               /* kirrecord_zhl (sem2d_psvsh.c:2116-2165) */
            float *tr = traces + (size_t)it * 3 * c->ntrace;
            for (i = 0; i < c->ntrace; i++) {
                size_t o = (size_t)c->iz1[i] * c->nx + c->ix1[i];
                tr[i] = (float)uvw[o].u[0];
                tr[c->ntrace + i] = (float)-uvw[o].u[2];
                tr[2 * c->ntrace + i] = (float)uvw[o].u[1];
            }
        }
        if (it < orc_clock_n) orc_clock_t[it] = orc_now();
    }
    if (final_u) memcpy(final_u, uvw, N * 3 * sizeof(real));
    for (k = 0; k < nstream; k++) {
        for (i = 0; i < s[k].n; i++) free(s[k].sy[i]);
        free(s[k].sy);
        fclose(s[k].f);
    }
    free(uvw0); free(uvw); free(uvw1); free(GS); free(g); free(g0); free(gv); free(gv0);
    return ORC_OK;
}

/* main()-equivalent: parse, set up, run nst steps, write seis.u/.w/.v into
   outdir when plot_trace=1.  Returns the reference's exit status. */
ORC_EXPORT int orc_main(int argc, char **argv, const char *outdir)
{
    orc_ctx *c;
    int rc = orc_create(argc, argv, &c);
    float *tr;
    if (rc) return rc;
    tr = (float *)xcalloc((size_t)c->nst * 3 * (c->ntrace ? c->ntrace : 1), sizeof(float));
    rc = orc_run(c, c->nst, tr, NULL);
    if (!rc && c->plot_trace) {
        const char *names[3] = {"seis.u", "seis.w", "seis.v"};
        int k, it;
        for (k = 0; k < 3; k++) {
            char path[2048];
            FILE *f;
            snprintf(path, sizeof path, "%s/%s", outdir ? outdir : ".", names[k]);
            f = fopen(path, "wb");
            if (!f) { rc = orc_fail(ORC_EIO, "Error opening file", path); break; }
            for (it = 0; it < c->nst; it++)
                fwrite(tr + ((size_t)it * 3 + k) * c->ntrace, sizeof(float), c->ntrace, f);
            fclose(f);
        }
    }
    free(tr);
    orc_destroy(c);
    return rc;
}
