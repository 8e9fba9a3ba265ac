/* semsks_internal.h -- declarations shared by the library's translation units. */
#ifndef SEMSKS_INTERNAL_H
#define SEMSKS_INTERNAL_H
#include "semsks.h"
#ifdef __cplusplus
extern "C" {
#endif
/* records an error text for semsks_last_error() and returns st */
semsks_status semsks_set_error(semsks_status st, const char *fmt, ...);

/* whether semsks_admit(e, p, -1, ...) would find a free slot and the HBM
   for p: SEMSKS_OK, SEMSKS_ENOSLOT or SEMSKS_ENOHBM, with no side effect.
   pending > 0: p's records are first generated on the device by
   semsks_planewave, which holds that many bytes until the admission */
semsks_status semsks_room(semsks_engine *e, const semsks_problem *p, size_t pending);

/* bytes semsks_planewave(e, ..., NULL, NULL) keeps on the device for nt
   records of a mesh of nex x nez elements of degree hpn */
size_t semsks_planewave_bytes(int nex, int nez, int hpn, int nt);

/* getpar semantics (semsyn/src/getpar/src/getpar.c) for the library's
   readers: open = setpar, semsks_par_* = getpar (1 if the key is present),
   semsks_mstpar_* = mstpar, close = endpar (EPAR for a missing mstpar key,
   ESTOP for STOP in the input) */
typedef struct semsks_pars semsks_pars;
semsks_pars *semsks_pars_open(int argc, char **argv);
int semsks_par_i(semsks_pars *q, const char *k, int *v);
int semsks_par_f(semsks_pars *q, const char *k, float *v);
int semsks_par_s(semsks_pars *q, const char *k, char *v, size_t cap);
void semsks_mstpar_i(semsks_pars *q, const char *k, int *v);
void semsks_mstpar_f(semsks_pars *q, const char *k, float *v);
void semsks_mstpar_s(semsks_pars *q, const char *k, char *v, size_t cap);
semsks_status semsks_pars_close(semsks_pars *q);

/* nez and eh of the SEM mesh from the raymodel thickness (readmodel.c:43-109,
   sem2d_psvsh.c:591-611) */
semsks_status semsks_raymodel_mesh(const char *fn, int flat, int nfil, float eh_in, int *nez, float *eh);
#ifdef __cplusplus
}
#endif
#endif
