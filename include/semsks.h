/*
 * semsks.h -- C ABI of the MI355X (gfx950) SEM time loop.
 *
 * Drop-in for the spectral-element solve of s-sks (semsyn/src/sem2d_psvsh.c,
 * sem2d.c, sem2dMP20140816.c).  The reference has no library API: it is a
 * process (`sem2d_psvsh par=run.par`, applications/test/prem_epi83.pl:81-82)
 * whose loop body is a set of K&R functions over globals.  Each entry point
 * below names the reference code it replaces.
 *
 * Conventions
 *   - Plain C types only; host buffers are owned by the caller, device memory
 *     by the library.  No entry point calls exit().
 *   - Fields are structure-of-arrays per component (u_x, u_y, u_z), node
 *     index iz*nx + ix with iz = 0 at the bottom of the box and iz = nz-1 at
 *     the free surface (sem2d_psvsh.c:1176,1214).
 *   - One host thread per engine; all device work of an engine is ordered on
 *     one HIP stream.
 */
#ifndef SEMSKS_H
#define SEMSKS_H

#include <stddef.h>
#include <stdint.h>

#ifdef __cplusplus
extern "C" {
#endif

/* Status codes.  The reference ends the process instead; the mapping is:
   EIO   <- open_file() exit(-1)          (readmodel.c:15-19)
   ENP   <- np > 1000 exit(-2)            (sem2d_psvsh.c:725-728)
   ERANGE<- lip() model x out of range    (sem2d_psvsh.c:835-839)
   EPAR  <- mstpar missing value exit(100)(getpar.c:398-434,715-725)
   ESTOP <- STOP in the input exit(101)   (getpar.c:378-390)            */
typedef enum {
    SEMSKS_OK = 0,
    SEMSKS_EINVAL = 1,
    SEMSKS_EIO = 2,
    SEMSKS_ENOHBM = 3,      /* device memory budget exceeded */
    SEMSKS_ENONFINITE = 4,  /* state became non-finite */
    SEMSKS_ERANGE = 5,
    SEMSKS_ENP = 6,
    SEMSKS_EDEVICE = 7,     /* HIP runtime error */
    SEMSKS_ENOSLOT = 8,     /* no free instance slot */
    SEMSKS_EPAR = 100,
    SEMSKS_ESTOP = 101
} semsks_status;

/* Which reference solver's boundary handling to reproduce. */
enum {
    SEMSKS_SOLVER_PSVSH = 0,  /* sem2d_psvsh.c: u,v,w injected, GRT_Interface4,
                                 P-SV and SH interpolation share one slot ring */
    SEMSKS_SOLVER_PSV = 1,    /* sem2d.c: u,w injected, GRT_Interface4 */
    SEMSKS_SOLVER_PSV_MP = 2  /* sem2dMP20140816.c: u,w injected, GRT_Interface3 */
};

/* problem flags */
enum {
    SEMSKS_F_NO_KAHAN = 1     /* plain FP32 displacement update, the reference's own arithmetic
                                 (par key kahan=0): no compensation words read or written */
};

/* Number of floats in one material record: rho followed by the 15 distinct
   elastic constants that enter 2-D stress, Voigt pairs
   (0,0)(0,2)(0,3)(0,4)(0,5)(2,2)(2,3)(2,4)(2,5)(3,3)(3,4)(3,5)(4,4)(4,5)(5,5)
   with Voigt 0=xx 1=yy 2=zz 3=yz 4=xz 5=xy (inv_index, sem2d_psvsh.c:1584-1597). */
#define SEMSKS_MAT_STRIDE 16

/* One SEM instance, fully described in memory (what main() of
   sem2d_psvsh.c:82-519 holds after Parts 1-6.0).  All pointers are host
   pointers owned by the caller; the library copies what it needs. */
typedef struct semsks_problem {
    int nex, nez;          /* elements in x and z */
    int hpn;               /* polynomial degree, 3..6 (NGLL = hpn+1) */
    int neh;               /* elements between edge and GRT interface */
    float eh;              /* element size in km after discret()'s adjustment */
    float sdt;             /* time step, s */
    int nst;               /* time levels recorded (steps run = nst-1) */
    int solver;            /* SEMSKS_SOLVER_* */
    int absorb;            /* Lysmer edges on (1) or A = 0 (0) */
    int flags;             /* SEMSKS_F_* */
    /* material: node -> record id, records of SEMSKS_MAT_STRIDE floats */
    int n_mat;
    const float *mat_table;     /* [n_mat][16] */
    const uint16_t *mat_id;     /* [nz][nx] */
    /* receivers (kirrecord_zhl node sampling) */
    int ntrace;
    const int *rec_ix, *rec_iz;
    /* GRT boundary records, greenfile layout (time-major, headerless):
       psv [nt][2(nx+nz)] = u_bot|w_bot|u_left|w_left (w positive down),
       sh  [nt][nx+nz]    = v_bot|v_left.  NULL means all zero. */
    int nt;
    float dp;
    const float *psv, *sh;
    /* exponential sponge on the edge strips (atten_wave2 / atten_taper,
       sem2d_psvsh.c:1725-1738,1780-1823): 0 off, 1 bottom, 2 +left,
       3 +right, 4 +top; ass = taper value at the inner end of a strip */
    int atten;
    float ass;
} semsks_problem;

/* ---------------------------------------------------------------------- */
/* Host-side inputs: the reference's parameter and file formats            */
/* ---------------------------------------------------------------------- */
typedef struct semsks_model semsks_model;

/* Replaces main()'s parameter block and Parts 1-2/6.0 input reading
   (sem2d_psvsh.c:154-205, discret 534-828, receivers 330-370, GRT files
   372-396): getpar semantics (name=value, par=file recursion, last
   definition wins, mstpar), raymodel (readmodel.c:43-109), semmodel
   (sem2d_psvsh.c:700-731) with per-node layer assignment, rec_file,
   greenfile/shgreenfile.  Extension keys: solver=0|1|2, semdip=file
   (six tilted-axis constants per semmodel layer), outdir=dir. */
semsks_status semsks_load_par(int argc, char **argv, semsks_model **out);
const semsks_problem *semsks_model_problem(const semsks_model *m);
/* derived sizes: nx, nz, mtotal=2(nx+nz), shtotal=nx+nz, nt_sem, plot_trace */
void semsks_model_dims(const semsks_model *m, int dims[8]);
const char *semsks_model_outdir(const semsks_model *m);
void semsks_model_free(semsks_model *m);

/* The setup diagnostics the reference writes with check=1, its default
   (sem2d_psvsh.c:47,179): grid.xz, rho on every second node column and row
   (:247-260); legendre.xz, the Lagrange basis on [-1, 1] in steps of 0.05
   (:262-270, lagrange.c:20-32,323-366); vel.xy, per node x in degrees of
   epicentral distance from xmin, depth and vp = sqrt(c11/rho), or the
   splitting (v44 - v55)/(v44 + v55)/2 with output_vd=1 (discret, :804-825).
   Written into the model's outdir before the solve; nothing when check=0. */
semsks_status semsks_write_check(const semsks_model *m);

/* The snapshot keys (sem2d_psvsh.c:149-150,202-204): k[0] plot_snap
   (default 0), k[1] startsnap (100), k[2] incresnap (50).  Snapshots
   out1..out26 are the levels startsnap + j incresnap, j = 0..25, that the
   run reaches (:483-510). */
void semsks_model_snapshots(const semsks_model *m, int k[3]);

/* Writes one snapshot file (snapshot(), sem2d_psvsh.c:1699-1715): for every
   second node column ix and row iz, "ix iz u v" (%d %d %e %e) from a field
   in semsks_get_field's layout [3][nz][nx] (components x, y, z); the
   reference prints components 0 and 1 in every solver. */
semsks_status semsks_write_snapshot(const char *outdir, const char *name, int nx, int nz, const float *field);

/* Writes seis.u / seis.w / seis.v (nst records of ntrace float32 each) from
   a trace block [nst][3][ntrace] ordered u, -w, v (kirrecord_zhl,
   sem2d_psvsh.c:2116-2165). */
semsks_status semsks_write_seis(const char *outdir, int nst, int ntrace, const float *traces);

/* The uvw2sac step after the solve (semsyn/plot/uvw2sac.c:40-526): reads
   seis.u/.v/.w and writes <k>.ce/.cn/.cz SAC files (u east, v north, w
   vertical; every second sample, x enlarge), pm.xy and log.txt, from the
   same parameter keys (raymodel nex nst hpn eh dp sdt zbeg idzplot xbeg
   idxplot zs Conv_Source ...; Conv_Source=1 convolves the trapezoid source
   of source.f:137-163).  Extension keys: indir=, outdir=. */
semsks_status semsks_uvw2sac(int argc, char **argv);

/* Human-readable text of the last error on this thread. */
const char *semsks_last_error(void);

/* ---------------------------------------------------------------------- */
/* Engine: device state for up to max_slots instances stepped together     */
/* ---------------------------------------------------------------------- */
typedef struct semsks_engine semsks_engine;

typedef struct semsks_engine_opts {
    size_t hbm_budget;     /* bytes the engine may allocate (0 = 90% of free) */
    int max_slots;         /* instance slots (>= 1) */
    int tile_z;            /* element rows per wave tile (0 = autotuned/default) */
    int waves_per_cu;      /* persistent grid size factor (0 = default) */
    int use_graph;         /* capture the step in a hipGraph (1) or not (0) */
    int steps_per_sweep;   /* 0 or 1: one time step per launch (other
                              values: SEMSKS_EINVAL; a sweep of two steps
                              moved 35% fewer bytes but issued ~1.4x the
                              vector instructions and was slower in every
                              measured case, so it is not offered) */
} semsks_engine_opts;

semsks_status semsks_engine_create(const semsks_engine_opts *opts, semsks_engine **out);
void semsks_engine_destroy(semsks_engine *e);

/* Admits one instance: uploads the problem, runs the FP64 one-off setup on
   the device (GLL tables, Jacobian check, lumped mass and its inverse,
   Lysmer coefficients: sem2d_psvsh.c:218-325), precomputes the interpolated
   and guarded boundary series (interpl.c:16-91, sem2d_psvsh.c:428-444) and
   sets time level 0 (GRT_Interface, sem2d_psvsh.c:446-452).  If share_slot
   >= 0 the instance reuses that slot's mesh and material (same nex, nez,
   hpn, eh, sdt and material required).  Fails with ENOHBM when the slot's
   footprint does not fit the budget. */
semsks_status semsks_admit(semsks_engine *e, const semsks_problem *p, int share_slot, int *slot);

/* Device bytes one problem keeps while it is admitted (for admission
   control). */
size_t semsks_footprint(const semsks_problem *p, int shared_material);

/* Device bytes an engine's admission scratch needs to admit p: its
   boundary records when they are generated on the device (device_records =
   1, semsks_planewave), the FP64 setup temporaries and the interpolation
   schedule, and the upload of host records.  The engine keeps the largest
   scratch it has needed (admissions then need no allocation that waits for
   running steps) and counts it in its HBM budget, on top of the slots'
   semsks_footprint. */
size_t semsks_admission_bytes(const semsks_problem *p, int device_records);

/* Advances every active instance by up to nsteps time steps (the loop body
   of sem2d_psvsh.c:418-512: GRT_Interface4/3, Stiff_Matrix,
   Time_diff_hybrid + Absorbs, the level rotation and kirrecord_zhl).
   Instances stop by themselves at their nst.  sync = 0: asynchronous;
   1: returns when the steps are done; 2: as 1, but launched eagerly (no
   graph) with every step-kernel launch timed (semsks_last_timing). */
semsks_status semsks_advance(semsks_engine *e, int nsteps, int sync);

/* Uploads the tile tables and captures the hipGraphs that a following
   semsks_advance(e, nsteps, 0 or 1) will replay, without running a step,
   so that call launches no capture (e.g. before a timed region).  The
   graphs stay valid until the instance set changes.  No counterpart in the
   reference. */
semsks_status semsks_prepare(semsks_engine *e, int nsteps);

/* Current time level of a slot (-1 if empty) and whether it finished. */
semsks_status semsks_slot_status(semsks_engine *e, int slot, int *level, int *done);

/* Copies traces [n][3][ntrace] of levels it0..it0+n-1 of a slot to host;
   returns SEMSKS_ENONFINITE (traces still copied) if any value is NaN or
   infinite, i.e. the run became unstable. */
semsks_status semsks_get_traces(semsks_engine *e, int slot, int it0, int n, float *out);

/* Copies the current displacement field of a slot to host as three SoA
   planes [3][nz][nx] (u_x, u_y, u_z).  Before the last level, the
   interface row and column already hold the injection of the next step
   (GRT_Interface4/3 act on u(t) in place, after its traces are recorded). */
semsks_status semsks_get_field(semsks_engine *e, int slot, float *out);

/* The slot's current level as the reference's loop leaves it at the end of
   an iteration, before the next step's interface injection: semsks_get_field
   with the interface row and column restored to their values before that
   injection (the field snapshot() writes, sem2d_psvsh.c:483-510,1699-1715).
   Same layout as semsks_get_field. */
semsks_status semsks_get_snapshot(semsks_engine *e, int slot, float *out);

/* Frees a slot (its device memory returns to the budget). */
semsks_status semsks_retire(semsks_engine *e, int slot);

/* Blocks until the engine's stream is idle. */
semsks_status semsks_sync(semsks_engine *e);

/* Milliseconds of device time of the last synchronous semsks_advance call,
   measured with HIP events on the engine's stream: out[0] = the whole call,
   out[1] = the element-force/update kernel launches only (summed over the
   timed steps; eager mode, use_graph = 0, first 1024 steps). */
semsks_status semsks_last_timing(semsks_engine *e, double out[2]);
/* Number of steps whose kernel time entered out[1] of semsks_last_timing. */
int semsks_last_timed_steps(semsks_engine *e);
/* Time steps one step-kernel launch advances: 1. */
int semsks_sweep_steps(semsks_engine *e);
/* Element rows per tile of the current tiling: the autotuned height once
   the autotuner has run for this instance set, the engine's default before,
   or semsks_engine_opts.tile_z when set. */
int semsks_tile_rows(semsks_engine *e);

/* ---------------------------------------------------------------------- */
/* Serving: continuous batching of arriving requests                       */
/* ---------------------------------------------------------------------- */
/* No counterpart in the reference, which runs one solve per process
   (applications/test/prem_epi83.pl:81-82 starts sem2d_psvsh per event). */
typedef struct semsks_request {
    const semsks_problem *problem;  /* in: the instance */
    double arrival_s;               /* in: seconds after semsks_serve starts; non-decreasing */
    int use_planewave;              /* in: 1 = boundary records from semsks_planewave on the
                                       device (problem->psv and ->sh must be NULL) */
    double planewave[4];            /* in: p (s/km), beta (km/s), f0 (Hz), psi (deg) */
    float *traces;                  /* out (optional): [nst][3][ntrace], u, -w, v */
    double admit_s, done_s;         /* out: admission and completion times (s after start) */
    int slot;                       /* out: the slot it ran in */
    int status;                     /* out: SEMSKS_OK, or why it was not run */
} semsks_request;

typedef struct semsks_serve_stats {
    double wall_s;                  /* start to last completion */
    double point_steps;             /* sum over completed requests of nglob * (nst - 1) */
    int completed, max_active, rounds;
} semsks_serve_stats;

/* Runs the requests to completion in real time: each is admitted as a slot
   once it has arrived and its footprint fits the engine's HBM budget (first
   come, first served), all admitted instances advance together
   steps_per_round steps at a time (one persistent launch per step), and
   each is retired -- traces copied out, memory freed -- as soon as it
   reaches its last level.  A request that cannot fit even into an idle
   engine ends with status ENOHBM/ENOSLOT. */
semsks_status semsks_serve(semsks_engine *e, semsks_request *req, int n, int steps_per_round,
                           semsks_serve_stats *stats);

/* Admission order of semsks_serve_ex among the requests that have arrived
   and wait: earliest arrival first, or least work (nglob * (nst - 1))
   first, earliest arrival on a tie. */
#define SEMSKS_SERVE_FCFS 0
#define SEMSKS_SERVE_SHORTEST 1

typedef struct semsks_serve_opts {
    int steps_per_round;            /* steps of every round (>= 1) */
    int order;                      /* SEMSKS_SERVE_FCFS or SEMSKS_SERVE_SHORTEST */
    double max_active_points;       /* admit only while the active instances hold at most this many
                                       GLL points (0: no cap; an idle engine always admits) */
} semsks_serve_opts;

/* semsks_serve with an admission policy (no counterpart in the reference,
   which runs one solve per process).  Under a load above what the GPU
   serves, admitting everything that fits the HBM budget makes every
   admitted request share the device until the backlog drains; a cap on the
   active points that still fills the GPU, with the least work first,
   finishes short requests early at the same throughput.  Admission stops at
   the first request that does not fit (head of line in the chosen order),
   so with SEMSKS_SERVE_SHORTEST a large request waits while smaller ones
   keep arriving: under an unbounded overload it can wait without bound.
   semsks_serve is semsks_serve_ex with FCFS and no cap. */
semsks_status semsks_serve_ex(semsks_engine *e, semsks_request *req, int n, const semsks_serve_opts *opts,
                              semsks_serve_stats *stats);

/* ---------------------------------------------------------------------- */
/* Single-run convenience (what the CLI uses)                              */
/* ---------------------------------------------------------------------- */
/* Runs a whole problem: engine with one slot, nst levels, traces to host
   [nst][3][ntrace]. */
semsks_status semsks_run_problem(const semsks_problem *p, float *traces);

/* ---------------------------------------------------------------------- */
/* Kernel-level entry points (parity tests of single functions)            */
/* ---------------------------------------------------------------------- */
/* K u of Stiff_Matrix (sem2d_psvsh.c:1173-1256) for an arbitrary field of
   slot's mesh and material: u, f host arrays [3][nz][nx].  Runs the same
   element-force code as the time step. */
semsks_status semsks_stiff(semsks_engine *e, int slot, const float *u, float *f);

/* K u as semsks_stiff, with the field and the force in DEVICE memory
   (u_dev, f_dev: [3][nz][nx] floats each, the caller's allocations), for
   callers that cannot allocate device memory (code under HIP graph capture,
   tight serving loops, harnesses that forbid hipMalloc).  No counterpart in
   the reference.  It allocates nothing: its scratch (the field and the force
   in the kernels' node order, 24 bytes per node) is a fixed per-GPU device
   array of SEMSKS_SCRATCH_BYTES (64 MiB, held until the process ends), so
   meshes up to about 2.7 million nodes per plane; a larger mesh gets
   SEMSKS_EINVAL.  Every call on one GPU uses that GPU's one array: calls
   that overlap on one GPU (two engines on two host threads) race.  Returns
   when f_dev holds the result. */
semsks_status semsks_stiff_dev(semsks_engine *e, int slot, const float *u_dev, float *f_dev);

/* Setup products of a slot, to host: lumped mass M (FP64, [nz][nx]), the
   FP32 time-loop factor dt^2/M ([nz][nx]), and the Lysmer coefficients A
   (FP64, [nz][nx][3], zero off the edges). */
semsks_status semsks_get_setup(semsks_engine *e, int slot, double *mass, float *mdt2, double *ab);

/* The interpolated and guarded boundary vectors the time loop injects:
   g [n][2(nx+nz)], gv [n][nx+nz] for levels 0..n-1 (n <= nst). */
semsks_status semsks_get_boundary_series(semsks_engine *e, int slot, int n, float *g, float *gv);

/* Fills a problem's boundary records on the device with a closed-form
   plane S wave (the stand-in for the GRT producer, aserpsv_sem.c:319-488):
   ray parameter p (s/km), bottom-layer S speed beta (km/s), Ricker f0 (Hz),
   polarisation psi (deg; SV share cos psi, SH share sin psi).  nt records
   spaced dp.  Writes the host arrays psv [nt][2(nx+nz)], sh [nt][nx+nz]
   when they are not NULL; with both NULL the records stay on the device and
   are used by the next semsks_admit on this engine with p->psv == NULL and
   p->sh == NULL and p->nt == nt. */
semsks_status semsks_planewave(semsks_engine *e, int nex, int nez, int hpn, int neh, float eh,
                               int nt, float dp, double p, double beta, double f0, double psi_deg,
                               float *psv, float *sh);

#ifdef __cplusplus
}
#endif
#endif /* SEMSKS_H */
