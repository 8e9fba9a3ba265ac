/*
 * semsks_kernels.h -- device data layout shared by the kernels and the engine.
 *
 * Layout in HBM, per instance (slot):
 *   u[l]      displacement of time level l (3 levels rotating): x, y, z
 *             per node, 12 bytes.  Nodes are
 *             stored in column blocks of bw node columns (the owned columns
 *             of one step-kernel tile, bw = (64/(hpn+1) - 1) * hpn), each
 *             block row after row: node (ix, iz) at
 *             (ix / bw) * bw * nz + iz * bw + ix % bw, so a wave sweeping a
 *             tile upward reads and writes one contiguous run per plane
 *   comp[l]   one 32-bit word per node: the Kahan compensations of the three
 *             components at level l and at level l-1, signed 5-bit fields in
 *             units of ulp(u)/32 (without Kahan there are no word planes,
             comp[l] is null and no kernel reads or writes words: a step
             moves 36 bytes of state per node instead of 44)
 *   mat_id    16-bit material record id per node (same node order)
 *   emat      16-bit record id per element when all its nodes share one
 *             record, 0xFFFF otherwise
 *   mat       records {rho, 15 constants}, 16 floats each
 *   mclass    dt^2/M per material record and node position class
 *             [n_mat][hpn+1][hpn+1] (FP32 from the FP64 lumped mass)
 *   edge_a/b  (M-A)/(M+A) and dt^2/(M+A) per absorbing-edge node, 3 comps
 *   G, Gv     interpolated, guarded boundary vectors per time level
 *   traces    [nst][3][ntrace] receiver samples (u, -w, v)
 */
#ifndef SEMSKS_KERNELS_H
#define SEMSKS_KERNELS_H

#include <stdint.h>

#define SEMSKS_MAXNP 7

/* tile-queue counters of the step kernel: SEMSKS_NQ per polynomial degree
   (index hpn * SEMSKS_NQ + queue), each on its own 128-byte line */
#define SEMSKS_NQ 8
#define SEMSKS_QSTRIDE 32
#define SEMSKS_QCOUNTERS (SEMSKS_MAXNP * SEMSKS_NQ)

/* bytes of the fixed per-GPU scratch array of semsks_stiff_dev: the field
   and the force of one plane, 24 bytes per node (with padding), so meshes
   of up to about 2.7 million nodes per plane */
#define SEMSKS_SCRATCH_BYTES ((size_t)64 << 20)

struct InstDesc {
    int nex, nez, nx, nz, bw;       /* bw: node columns per storage block */
    int hpn, neh, nst, nt_sem, solver, kahan, ntrace, active, nlev;
    /* displacement of nlev = 3 time levels (entries 0-2 used), level it
       in plane it % nlev: u[l] holds x, y, z per node (node q at
       u[l][3q..3q+2]), comp[l] one 32-bit word per node with the packed
       Kahan compensations of the three components at that level (bits
       0-14) and at the level before (bits 15-29) */
    float *u[4];
    uint32_t *comp[4];
    float *fout[3];                 /* force-only mode output planes, else null */
    const uint16_t *mat_id;
    const uint16_t *emat;
    const float *mat;
    const float *mclass;
    const float *edge_a, *edge_b;   /* [n_edge][3] */
    const float *G, *Gv;            /* [nt_sem][2(nx+nz)], [nt_sem][nx+nz] */
    const int *rec;                 /* node offsets of the receivers */
    const int *inj_rec;             /* per interface node: first receiver on it, or -1 */
    const int *rec_next;            /* per receiver: next one on its interface node (-1 at the
                                       end), -2 if it is not on an interface node */
    const float *taper;             /* sponge factors [nh] (atten > 0) */
    /* values of the interface nodes (inject_node's order) of the newest
       level before its injection, [n_inject][3]: the field as the
       reference's loop leaves it at the end of an iteration */
    float *pre_inj;
    int atten, nh;
    float *traces;
    /* element operators (FP32 from FP64):
       Dg[a][t] = l'_t(xi_a) * 2/eh      (gradient, dlagP_JeI)
       Dw[a][k] = l'_k(xi_a) * w_a       (divergence, dlagP_Gllw)
       cw[k]    = Je * Je^-1 * w_k = (eh/2) w_k                       */
    float Dg[SEMSKS_MAXNP][SEMSKS_MAXNP];
    float Dw[SEMSKS_MAXNP][SEMSKS_MAXNP];
    float cw[SEMSKS_MAXNP];
    int ortho;                      /* all material records orthotropic in the x, y, z axes */
};

#if defined(__HIPCC__)
/* Pointers taken out of an InstDesc are generic to the compiler; casting
   them to the global address space makes the hot loops use global_load /
   global_store instead of FLAT memory operations (which also occupy the
   LDS counter and so serialise with cross-lane shuffles). */
#if defined(__HIP_DEVICE_COMPILE__)
#define SEMSKS_GMEM __attribute__((address_space(1)))
#else
#define SEMSKS_GMEM
#endif
template <class T>
__device__ __forceinline__ SEMSKS_GMEM T *gp(T *p) { return (SEMSKS_GMEM T *)p; }
template <class T>
__device__ __forceinline__ const SEMSKS_GMEM T *gp(const T *p) { return (const SEMSKS_GMEM T *)p; }
/* read-only data at a wave-uniform address (material records, element
   record ids): the constant address space lets the compiler load it with
   scalar loads into scalar registers */
#if defined(__HIP_DEVICE_COMPILE__)
#define SEMSKS_CMEM __attribute__((address_space(4)))
#else
#define SEMSKS_CMEM
#endif
template <class T>
__device__ __forceinline__ const SEMSKS_CMEM T *cp(const T *p) { return (const SEMSKS_CMEM T *)p; }
#endif

/* edge-node numbering: bottom row iz=0 (nx) | left column iz>=1 (nz-1) |
   right column iz>=1 (nz-1) | top row iz=nz-1, ix=1..3*hpn (3*hpn) */
#ifdef __HIPCC__
#define SEMSKS_HD __host__ __device__
#else
#define SEMSKS_HD
#endif
static inline SEMSKS_HD int semsks_n_edge(int nx, int nz, int hpn) { return nx + 2 * (nz - 1) + 3 * hpn; }

#endif
