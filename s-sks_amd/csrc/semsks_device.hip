/*
 * semsks_device.hip -- gfx950 kernels and the engine behind include/semsks.h.
 *
 * One time step of one instance (semsyn/src/sem2d_psvsh.c:418-512):
 *   step_kernel    Stiff_Matrix (sem2d_psvsh.c:1173-1256: dUj_di_2D, Stress_ij,
 *                  F_ik, Grid_Stiff) fused with Time_diff_hybrid + Absorbs
 *                  (sem2d_psvsh.c:1004-1029, 1834-1889) and the level rotation
 *                  (Copy_Displacement x2, :468-469, replaced by rotating planes)
 *   post_kernel    kirrecord_zhl (sem2d_psvsh.c:2116-2165) on the new level,
 *                  then GRT_Interface4 / GRT_Interface3 (sem2d_psvsh.c:
 *                  2076-2105, sem2dMP20140816.c:2001-2037) on the interface
 *                  row/column of that level, for the next step (level 0 gets
 *                  its injection at admission, inject0_kernel)
 *
 * The update is written in increment form: with delta_n = u_{n+1} - u_n,
 *   interior : delta_n = delta_{n-1} - (dt^2/M) f(u_n)
 *   edges    : delta_n = ((M-A) delta_{n-1} - dt^2 f) / (M+A)
 * which is algebraically the reference's u+ = -(dt^2/M) f + 2u - u- and its
 * absorbing rows.  delta_{n-1} comes from the compensated values of u_n and
 * u_{n-1} (each u + packed compensation), and u_{n+1} = u_n + delta_n is a
 * Kahan-compensated sum.
 *
 * Element force (row-per-lane layout): a 64-lane wave owns a tile of
 * NE-1 = 64/(N+1)-1 element columns (11 for N = 4) x TZ element rows.  The
 * wave holds NE element columns side by side, N+1 lanes per element, one
 * lane per node row of the element; column 0 is a halo whose only output is
 * its shared right node column.  The wave sweeps the element rows upward
 * from z0-1 (a halo row whose only output is its shared top node row).
 * x-derivatives and row tractions stay inside a lane; z-derivatives and
 * column tractions go through a small per-wave LDS block.  Shared nodes are
 * assembled without atomics or colouring: the raw top-row force of an
 * element is carried to the next element row through LDS, the right node
 * column is handed to the next element with a cross-lane shift, and every
 * node is written by exactly one lane, which applies the update in the same
 * pass while the row of u is still in registers.
 *
 * State in HBM: three time levels of u rotating (level it in plane
 * it % 3), each an xyz plane (12 bytes per node) and a plane of 32-bit
 * compensation words that hold the packed Kahan compensations of that level
 * and of the level before.  A step reads the xyz of level it (with halos),
 * the xyz of level it-1 and the words of level it, and writes the xyz and
 * words of level it+1.  Planes are padded to a multiple of 64 nodes.
 */
#include <hip/hip_runtime.h>
#include <math.h>
#include <stdio.h>
#include <stdlib.h>
#include <string.h>
#include <vector>
#include <map>
#include <mutex>
#include <algorithm>

#include "semsks.h"
#include "semsks_internal.h"
#include "semsks_kernels.h"

/* node offset in a plane: column blocks of bw node columns, each stored row
   after row (one block holds the owned columns of one step-kernel tile) */
__host__ __device__ static inline long node_col(int bw, int nz, int ix)
{
    const int blk = ix / bw;
    return (long)blk * bw * nz + (ix - blk * bw);
}
__host__ __device__ static inline long node_at(int bw, int nz, int ix, int iz)
{
    return node_col(bw, nz, ix) + (long)iz * bw;
}
__host__ __device__ static inline long node_at(const InstDesc &D, int ix, int iz)
{
    return node_at(D.bw, D.nz, ix, iz);
}

#define HIPCHK(x)                                                                         \
    do {                                                                                  \
        hipError_t e_ = (x);                                                              \
        if (e_ != hipSuccess)                                                             \
            return semsks_set_error(SEMSKS_EDEVICE, "%s:%d %s: %s", __FILE__, __LINE__, #x, \
                                    hipGetErrorString(e_));                               \
    } while (0)

/* ========================================================================= */
/* FP64 one-off setup                                                        */
/* ========================================================================= */

/* GLL nodes (closed forms, lagrange.c:323-366) and weights
   2/(n(n+1) P_n(x)^2) (lagrange.c:256-278), in double */
__host__ __device__ static void gll64(int n, double *x, double *w)
{
    if (n == 3) { double r = sqrt(1.0 / 5.0); x[0] = -1; x[1] = -r; x[2] = r; x[3] = 1; }
    else if (n == 4) { double r = sqrt(15.0 / 35.0); x[0] = -1; x[1] = -r; x[2] = 0; x[3] = r; x[4] = 1; }
    else {
        double a = n == 5 ? 315.0 : 1386.0, c = n == 5 ? -210.0 : -1260.0, e = n == 5 ? 15.0 : 210.0;
        double d = sqrt(c * c - 4 * a * e);
        double r1 = sqrt((-c + d) / 2 / a), r2 = sqrt((-c - d) / 2 / a);
        x[0] = -1; x[1] = -r1; x[2] = -r2;
        if (n == 5) { x[3] = r2; x[4] = r1; x[5] = 1; }
        else { x[3] = 0; x[4] = r2; x[5] = r1; x[6] = 1; }
    }
    for (int i = 0; i <= n; ++i) {
        if (i == 0 || i == n) { w[i] = 2.0 / (n * (n + 1)); continue; }
        double p2 = 1, p1 = x[i], pn = 0;
        for (int k = 2; k <= n; ++k) { pn = ((2 * k - 1) * x[i] * p1 - (k - 1) * p2) / k; p2 = p1; p1 = pn; }
        w[i] = 2.0 / (n * (n + 1) * pn * pn);
    }
}

/* l'_k(x_j) by the product rule (dLag_N_i, lagrange.c:42-67), in double */
__host__ __device__ static double dlag64(int n, int k, const double *x, double s)
{
    double down = 1.0, v = 0.0;
    for (int i = 0; i <= n; ++i) if (i != k) down *= (x[k] - x[i]);
    for (int i = 0; i <= n; ++i) {
        if (i == k) continue;
        double L = 1.0;
        for (int m = 0; m <= n; ++m) if (m != k && m != i) L *= (s - x[m]);
        v += L;
    }
    return v / down;
}

/* element operators of one instance, FP64 -> FP32 (one thread) */
__global__ void setup_tables_kernel(InstDesc *D, double *w64, double eh)
{
    if (threadIdx.x != 0 || blockIdx.x != 0) return;
    const int n = D->hpn;
    double x[SEMSKS_MAXNP], w[SEMSKS_MAXNP];
    gll64(n, x, w);
    const double jei = 2.0 / eh, half = eh / 2.0;
    for (int a = 0; a <= n; ++a) {
        w64[a] = w[a];
        D->cw[a] = (float)(half * w[a]);
        for (int t = 0; t <= n; ++t) {
            double d = dlag64(n, t, x, x[a]);       /* l'_t(x_a) */
            D->Dg[a][t] = (float)(d * jei);
            D->Dw[a][t] = (float)(d * w[a]);
        }
    }
}

/* bilinear Jacobian of every element at every GLL point, FP64
   (All_Je + Jacobian_4, sem2d_psvsh.c:857-889, lagrange.c:216-248);
   records the largest deviation from eh^2/4 */
__global__ void jacobian_kernel(int nex, int nez, int n, double eh, unsigned long long *maxdev_bits)
{
    double x[SEMSKS_MAXNP], w[SEMSKS_MAXNP];
    gll64(n, x, w);
    long nel = (long)nex * nez;
    double dev = 0.0, je0 = eh * eh / 4.0;
    for (long e = blockIdx.x * (long)blockDim.x + threadIdx.x; e < nel; e += (long)gridDim.x * blockDim.x) {
        int ex = (int)(e / nez), ez = (int)(e % nez);
        double px[4] = {ex * eh, (ex + 1) * eh, ex * eh, (ex + 1) * eh};
        double pz[4] = {ez * eh, ez * eh, (ez + 1) * eh, (ez + 1) * eh};
        for (int m = 0; m <= n; ++m)
            for (int k = 0; k <= n; ++k) {
                double rx = x[m], rz = x[k];
                double Nrx[4] = {-1 + rz, 1 - rz, -1 - rz, 1 + rz}, Nrz[4] = {-1 + rx, -1 - rx, 1 - rx, 1 + rx};
                double j0 = 0, j1 = 0, j2 = 0, j3 = 0;
                for (int i = 0; i < 4; ++i) { j0 += px[i] * Nrx[i] / 4; j1 += px[i] * Nrz[i] / 4; j2 += pz[i] * Nrx[i] / 4; j3 += pz[i] * Nrz[i] / 4; }
                double je = j0 * j3 - j1 * j2;
                dev = fmax(dev, fabs(je - je0));
            }
    }
    atomicMax(maxdev_bits, (unsigned long long)__double_as_longlong(dev));
}

/* lumped mass M = sum_e w_m w_n Je rho in the element order of Mass_Matrix
   (sem2d_psvsh.c:1042-1071; every term at a node is equal, so the sum is
   k additions of one term), dt^2/M to FP32 */
__global__ void mass_kernel(InstDesc D, const double *w64, double je, float sdt, double *mass)
{
    const int n = D.hpn;
    const long N = (long)D.nx * D.nz;
    for (long q = blockIdx.x * (long)blockDim.x + threadIdx.x; q < N; q += (long)gridDim.x * blockDim.x) {
        int iz = (int)(q / D.nx), ix = (int)(q % D.nx);
        int m = ix % n, k = iz % n;
        int cx = 1, cz = 1;
        if (m == 0 && ix > 0 && ix < D.nx - 1) cx = 2;
        if (k == 0 && iz > 0 && iz < D.nz - 1) cz = 2;
        if (ix == D.nx - 1) m = n;
        if (iz == D.nz - 1) k = n;
        double rho = (double)D.mat[(long)D.mat_id[node_at(D, ix, iz)] * SEMSKS_MAT_STRIDE];
        double t = w64[m] * w64[k] * je * rho, s = 0.0;
        for (int c = 0; c < cx * cz; ++c) s = s + t;
        if (mass) mass[q] = s;
    }
}

/* dt^2/M per material record and position class (cz, cx): class 0 = node
   shared by two elements (two equal terms w_0), 1..N-1 = interior GLL
   point, N = first/last node of the mesh (one term w_0 = w_N).  The sum is
   formed term by term as in Mass_Matrix, so every entry equals dt^2/M of
   every node of that class, bit for bit. */
__global__ void mclass_kernel(const float *mat, int n_mat, int n, const double *w64, double je, float sdt, float *out)
{
    const int NC = n + 1;
    for (int id = blockIdx.x * blockDim.x + threadIdx.x; id < n_mat * NC * NC; id += gridDim.x * blockDim.x) {
        const int m = id / (NC * NC), r = id % (NC * NC), cz = r / NC, cx = r % NC;
        const int mx = cx == n ? 0 : cx, mz = cz == n ? 0 : cz;
        const int cnt = (cx == 0 ? 2 : 1) * (cz == 0 ? 2 : 1);
        const double rho = (double)mat[(long)m * SEMSKS_MAT_STRIDE];
        const double t = w64[mx] * w64[mz] * je * rho;
        double s = 0.0;
        for (int c = 0; c < cnt; ++c) s = s + t;
        out[id] = (float)((double)sdt * (double)sdt / s);
    }
}

/* per-node dt^2/M from the class table (for inspection) */
__global__ void expand_mdt2_kernel(InstDesc D, float *out)
{
    const int n = D.hpn, NC = n + 1;
    const long N = (long)D.nx * D.nz;
    for (long q = blockIdx.x * (long)blockDim.x + threadIdx.x; q < N; q += (long)gridDim.x * blockDim.x) {
        const int iz = (int)(q / D.nx), ix = (int)(q % D.nx);
        const int cx = (ix == 0 || ix == D.nx - 1) ? n : ix % n;
        const int cz = (iz == 0 || iz == D.nz - 1) ? n : iz % n;
        const int mid = D.mat_id[node_at(D, ix, iz)];
        out[q] = D.mclass[(long)mid * NC * NC + cz * NC + cx];
    }
}

__device__ static double lysmer_term(const InstDesc &D, long node, double w, double jb, float sdt, int normal, int k)
{
    const float *r = D.mat + (long)D.mat_id[node] * SEMSKS_MAT_STRIDE;
    double rho = r[0];
    double v = (k == normal) ? sqrt((double)r[1] / rho) : sqrt((double)r[10] / rho);
    return w * rho * v * jb * sdt / 2.;
}

/* Lysmer coefficients A (Absorb_Matrix, sem2d_psvsh.c:1093-1162, summed in
   its loop order: left, bottom, right) and the edge update factors */
__global__ void absorb_kernel(InstDesc D, const double *w64, double jb, float sdt, int absorb, const double *mass,
                              double *ab_out, float *ea, float *eb)
{
    const int n = D.hpn, nx = D.nx, nz = D.nz;
    const int ne = semsks_n_edge(nx, nz, n);
    for (int e = blockIdx.x * blockDim.x + threadIdx.x; e < ne; e += gridDim.x * blockDim.x) {
        int ix, iz;
        if (e < nx) { ix = e; iz = 0; }
        else if (e < nx + nz - 1) { ix = 0; iz = e - nx + 1; }
        else if (e < nx + 2 * (nz - 1)) { ix = nx - 1; iz = e - nx - (nz - 1) + 1; }
        else { ix = e - nx - 2 * (nz - 1) + 1; iz = nz - 1; }
        if (ix >= nx) continue;
        long node = node_at(D, ix, iz);
        double A[3] = {0, 0, 0};
        if (absorb) {
            for (int k = 0; k < 3; ++k) {
                double a = 0.0;
                if (ix == 0) {           /* left: elements ez with ez*n+be == iz, ez ascending */
                    for (int ez = 0; ez < D.nez; ++ez) {
                        int be = iz - ez * n;
                        if (be >= 0 && be <= n) a = a + lysmer_term(D, node, w64[be], jb, sdt, 0, k);
                    }
                }
                if (iz == 0) {           /* bottom */
                    for (int ex = 0; ex < D.nex; ++ex) {
                        int al = ix - ex * n;
                        if (al >= 0 && al <= n) a = a + lysmer_term(D, node, w64[al], jb, sdt, 2, k);
                    }
                }
                if (ix == nx - 1) {      /* right */
                    for (int ez = 0; ez < D.nez; ++ez) {
                        int be = iz - ez * n;
                        if (be >= 0 && be <= n) a = a + lysmer_term(D, node, w64[be], jb, sdt, 0, k);
                    }
                }
                A[k] = a;
            }
        }
        double M = mass[(long)iz * nx + ix];
        double dt2 = (double)sdt * (double)sdt;
        for (int k = 0; k < 3; ++k) {
            if (ab_out) ab_out[((long)iz * nx + ix) * 3 + k] = A[k];
            ea[e * 3 + k] = (float)((M - A[k]) / (M + A[k]));
            eb[e * 3 + k] = (float)(dt2 / (M + A[k]));
        }
    }
}

/* ========================================================================= */
/* boundary series: 4-tap interpolation of the records + guard               */
/* ========================================================================= */
struct SchedEntry { int rec[4]; float w[4]; };

/* G[k][i] = ((((0 + w0 s0) + w1 s1) + w2 s2) + w3 s3) in FP32 without
   contraction (interpl.c:43-48), then the sequential non-finite / |g|>=10
   repair of sem2d_psvsh.c:433-443 when a row needs it */
__global__ void series_kernel(const float *rec, int width, int nrec, const SchedEntry *sched, int nlev, float *G)
{
    for (int k = blockIdx.x; k < nlev; k += gridDim.x) {
        const SchedEntry s = sched[k];
        float *g = G + (long)k * width;
        int mybad = 0;
        for (int i = threadIdx.x; i < width; i += blockDim.x) {
            float acc = 0.0f;
#pragma unroll
            for (int j = 0; j < 4; ++j) {
                float v = (s.rec[j] >= 0 && s.rec[j] < nrec) ? rec[(long)s.rec[j] * width + i] : 0.0f;
                acc = __fadd_rn(acc, __fmul_rn(s.w[j], v));
            }
            g[i] = acc;
            if (!(acc < 10.0f && acc > -10.0f)) mybad = 1;
        }
        mybad = __syncthreads_or(mybad);
        if (mybad && threadIdx.x == 0) {
            for (int i = 0; i < width; ++i)
                if (!(g[i] < 10.0f && g[i] > -10.0f)) {
                    float l = i > 0 ? g[i - 1] : 0.0f, r = i + 1 < width ? g[i + 1] : 0.0f;
                    g[i] = (l + r) / 2;
                }
        }
        __syncthreads();
    }
}

/* ========================================================================= */
/* level 0, injection, receivers                                             */
/* ========================================================================= */

/* GRT_Interface at it == 0 (sem2d_psvsh.c:1953-2005): u on the interface row
   and column from g(0); u(-dt) = 0 (the cleared buffer of level -1) */
__global__ void init_state_kernel(InstDesc D)
{
    const int nbl = D.neh * D.hpn, nx = D.nx, nz = D.nz;
    const float *g = D.G, *gv = D.Gv;
    const bool sh = D.solver == SEMSKS_SOLVER_PSVSH;
    for (int i = blockIdx.x * blockDim.x + threadIdx.x; i < nx + nz; i += gridDim.x * blockDim.x) {
        long q;
        float u, v = 0.f, w;
        if (i < nx) {
            int ix = i;
            if (ix < nbl) continue;
            if (ix == nbl) continue;                 /* the left pass writes the corner */
            q = node_at(D, ix, nbl);
            u = g[ix]; w = -g[nx + ix]; if (sh) v = gv[ix];
        } else {
            int iz = i - nx;
            if (iz < nbl) continue;
            q = node_at(D, nbl, iz);
            u = g[2 * nx + iz]; w = -g[2 * nx + nz + iz]; if (sh) v = gv[nx + iz];
        }
        D.u[0][3 * q] = u; D.u[0][3 * q + 1] = v; D.u[0][3 * q + 2] = w;
    }
}

__device__ __forceinline__ void kahan_add(float &u, float &c, float x)
{
    float y = x - c;
    float t = u + y;
    c = (t - u) - y;
    u = t;
}

/* The Kahan compensation c of u is the rounding error of the last addition,
   |c| <= ulp(u)/2 (true value u - c).  It is stored per component as a
   signed 5-bit field in units of ulp(u)/32 (2^(e-29) with u = m 2^e,
   0.5 <= |m| < 1), clamped to +-15; the residual it leaves is below
   ulp(u)/64 per step.  The 32-bit word of level n holds the fields of level
   n (bits 0-14, component k at 5k) and those of level n-1 (bits 15-29), so
   a step reads the word of level n only: the previous increment
   U_n - U_{n-1} needs both, the Kahan sum the first. */
constexpr uint32_t C_OWN = 0x7fffu;          /* fields of the word's own level */

__device__ __forceinline__ float c_field(uint32_t w, int f, float u)
{
    const int q = (int)(w << (27 - 5 * f)) >> 27;
    return ldexpf((float)q, __builtin_amdgcn_frexp_expf(u) - 29);
}

__device__ __forceinline__ int c_quant(float c, float u)
{
    const int q = (int)rintf(ldexpf(c, 29 - __builtin_amdgcn_frexp_expf(u)));
    return q > 15 ? 15 : (q < -15 ? -15 : q);
}

__device__ __forceinline__ uint32_t c_pack3(int q0, int q1, int q2)
{
    return ((uint32_t)q0 & 31u) | (((uint32_t)q1 & 31u) << 5) | (((uint32_t)q2 & 31u) << 10);
}

/* u += d for the three components as an in-place change of one level (the
   interface injection): w is the word of u's level, its own fields take the
   new compensations and the previous level's fields stay.  Without
   compensation w stays 0.  (A step to the next level is advance3.) */
__device__ __forceinline__ void kahan3(bool kahan, float (&u)[3], uint32_t &w, const float (&d)[3])
{
    if (kahan) {
        int q[3];
#pragma unroll
        for (int k = 0; k < 3; ++k) {
            float c = c_field(w, k, u[k]);
            kahan_add(u[k], c, d[k]);
            q[k] = c_quant(c, u[k]);
        }
        w = c_pack3(q[0], q[1], q[2]) | (w & ~C_OWN);
    } else {
#pragma unroll
        for (int k = 0; k < 3; ++k) u[k] = u[k] + d[k];
        w = 0u;
    }
}

/* a node's xyz at float offset 3q of a level plane */
__device__ __forceinline__ void xyz_load(const SEMSKS_GMEM float *P, long q, float (&v)[3])
{
    v[0] = P[3 * q]; v[1] = P[3 * q + 1]; v[2] = P[3 * q + 2];
}

__device__ __forceinline__ void xyz_store(SEMSKS_GMEM float *P, long q, const float (&v)[3])
{
    P[3 * q] = v[0]; P[3 * q + 1] = v[1]; P[3 * q + 2] = v[2];
}

/* GRT_Interface4: u += g(t) - g(t-dt) (w sign flipped); GRT_Interface3:
   u += -u(t-dt) + g(t).  Both change u(t) in place before the force; the
   changed u(t) is the next step's u(t-dt), as after the reference's copy.
   inject_delta gives the increment at one interface node for the step that
   leaves level it: the bottom row iz = nbl (ix >= nbl, the corner included)
   or the left column ix = nbl (iz > nbl); up0/up2 are u_x/u_z of level it-1
   at the node (GRT_Interface3 only). */
__device__ __forceinline__ void inject_delta(const InstDesc &D, int it, int ix, int iz, float up0, float up2, float (&d)[3])
{
    const int nbl = D.neh * D.hpn, nx = D.nx, nz = D.nz, mt = 2 * (nx + nz), sht = nx + nz;
    const int lev = it + 1;
    const float *g = D.G + (long)lev * mt, *g0 = D.G + (long)(lev - 1) * mt;
    const float *gv = D.Gv + (long)lev * sht, *gv0 = D.Gv + (long)(lev - 1) * sht;
    const bool sh = D.solver == SEMSKS_SOLVER_PSVSH, mp = D.solver == SEMSKS_SOLVER_PSV_MP;
    if (iz == nbl) {
        if (mp) {
            d[0] = -up0 + g[ix];
            d[1] = 0.f;
            d[2] = -up2 - g[ix + nx];
        } else {
            d[0] = -g0[ix] + g[ix];
            d[1] = sh ? -gv0[ix] + gv[ix] : 0.f;
            d[2] = g0[ix + nx] - g[ix + nx];
        }
    } else {
        if (mp) {
            d[0] = -up0 + g[2 * nx + iz];
            d[1] = 0.f;
            d[2] = -up2 - g[2 * nx + nz + iz];
        } else {
            d[0] = -g0[2 * nx + iz] + g[2 * nx + iz];
            d[1] = sh ? -gv0[nx + iz] + gv[nx + iz] : 0.f;
            d[2] = g0[2 * nx + nz + iz] - g[2 * nx + nz + iz];
        }
    }
}

__device__ __forceinline__ bool is_inject_node(int ix, int iz, int nbl)
{
    return (iz == nbl && ix >= nbl) || (ix == nbl && iz > nbl);
}

/* The interface injection of one node, in place on level it (plane p):
   interface index i runs over the bottom row iz = nbl (ix >= nbl, the
   corner included) and then the left column ix = nbl (iz > nbl). */
__device__ __forceinline__ void inject_node(const InstDesc &D, int it, int i, int nb)
{
    const int L = D.nlev, p = it % L, nbl = D.neh * D.hpn;
    const int ix = i < nb ? nbl + i : nbl;
    const int iz = i < nb ? nbl : nbl + 1 + (i - nb);
    const long q = node_at(D, ix, iz);
    const bool mp = D.solver == SEMSKS_SOLVER_PSV_MP;
    float *U = D.u[p];
    const float *Up = D.u[(p + L - 1) % L];
    float d[3];
    inject_delta(D, it, ix, iz, mp ? Up[3 * q] : 0.f, mp ? Up[3 * q + 2] : 0.f, d);
    float u[3] = {U[3 * q], U[3 * q + 1], U[3 * q + 2]};
    for (int k = 0; k < 3; ++k) D.pre_inj[3 * i + k] = u[k];
    uint32_t w = D.kahan ? D.comp[p][q] : 0u;
    kahan3(D.kahan != 0, u, w, d);
    for (int k = 0; k < 3; ++k) U[3 * q + k] = u[k];
    if (D.kahan) D.comp[p][q] = w;
}

/* does level it get an injection before the step that leaves it? */
__device__ __forceinline__ bool injects(const InstDesc &D, int it)
{
    return it < D.nst - 1 && it + 1 < D.nt_sem;
}

/* the injection of level 0 at admission, before the instance's first step
   (every later level gets its injection in post_kernel) */
__global__ void inject0_kernel(InstDesc D)
{
    if (!injects(D, 0)) return;
    const int nbl = D.neh * D.hpn, nb = D.nx - nbl, nl = D.nz - nbl - 1;
    for (int i = blockIdx.x * blockDim.x + threadIdx.x; i < nb + nl; i += gridDim.x * blockDim.x) inject_node(D, 0, i, nb);
}

/* atten_wave2 (sem2d_psvsh.c:1780-1823): after the update, u(t+dt) on the
   edge strips of nh = neh*hpn nodes is multiplied by the taper (bottom for
   atten >= 1, left >= 2, right >= 3, top >= 4; each node at most once).
   The compensation of the rescaled value restarts. */
__global__ void sponge_kernel(const InstDesc *insts, const int *levels)
{
    const InstDesc &D = insts[blockIdx.y];
    if (!D.active || D.atten <= 0) return;
    const int it = levels[blockIdx.y];
    if (it >= D.nst - 1) return;
    const int pn = (it + 1) % D.nlev, nh = D.nh, nx = D.nx, nz = D.nz, at = D.atten;
    const long nb = (long)nh * nx, ns = (long)(nz - nh) * nh, ntop = (long)nh * (nx - 2 * nh);
    const long total = nb + (at > 1 ? ns : 0) + (at > 2 ? ns : 0) + (at > 3 ? ntop : 0);
    for (long i = blockIdx.x * (long)blockDim.x + threadIdx.x; i < total; i += (long)gridDim.x * blockDim.x) {
        int ix, iz;
        float f;
        long j = i;
        if (j < nb) {
            iz = (int)(j / nx); ix = (int)(j - (long)iz * nx);
            f = D.taper[nh - iz - 1];
        } else if ((j -= nb) < ns) {
            iz = nh + (int)(j / nh); ix = (int)(j % nh);
            f = D.taper[nh - ix - 1];
        } else if ((j -= ns) < ns) {
            iz = nh + (int)(j / nh); ix = nx - nh + (int)(j % nh);
            f = D.taper[ix - nx + nh];
        } else {
            j -= ns;
            iz = nz - nh + (int)(j / (nx - 2 * nh)); ix = nh + (int)(j % (nx - 2 * nh));
            f = D.taper[iz - nz + nh];
        }
        const long q = node_at(D, ix, iz);
        float *U1 = D.u[pn] + 3 * q;
        for (int k = 0; k < 3; ++k) U1[k] = f * U1[k];
        if (D.kahan) D.comp[pn][q] &= ~C_OWN;          /* the previous level's fields stay */
    }
}

/* kirrecord_zhl: u to seis.u, -w to seis.w, v to seis.v */
__device__ __forceinline__ void record_one(const InstDesc &D, int lev, int r)
{
    const int nt = D.ntrace;
    const float *U = D.u[lev % D.nlev];
    float *tr = D.traces + (long)lev * 3 * nt;
    const long q = D.rec[r];
    tr[r] = U[3 * q];
    tr[nt + r] = -U[3 * q + 2];
    tr[2 * nt + r] = U[3 * q + 1];
}

/* After a step launch, per slot (blockIdx.y), grid-wide over blockIdx.x:
   records every level the launch produced (up to smax), then applies the
   injection that the next step needs to the newest level in place
   (GRT_Interface4 / GRT_Interface3, sem2d_psvsh.c:2076-2105,
   sem2dMP20140816.c:2001-2037, which the reference applies to u(t) at the
   top of the next iteration, after kirrecord_zhl has written it), and
   advances the slot's level.  A receiver on an interface node is recorded
   by the thread that injects that node, before the injection
   (D.inj_rec[i]: first receiver at interface node i, D.rec_next[r]: next
   one at the same node, -1 at the end of a list, -2 for a receiver off the
   interface).  The last block of a slot to finish advances its level
   (pctr counts the finished blocks).  It also clears the step kernel's
   tile-queue counters for the next launch. */
__global__ void post_kernel(const InstDesc *insts, int *levels, int smax, int *qctr, int *pctr)
{
    if (blockIdx.x == 0 && blockIdx.y == 0 && threadIdx.x < SEMSKS_QCOUNTERS) qctr[threadIdx.x * SEMSKS_QSTRIDE] = 0;
    const int s = blockIdx.y;
    const InstDesc &D = insts[s];
    if (!D.active) return;
    const int it = levels[s];
    if (it >= D.nst - 1) return;
    const int adv = min(smax, D.nst - 1 - it), lev = it + adv, nt = D.ntrace;
    const bool inj = injects(D, lev);
    const int tid = blockIdx.x * blockDim.x + threadIdx.x, nth = gridDim.x * blockDim.x;
    for (int l = 1; l <= adv; ++l)
        for (int r = tid; r < nt; r += nth)
            if (!(inj && l == adv && D.rec_next[r] != -2)) record_one(D, it + l, r);
    if (inj) {
        const int nbl = D.neh * D.hpn, nb = D.nx - nbl, nl = D.nz - nbl - 1;
        for (int i = tid; i < nb + nl; i += nth) {
            for (int r = D.inj_rec[i]; r >= 0; r = D.rec_next[r]) record_one(D, lev, r);
            inject_node(D, lev, i, nb);
        }
    }
    /* every thread of the block has used the level it read; no data goes
       from one block to another here (the next launch sees this kernel's
       stores), so the count needs no fence */
    __syncthreads();
    if (threadIdx.x == 0) {
        if (atomicAdd(pctr + s, 1) == (int)gridDim.x - 1) {
            levels[s] = lev;
            pctr[s] = 0;
        }
    }
}

__global__ void record0_kernel(InstDesc D)
{
    const int nt = D.ntrace;
    for (int r = threadIdx.x; r < nt; r += blockDim.x) {
        long q = D.rec[r];
        D.traces[r] = D.u[0][3 * q];
        D.traces[nt + r] = -D.u[0][3 * q + 2];
        D.traces[2 * nt + r] = D.u[0][3 * q + 1];
    }
}

/* ========================================================================= */
/* the element force + update kernel                                         */
/* ========================================================================= */
/* Compile-time element operators.  For square elements the 2-D stiffness
   is independent of the element size (gradient ~ 2/eh, quadrature weight ~
   (eh/2)^2), so the kernel uses the reference-element operators
     D[a][t]  = l'_t(x_a)                  (dLag_N_i, lagrange.c:42-67)
     Dw[a][k] = l'_k(x_a) w_a              (dlagP_Gllw, sem2d_psvsh.c:299-300)
   evaluated in double at compile time from the GLL nodes and rounded to
   FP32.  Only the mass (dt^2/M) carries eh. */
template <int N> struct GllNodes;
template <> struct GllNodes<3> { static constexpr double x[4] = {-1.0, -0.44721359549995793928, 0.44721359549995793928, 1.0}; };
template <> struct GllNodes<4> { static constexpr double x[5] = {-1.0, -0.65465367070797714380, 0.0, 0.65465367070797714380, 1.0}; };
template <> struct GllNodes<5> { static constexpr double x[6] = {-1.0, -0.76505532392946469285, -0.28523151648064509632, 0.28523151648064509632, 0.76505532392946469285, 1.0}; };
template <> struct GllNodes<6> { static constexpr double x[7] = {-1.0, -0.83022389627856692987, -0.46884879347071421381, 0.0, 0.46884879347071421381, 0.83022389627856692987, 1.0}; };

template <int N>
struct ElemOps {
    float D[N + 1][N + 1];
    float Dw[N + 1][N + 1];
    float w[N + 1];
    constexpr ElemOps() : D(), Dw(), w()
    {
        const double *x = GllNodes<N>::x;
        double wd[N + 1] = {};
        for (int i = 0; i <= N; ++i) {
            if (i == 0 || i == N) { wd[i] = 2.0 / (N * (N + 1)); continue; }
            double p2 = 1.0, p1 = x[i], pn = 0.0;
            for (int k = 2; k <= N; ++k) { pn = ((2 * k - 1) * x[i] * p1 - (k - 1) * p2) / k; p2 = p1; p1 = pn; }
            wd[i] = 2.0 / (N * (N + 1) * pn * pn);
        }
        for (int a = 0; a <= N; ++a) {
            w[a] = (float)wd[a];
            for (int t = 0; t <= N; ++t) {
                double down = 1.0, v = 0.0;
                for (int i = 0; i <= N; ++i) if (i != t) down *= (x[t] - x[i]);
                for (int i = 0; i <= N; ++i) {
                    if (i == t) continue;
                    double L = 1.0;
                    for (int m = 0; m <= N; ++m) if (m != t && m != i) L *= (x[a] - x[m]);
                    v += L;
                }
                D[a][t] = (float)(v / down);
                Dw[a][t] = (float)(v / down * wd[a]);
            }
        }
    }
};

template <int N>
__device__ __forceinline__ constexpr float opD(int a, int t) { constexpr ElemOps<N> o{}; return o.D[a][t]; }
template <int N>
__device__ __forceinline__ constexpr float opDw(int a, int t) { constexpr ElemOps<N> o{}; return o.Dw[a][t]; }
template <int N>
__device__ __forceinline__ constexpr float opW(int a) { constexpr ElemOps<N> o{}; return o.w[a]; }

/* stress at one node from the 5x5 block C over {xx, zz, yz, xz, xy} and the
   2-D strains (uy,y = ux,y = uz,y = 0, Stress_ij sem2d_psvsh.c:1345-1379);
   C: 0:00 1:02 2:03 3:04 4:05 5:22 6:23 7:24 8:25 9:33 10:34 11:35 12:44 13:45 14:55 */
template <bool ORTHO>
__device__ __forceinline__ void stress5(const float *C, float e0, float e2, float e3, float e4, float e5,
                                        float &s0, float &s2, float &s3, float &s4, float &s5)
{
    if constexpr (ORTHO) {
        /* orthotropic in the x-y-z axes (isotropic, vertical or horizontal
           axis along x or z): C 2,3,4,6,7,8,10,11,13 are zero */
        s0 = C[0] * e0 + C[1] * e2;
        s2 = C[1] * e0 + C[5] * e2;
        s3 = C[9] * e3;
        s4 = C[12] * e4;
        s5 = C[14] * e5;
    } else {
        s0 = C[0] * e0 + C[1] * e2 + C[2] * e3 + C[3] * e4 + C[4] * e5;
        s2 = C[1] * e0 + C[5] * e2 + C[6] * e3 + C[7] * e4 + C[8] * e5;
        s3 = C[2] * e0 + C[6] * e2 + C[9] * e3 + C[10] * e4 + C[11] * e5;
        s4 = C[3] * e0 + C[7] * e2 + C[10] * e3 + C[12] * e4 + C[13] * e5;
        s5 = C[4] * e0 + C[8] * e2 + C[11] * e3 + C[13] * e4 + C[14] * e5;
    }
}

/* record entries that are zero for an orthotropic medium in these axes */
__host__ __device__ inline bool semsks_ortho_record(const float *C)
{
    const int z[9] = {2, 3, 4, 6, 7, 8, 10, 11, 13};
    for (int i = 0; i < 9; ++i)
        if (C[z[i]] != 0.f) return false;
    return true;
}

__device__ __forceinline__ int edge_index(int ix, int iz, int nx, int nz, int n)
{
    if (iz == 0) return ix;
    if (ix == 0) return nx + iz - 1;
    if (ix == nx - 1) return nx + (nz - 1) + iz - 1;
    if (iz == nz - 1 && ix >= 1 && ix <= 3 * n) return nx + 2 * (nz - 1) + ix - 1;
    return -1;
}

/* dt^2/M of a node from its material record and its position class: the
   lumped mass is rho * Je * Wx * Wz with Wx the GLL weight of the node's
   column inside its element(s) -- class 0: shared by two elements, 1..N-1:
   interior, N: the domain's first or last column (one element) */
template <int N>
__device__ __forceinline__ int mass_class(int i, int n_last)
{
    if (i == 0 || i == n_last) return N;
    const int r = i % N;
    return r;
}

template <int N>
__device__ __forceinline__ float mdt2_of(const InstDesc &D, int mid, int ix, int iz)
{
    constexpr int NC = N + 1;
    return gp(D.mclass)[(long)mid * NC * NC + mass_class<N>(iz, D.nz - 1) * NC + mass_class<N>(ix, D.nx - 1)];
}

/* u(t+dt) of one node, in place: u = u(t) with its word w, up = u(t-dt).
   The previous increment (u - up) - (c - cp) (c, cp the compensations of
   levels t and t-dt, both in w; true value u - c), the step's increment
   (interior: minus m f with m = dt^2/M; edge: ea * delta - eb * f, the
   Lysmer rows), then the Kahan-compensated sum; w becomes the word of
   level t+dt.  Each compensation is unpacked once. */
__device__ __forceinline__ void advance3(bool kahan, bool edge, float (&u)[3], uint32_t &w, const float (&up)[3],
                                         const float (&f)[3], float m, const float (&ea)[3], const float (&eb)[3])
{
    if (!kahan) {
#pragma unroll
        for (int k = 0; k < 3; ++k) {
            float d = u[k] - up[k];
            d = edge ? ea[k] * d - eb[k] * f[k] : d - m * f[k];
            u[k] = u[k] + d;
        }
        w = 0u;
        return;
    }
    int q[3];
#pragma unroll
    for (int k = 0; k < 3; ++k) {
        const float c = c_field(w, k, u[k]);
        float d = (u[k] - up[k]) - (c - c_field(w, k + 3, up[k]));
        d = edge ? ea[k] * d - eb[k] * f[k] : d - m * f[k];
        const float y = d - c;
        const float t = u[k] + y;
        q[k] = c_quant((t - u[k]) - y, t);
        u[k] = t;
    }
    w = c_pack3(q[0], q[1], q[2]) | (w & C_OWN) << 15;
}

/* u(t+dt) of one node: interior nodes (eidx < 0) by dt^2/M = m, edge nodes
   by their Lysmer factors (word wn of level t+dt) */
__device__ __forceinline__ void next_node(const InstDesc &D, bool kahan, int eidx, float m, const float (&f)[3],
                                          const float (&uc)[3], uint32_t w, const float (&up)[3],
                                          float (&nu)[3], uint32_t &wn)
{
    float ea[3] = {1.f, 1.f, 1.f}, eb[3] = {0.f, 0.f, 0.f};
    if (eidx >= 0) {
#pragma unroll
        for (int k = 0; k < 3; ++k) { ea[k] = gp(D.edge_a)[eidx * 3 + k]; eb[k] = gp(D.edge_b)[eidx * 3 + k]; }
    }
#pragma unroll
    for (int k = 0; k < 3; ++k) nu[k] = uc[k];
    wn = w;
    advance3(kahan, eidx >= 0, nu, wn, up, f, m, ea, eb);
}

/* one node: delta, Kahan-compensated u, edge rows by their Lysmer factors;
   uc: the node's u(t); p: plane of level t */
template <int N>
__device__ __forceinline__ void update_one(const InstDesc &D, int p, long q, int ix, int iz, int mid,
                                           const float (&f)[3], const float (&uc)[3])
{
    const int e = edge_index(ix, iz, D.nx, D.nz, D.hpn);
    float up[3];
    xyz_load(gp(D.u[(p + D.nlev - 1) % D.nlev]), q, up);
    const uint32_t w = D.kahan ? gp(D.comp[p])[q] : 0u;
    const float m = e < 0 ? mdt2_of<N>(D, mid, ix, iz) : 0.f;
    float nu[3];
    uint32_t wn;
    next_node(D, D.kahan != 0, e, m, f, uc, w, up, nu, wn);
    xyz_store(gp(D.u[(p + 1) % D.nlev]), q, nu);
    if (D.kahan) gp(D.comp[(p + 1) % D.nlev])[q] = wn;
}

/* Lane layout of the element-force/update kernel: a 64-lane wave holds
   NE = 64/(N+1) element columns side by side, N+1 lanes per element, one
   lane per element ROW (lane = e*(N+1) + b).  A lane keeps only its row of u
   and of the element force in registers; derivatives and tractions along x
   stay inside the lane, the ones along z go through small per-wave LDS
   blocks.  The wave sweeps its tile upward one element row per iteration. */
template <int N> struct RowLayout {
    static constexpr int NP = N + 1;
    static constexpr int NE = 64 / NP;              /* element columns per wave */
    static constexpr int NL = NE * NP;              /* active lanes */
    static constexpr int XCHUNK = NE - 1;           /* owned columns per tile (one halo column) */
    static constexpr int CB = (3 * NP + 3) / 4 * 4; /* floats per column block [k][t], 16-byte multiple */
    static constexpr int ESTR = NP * CB + 4;        /* floats per element block (u or q) */
    /* floats of a carried force row; an odd number of 16-byte groups keeps
       the 16-byte reads of different element columns on distinct banks */
    static constexpr int CFS = ((3 * NP + 3) / 4 | 1) * 4;
    static constexpr int WAVE_FLOATS = 2 * NE * ESTR + NE * CFS;
};

__host__ __device__ static inline int xchunk_of(int n) { return 64 / (n + 1) - 1; }

/* per-lane constants of the element row b the lane holds */
template <int N> struct RowCoef {
    float dz[N + 1];    /* D[b][t]   : z-derivative at row b from row t   */
    float az[N + 1];    /* Dw[t][b]  : z-divergence into row b from row t */
    float wb;           /* GLL weight of row b                            */
};

template <int N>
__device__ __forceinline__ RowCoef<N> row_coef(int b)
{
    RowCoef<N> c;
#pragma unroll
    for (int t = 0; t <= N; ++t) {
        float dz = 0.f, az = 0.f;
#pragma unroll
        for (int bb = 0; bb <= N; ++bb) {
            dz = (b == bb) ? opD<N>(bb, t) : dz;
            az = (b == bb) ? opDw<N>(t, bb) : az;
        }
        c.dz[t] = dz;
        c.az[t] = az;
    }
    float wb = 0.f;
#pragma unroll
    for (int bb = 0; bb <= N; ++bb) wb = (b == bb) ? opW<N>(bb) : wb;
    c.wb = wb;
    return c;
}

template <bool UNIFORM, bool ORTHO>
__device__ __forceinline__ void node_stress(const float *Cu, const SEMSKS_GMEM float *crec, float e0, float e2, float e3,
                                            float e4, float e5, float &s0, float &s2, float &s3, float &s4, float &s5)
{
    float C[15];
    if constexpr (UNIFORM) {
#pragma unroll
        for (int k = 0; k < 15; ++k) C[k] = Cu[k];
    } else {
        const SEMSKS_GMEM float4 *r = reinterpret_cast<const SEMSKS_GMEM float4 *>(crec);
        const float4 r0 = r[0], r1 = r[1], r2 = r[2], r3 = r[3];
        C[0] = r0.y; C[1] = r0.z; C[2] = r0.w; C[3] = r1.x; C[4] = r1.y; C[5] = r1.z; C[6] = r1.w;
        C[7] = r2.x; C[8] = r2.y; C[9] = r2.z; C[10] = r2.w; C[11] = r3.x; C[12] = r3.y; C[13] = r3.z;
        C[14] = r3.w;
    }
    stress5<ORTHO>(C, e0, e2, e3, e4, e5, s0, s2, s3, s4, s5);
}

/* one element row of K_e u_e (Stiff_Matrix: dUj_di_2D, Stress_ij, F_ik,
   Grid_Stiff) for the lane's row b: f[a][k] = force on node (a, b).  The
   element's LDS blocks are laid out [a][k][t] (CB floats per column a), so a
   lane reads column a of all rows with a few 16-byte loads: Ub holds the
   rows of u, Qb the column tractions. */
template <int N, bool UNIFORM, bool ORTHO>
__device__ __forceinline__ void row_force(const float (&u)[N + 1][3], float (&f)[N + 1][3], int b,
                                          const RowCoef<N> &rc, float *Ub, float *Qb, const float *Cu,
                                          const SEMSKS_GMEM float *mat, const SEMSKS_GMEM uint16_t *idrow,
                                          const SEMSKS_GMEM uint16_t *idn, bool wr)
{
    constexpr int NP = N + 1;
    constexpr int CB = RowLayout<N>::CB;
    if (wr) {
#pragma unroll
        for (int a = 0; a < NP; ++a)
#pragma unroll
            for (int k = 0; k < 3; ++k) Ub[a * CB + k * NP + b] = u[a][k];
    }
    __builtin_amdgcn_wave_barrier();
#pragma unroll
    for (int a = 0; a < NP; ++a) {
        float col[CB];
#pragma unroll
        for (int v = 0; v < CB / 4; ++v) {
            const float4 t4 = *reinterpret_cast<const float4 *>(Ub + a * CB + 4 * v);
            col[4 * v] = t4.x; col[4 * v + 1] = t4.y; col[4 * v + 2] = t4.z; col[4 * v + 3] = t4.w;
        }
        float gx0 = 0.f, gx1 = 0.f, gx2 = 0.f, gz0 = 0.f, gz1 = 0.f, gz2 = 0.f;
#pragma unroll
        for (int t = 0; t < NP; ++t) {
            const float dx = opD<N>(a, t);
            gx0 += dx * u[t][0];
            gx1 += dx * u[t][1];
            gx2 += dx * u[t][2];
            gz0 += rc.dz[t] * col[t];
            gz1 += rc.dz[t] * col[NP + t];
            gz2 += rc.dz[t] * col[2 * NP + t];
        }
        float s0, s2, s3, s4, s5;
        const SEMSKS_GMEM float *crec =
            UNIFORM ? nullptr : mat + (long)(a < N ? idrow[a] : idn[0]) * SEMSKS_MAT_STRIDE;
        node_stress<UNIFORM, ORTHO>(Cu, crec, gx0, gz2, gz1, gx2 + gz0, gx1, s0, s2, s3, s4, s5);
        const float wa = opW<N>(a);
        const float p0 = rc.wb * s0, p1 = rc.wb * s5, p2 = rc.wb * s4;   /* row traction    T_i0 */
#pragma unroll
        for (int k = 0; k < NP; ++k) {
            const float ax = opDw<N>(a, k);
            f[k][0] += ax * p0;
            f[k][1] += ax * p1;
            f[k][2] += ax * p2;
        }
        if (wr) {                                                        /* column traction T_i2 */
            Qb[a * CB + 0 * NP + b] = wa * s4;
            Qb[a * CB + 1 * NP + b] = wa * s3;
            Qb[a * CB + 2 * NP + b] = wa * s2;
        }
        __builtin_amdgcn_sched_barrier(0);
    }
    __builtin_amdgcn_wave_barrier();
#pragma unroll
    for (int a = 0; a < NP; ++a) {
        float col[CB];
#pragma unroll
        for (int v = 0; v < CB / 4; ++v) {
            const float4 t4 = *reinterpret_cast<const float4 *>(Qb + a * CB + 4 * v);
            col[4 * v] = t4.x; col[4 * v + 1] = t4.y; col[4 * v + 2] = t4.z; col[4 * v + 3] = t4.w;
        }
#pragma unroll
        for (int t = 0; t < NP; ++t) {
            f[a][0] += rc.az[t] * col[t];
            f[a][1] += rc.az[t] * col[NP + t];
            f[a][2] += rc.az[t] * col[2 * NP + t];
        }
    }
    __builtin_amdgcn_wave_barrier();
}

/* What a lane fetches from global memory for one element row: its row of
   u(t) with the packed compensations (columns 0..N-1, plus column N on
   lanes whose right neighbour lane does not hold it), and, for the update,
   the row of u(t-dt) and dt^2/M.  The next row's fetch is issued before the
   current row's stores, so waiting for it never waits for those stores. */
template <int N> struct RowFetch {
    float u[N][3];
    float un[3];
    uint32_t w[N];     /* compensation words of level t of the owned nodes */
    uint32_t wn;
    float pv[4][3];    /* per owned node: u(t-dt) */
    float m[4];
    int rid;           /* single-record tiles: the row's material record */
};

/* N consecutive nodes' xyz from a level plane (P at the first node); zero
   when off.  For N = 4 the 48 bytes start on a 16-byte boundary (rows are
   line-aligned and element columns 48 bytes apart). */
template <int N>
__device__ __forceinline__ void row_load(const SEMSKS_GMEM float *P, bool on, float (&v)[N][3])
{
    if constexpr (N == 4) {
        float4 t0 = make_float4(0.f, 0.f, 0.f, 0.f), t1 = t0, t2 = t0;
        if (on) {
            const SEMSKS_GMEM float4 *p4 = reinterpret_cast<const SEMSKS_GMEM float4 *>(P);
            t0 = p4[0]; t1 = p4[1]; t2 = p4[2];
        }
        v[0][0] = t0.x; v[0][1] = t0.y; v[0][2] = t0.z; v[1][0] = t0.w;
        v[1][1] = t1.x; v[1][2] = t1.y; v[2][0] = t1.z; v[2][1] = t1.w;
        v[2][2] = t2.x; v[3][0] = t2.y; v[3][1] = t2.z; v[3][2] = t2.w;
    } else {
#pragma unroll
        for (int a = 0; a < N; ++a)
#pragma unroll
            for (int k = 0; k < 3; ++k) v[a][k] = on ? P[3 * a + k] : 0.f;
    }
}

/* four nodes' xyz to a level plane (16-byte aligned, as row_load) */
__device__ __forceinline__ void row_store4(SEMSKS_GMEM float *P, const float (&v)[4][3])
{
    SEMSKS_GMEM float4 *p4 = reinterpret_cast<SEMSKS_GMEM float4 *>(P);
    p4[0] = make_float4(v[0][0], v[0][1], v[0][2], v[1][0]);
    p4[1] = make_float4(v[1][1], v[1][2], v[2][0], v[2][1]);
    p4[2] = make_float4(v[2][2], v[3][0], v[3][1], v[3][2]);
}

/* MIXED = false: every element row of the tile (over the wave's element
   columns) holds one material record, which the host has checked; the
   record sits in scalar registers.  MIXED = true: material per node. */
template <int N, bool FORCE_ONLY, bool MIXED, bool ORTHO>
__device__ void process_tile(const InstDesc &D, int p, int x0, int z0, int rows, int lane, float *lds)
{
    using RL = RowLayout<N>;
    constexpr int NP = RL::NP;
    const int nex = D.nex, nez = D.nez, bw = D.bw;
    const int e = lane / NP, b = lane - e * NP;
    const bool act = lane < RL::NL;
    const int ex = x0 - 1 + e;
    const bool valid = act && ex >= 0 && ex < nex;
    const bool own_x = valid && e >= 1;
    const bool last_x = ex == nex - 1;
    const bool direct = valid && (e == RL::NE - 1 || last_x);
    const bool kahan = D.kahan != 0;
    const int ec = act ? e : 0;
    float *Ub = lds + ec * RL::ESTR;
    float *Qb = lds + RL::NE * RL::ESTR + ec * RL::ESTR;
    float *CF = lds + 2 * RL::NE * RL::ESTR + ec * RL::CFS;
    const int exl = valid ? ex : 0;
    const int ex_first = x0 > 0 ? x0 - 1 : 0;
    const RowCoef<N> rc = row_coef<N>(b);
    /* storage columns of the lane's nodes 0..N-1 and of node N, the first
       node of the next element, which may open the next block */
    const uint32_t colb = (uint32_t)node_col(bw, D.nz, exl * N), colbn = (uint32_t)node_col(bw, D.nz, exl * N + N);

    const SEMSKS_GMEM float *U = gp(D.u[p]);
    const int pp = (p + D.nlev - 1) % D.nlev, pn = (p + 1) % D.nlev;   /* buffers of levels t-dt, t+dt */

    const int ez0 = z0 > 0 ? z0 - 1 : 0;
    const int ez1 = z0 + rows;

    /* share: lane b = 0 skips reading node row 0 of element row ez, the
       row that lane b = N of the row below already holds (take_top then
       moves it across lanes) */
    auto fetch = [&](int ez, RowFetch<N> &r, bool share) {
        const int iz = ez * N + b;
        const long rowoff = (long)(colb + (uint32_t)(iz * bw));
        const bool mem = !(share && b == 0);
        row_load<N>(U + 3 * rowoff, valid && mem, r.u);
        r.un[0] = r.un[1] = r.un[2] = 0.f;
        if (direct && mem) xyz_load(U, (long)(colbn + (uint32_t)(iz * bw)), r.un);
        r.rid = 0;
        if constexpr (!MIXED) r.rid = (int)cp(D.emat)[(long)ez * nex + ex_first];
        if constexpr (N == 4 && !FORCE_ONLY) {
            const bool own = own_x && ez >= z0 && b < N;
            const bool vec = own && iz != 0 && ex != 0;
            r.m[0] = r.m[1] = r.m[2] = r.m[3] = 0.f;
            uint4 w4 = make_uint4(0u, 0u, 0u, 0u);
            if (vec && kahan) w4 = *reinterpret_cast<const SEMSKS_GMEM uint4 *>(gp(D.comp[p]) + rowoff);
            r.w[0] = w4.x; r.w[1] = w4.y; r.w[2] = w4.z; r.w[3] = w4.w;
            row_load<4>(gp(D.u[pp]) + 3 * rowoff, vec, r.pv);
            if (vec) {
                const SEMSKS_GMEM uint16_t *idrow = gp(D.mat_id) + rowoff;
                /* mass class: cz = b (0 = row shared with the element below) */
#pragma unroll
                for (int a = 0; a < 4; ++a) {
                    const int mid = MIXED ? (int)idrow[a] : r.rid;
                    r.m[a] = gp(D.mclass)[(long)mid * 25 + b * 5 + a];
                }
            }
        }
    };
    auto take_top = [&](const RowFetch<N> &below, RowFetch<N> &r) {
#pragma unroll
        for (int a = 0; a < N; ++a) {
#pragma unroll
            for (int k = 0; k < 3; ++k) {
                const float v = __shfl(below.u[a][k], lane + N, 64);
                if (b == 0) r.u[a][k] = v;
            }
        }
#pragma unroll
        for (int k = 0; k < 3; ++k) {
            const float v = __shfl(below.un[k], lane + N, 64);
            if (b == 0) r.un[k] = v;
        }
    };

    /* carried raw top-row force of the element below, one block per element */
    if (act && b == N) {
#pragma unroll
        for (int i = 0; i < RL::CFS; ++i) CF[i] = 0.f;
    }
    __builtin_amdgcn_wave_barrier();

    /* single-record tiles fetch one row ahead; the mixed path, which holds
       per-node material records in registers, fetches at the top of a row */
    constexpr bool AHEAD = !MIXED;
    RowFetch<N> cur;
    fetch(ez0, cur, false);
    for (int ez = ez0; ez < ez1; ++ez) {
        if constexpr (!AHEAD) {
            if (ez > ez0) {
                RowFetch<N> nx;
                fetch(ez, nx, true);
                take_top(cur, nx);
                cur = nx;
            }
        }
        const int iz = ez * N + b;
        const bool top = ez == nez - 1;
        const bool own = own_x && ez >= z0 && (b < N || top);
        const long rowoff = (long)(colb + (uint32_t)(iz * bw)), offn = (long)(colbn + (uint32_t)(iz * bw));
        const SEMSKS_GMEM uint16_t *idrow = gp(D.mat_id) + rowoff, *idn = gp(D.mat_id) + offn;
        const int rid = cur.rid;
        /* this lane's row of u(t); column N from the right neighbour lane */
        float u[NP][3];
#pragma unroll
        for (int k = 0; k < 3; ++k) {
#pragma unroll
            for (int a = 0; a < N; ++a) u[a][k] = cur.u[a][k];
            const float nb = __shfl(cur.u[0][k], lane + NP, 64);
            u[N][k] = direct ? cur.un[k] : nb;
        }
        float f[NP][3];
#pragma unroll
        for (int a = 0; a < NP; ++a)
#pragma unroll
            for (int k = 0; k < 3; ++k) f[a][k] = 0.f;
        if constexpr (!MIXED) {
            const SEMSKS_CMEM float *r = cp(D.mat) + (long)rid * SEMSKS_MAT_STRIDE;
            float C[15];
#pragma unroll
            for (int k = 0; k < 15; ++k) C[k] = r[1 + k];
            row_force<N, true, ORTHO>(u, f, b, rc, Ub, Qb, C, gp(D.mat), idrow, idn, act);
        } else {
            row_force<N, false, ORTHO>(u, f, b, rc, Ub, Qb, nullptr, gp(D.mat), idrow, idn, act);
        }
        if (!valid) {
#pragma unroll
            for (int a = 0; a < NP; ++a)
#pragma unroll
                for (int k = 0; k < 3; ++k) f[a][k] = 0.f;
        }
        /* z assembly: the bottom row adds the raw top row of the element
           below, then the raw top row is handed up through LDS */
        if (b == 0) {
            float cf[RL::CFS];
#pragma unroll
            for (int v = 0; v < (3 * NP + 3) / 4; ++v) {
                const float4 t4 = *reinterpret_cast<const float4 *>(CF + 4 * v);
                cf[4 * v] = t4.x; cf[4 * v + 1] = t4.y; cf[4 * v + 2] = t4.z; cf[4 * v + 3] = t4.w;
            }
#pragma unroll
            for (int a = 0; a < NP; ++a)
#pragma unroll
                for (int k = 0; k < 3; ++k) f[a][k] += cf[a * 3 + k];
        }
        __builtin_amdgcn_wave_barrier();
        if (act && b == N) {
#pragma unroll
            for (int a = 0; a < NP; ++a)
#pragma unroll
                for (int k = 0; k < 3; ++k) CF[a * 3 + k] = f[a][k];
        }
        __builtin_amdgcn_wave_barrier();
        /* x assembly: the left column adds the right column of the element
           to the left */
#pragma unroll
        for (int k = 0; k < 3; ++k) {
            const float s = __shfl(f[N][k], lane - NP, 64);
            if (e > 0) f[0][k] += s;
        }
        /* next row's loads go out before this row's stores */
        RowFetch<N> nxt;
        if constexpr (AHEAD) {
            __builtin_amdgcn_sched_barrier(0);
            if (ez + 1 < ez1) {
                fetch(ez + 1, nxt, true);
                take_top(cur, nxt);
            }
            __builtin_amdgcn_sched_barrier(0);
        }
        if (own) {
            const int ixb = ex * N;
            const long q0 = rowoff;
            const int a1 = last_x ? NP : N;
            const bool vec = N == 4 && b < N && iz != 0 && ex != 0;
            if constexpr (FORCE_ONLY) {
#pragma unroll
                for (int a = 0; a < NP; ++a)
                    if (a < a1)
#pragma unroll
                        for (int k = 0; k < 3; ++k) gp(D.fout[k])[a < N ? q0 + a : offn] = f[a][k];
            } else if (N == 4 && vec) {
                float nv[4][3];
                uint32_t nw[4];
                const float one[3] = {1.f, 1.f, 1.f}, zero[3] = {0.f, 0.f, 0.f};
#pragma unroll
                for (int a = 0; a < 4; ++a) {
#pragma unroll
                    for (int k = 0; k < 3; ++k) nv[a][k] = u[a][k];
                    nw[a] = cur.w[a];
                    advance3(kahan, false, nv[a], nw[a], cur.pv[a], f[a], cur.m[a], one, zero);
                }
                row_store4(gp(D.u[pn]) + 3 * q0, nv);
                if (kahan)
                    *reinterpret_cast<SEMSKS_GMEM uint4 *>(gp(D.comp[pn]) + q0) = make_uint4(nw[0], nw[1], nw[2], nw[3]);
                if (last_x) {
                    float fa[3] = {f[N][0], f[N][1], f[N][2]};
                    float ua[3] = {u[N][0], u[N][1], u[N][2]};
                    update_one<N>(D, p, offn, ixb + N, iz, MIXED ? (int)idn[0] : rid, fa, ua);
                }
            } else {
#pragma unroll
                for (int a = 0; a < NP; ++a) {
                    if (a >= a1) continue;
                    float fa[3] = {f[a][0], f[a][1], f[a][2]};
                    float ua[3] = {u[a][0], u[a][1], u[a][2]};
                    update_one<N>(D, p, a < N ? q0 + a : offn, ixb + a, iz,
                                  MIXED ? (int)(a < N ? idrow[a] : idn[0]) : rid, fa, ua);
                }
            }
        }
        if constexpr (AHEAD) cur = nxt;
    }
}

/* Blocks of a launch are dealt round-robin to the 8 XCDs of an MI355X
   (block b runs on XCD b % 8). */
constexpr int NXCD = SEMSKS_NQ;

/* one tile of the step kernel: (slot, x0, z0, rows | mixed << 16) */
template <int N, bool FORCE_ONLY>
__device__ __forceinline__ void step_tile(const InstDesc *__restrict__ insts, const int *__restrict__ levels, int4 tile,
                                          int lane, float *lds)
{
    tile.x = __builtin_amdgcn_readfirstlane(tile.x);
    tile.y = __builtin_amdgcn_readfirstlane(tile.y);
    tile.z = __builtin_amdgcn_readfirstlane(tile.z);
    tile.w = __builtin_amdgcn_readfirstlane(tile.w);
    const InstDesc &D = insts[tile.x];
    if (!D.active || D.hpn != N) return;
    int p = 0;
    if constexpr (!FORCE_ONLY) {
        const int it = levels[tile.x];
        if (it >= D.nst - 1) return;
        p = it % D.nlev;
    }
    const int rows = tile.w & 0xFFFF;
    if (D.ortho) {
        if (tile.w >> 16) process_tile<N, FORCE_ONLY, true, true>(D, p, tile.y, tile.z, rows, lane, lds);
        else process_tile<N, FORCE_ONLY, false, true>(D, p, tile.y, tile.z, rows, lane, lds);
    } else {
        if (tile.w >> 16) process_tile<N, FORCE_ONLY, true, false>(D, p, tile.y, tile.z, rows, lane, lds);
        else process_tile<N, FORCE_ONLY, false, false>(D, p, tile.y, tile.z, rows, lane, lds);
    }
}

/* Persistent grid over the tile list (slot, x0, z0, rows | mixed << 16); a
   mixed tile takes the per-node material path.  The list is cut into NXCD
   contiguous runs, one per XCD, so that consecutive tiles -- the
   neighbours of one instance, which read the same 128-byte lines along
   their shared edges and stream one storage block -- run at the same time
   under the same L2.
   Time steps with more tiles than waves (qctr != nullptr): each run is a
   queue with its counter in qctr[q * SEMSKS_QSTRIDE] (zero at launch:
   post_kernel clears them).  A wave takes the next tile of its own XCD's
   queue, which keeps neighbours together however far the waves drift
   apart; a wave whose queue is empty takes from the next queues in turn
   and leaves once all NXCD are exhausted.  Otherwise the waves of an XCD
   deal its run statically (local wave w takes w, w + W, ...). */
template <int N, bool FORCE_ONLY>
__global__ __launch_bounds__(256, N <= 4 ? 3 : 2) void step_kernel(const InstDesc *__restrict__ insts, const int *__restrict__ levels,
                                                      const int4 *__restrict__ tiles, const int *__restrict__ ntiles,
                                                      int *__restrict__ qctr)
{
    __shared__ float lds_blocks[4][RowLayout<N>::WAVE_FLOATS];
    const int lane = threadIdx.x & 63;
    float *lds = lds_blocks[threadIdx.x >> 6];
    const int wpb = blockDim.x >> 6;
    const int nt = *ntiles;
    const bool split = gridDim.x % NXCD == 0;
    const int nq = split ? NXCD : 1;
    const int q0 = split ? blockIdx.x % NXCD : 0;
    const int W = (split ? gridDim.x / NXCD : gridDim.x) * wpb;        /* waves per run */
    const bool queued = qctr != nullptr && nt > W * nq;
    int qi = 0;
    /* run q is [nt * q / nq, nt * (q + 1) / nq) (nt < 2^28: one sweep holds at most a few million tiles) */
    int t = nt * q0 / nq + (split ? blockIdx.x / NXCD : blockIdx.x) * wpb + (threadIdx.x >> 6) - W;
    t = __builtin_amdgcn_readfirstlane(t);
    const int shi = nt * (q0 + 1) / nq;
    for (;;) {
        if (queued) {
            t = -1;
            while (qi < nq) {
                const int q = (q0 + qi) % nq;
                const int lo = nt * q / nq, hi = nt * (q + 1) / nq;
                int k = 0;
                if (lane == 0) k = atomicAdd(qctr + q * SEMSKS_QSTRIDE, 1);
                k = __builtin_amdgcn_readfirstlane(k);
                if (lo + k < hi) {
                    t = lo + k;
                    break;
                }
                ++qi;
            }
            if (t < 0) break;
        } else {
            t += W;
            if (t >= shi) break;
        }
        step_tile<N, FORCE_ONLY>(insts, levels, tiles[t], lane, lds);
    }
}

/* ========================================================================= */
/* closed-form plane wave records (stand-in for the GRT producer)            */
/* ========================================================================= */
__device__ static double node_coord(int i, int nel, double eh, int n, const double *xi)
{
    int e = i / n;
    if (e > nel - 1) e = nel - 1;
    int a = i - e * n;
    return e * eh + (xi[a] + 1.0) / 2.0 * eh;
}

__global__ void planewave_kernel(int nex, int nez, int n, int neh, double eh, int nt, double dp, double p, double beta,
                                 double f0, double psi, float *psv, float *sh)
{
    double xi[SEMSKS_MAXNP], w[SEMSKS_MAXNP];
    gll64(n, xi, w);
    const int nx = nex * n + 1, nz = nez * n + 1, nbl = neh * n;
    const int mt = 2 * (nx + nz), sht = nx + nz;
    const double s = p * beta, ci = sqrt(1.0 - s * s), eta = ci / beta, lead = 1.2 / f0;
    const double pf = (M_PI * f0) * (M_PI * f0), cp = cos(psi * (M_PI / 180.0)), sp = sin(psi * (M_PI / 180.0));
    const long total = (long)nt * sht;
    for (long id = blockIdx.x * (long)blockDim.x + threadIdx.x; id < total; id += (long)gridDim.x * blockDim.x) {
        const int k = (int)(id / sht), j = (int)(id % sht);
        double x, z;
        if (j < nx) { x = node_coord(j, nex, eh, n, xi); z = node_coord(nbl, nez, eh, n, xi); }
        else { x = node_coord(nbl, nex, eh, n, xi); z = node_coord(j - nx, nez, eh, n, xi); }
        const double t = k * dp;
        const double tau = t - lead - (p * x + eta * z);
        const double a = pf * tau * tau;
        const double r = (1.0 - 2.0 * a) * exp(-a);
        const double sv = r * cp;
        const float uu = (float)(sv * ci), wd = (float)(-(-sv * s)), vv = (float)(r * sp);
        float *rp = psv + (long)k * mt;
        if (j < nx) { rp[j] = uu; rp[nx + j] = wd; }
        else { rp[2 * nx + (j - nx)] = uu; rp[2 * nx + nz + (j - nx)] = wd; }
        sh[(long)k * sht + j] = vv;
    }
}

/* ========================================================================= */
/* fixed device scratch of semsks_stiff_dev                                  */
/* ========================================================================= */
/* The field and force of semsks_stiff_dev in the kernels' node order live
   in this array, one copy per GPU, held until the process ends: an entry
   point that allocates nothing.  Host code never passes the array's name
   to a kernel (on the host it names a stand-in, not device memory); it uses
   the per-device address from hipGetSymbolAddress (scratch_ptr). */
static __device__ __align__(256) unsigned char scratch[SEMSKS_SCRATCH_BYTES];

/* elements of the force-only tile height used by semsks_stiff_dev */
constexpr int FORCE_TZ = 8;

__global__ void fill_kernel(float *p, size_t n, float v)
{
    for (size_t i = blockIdx.x * (size_t)blockDim.x + threadIdx.x; i < n; i += (size_t)gridDim.x * blockDim.x) p[i] = v;
}

/* SoA planes [3][nz][nx] -> xyz per node in the kernels' node order */
__global__ void soa_to_blocks_kernel(int nx, int nz, int bw, const float *u, float *aos)
{
    const size_t N = (size_t)nx * nz;
    for (size_t q = blockIdx.x * (size_t)blockDim.x + threadIdx.x; q < N; q += (size_t)gridDim.x * blockDim.x) {
        const int iz = (int)(q / nx), ix = (int)(q - (size_t)iz * nx);
        const long n = node_at(bw, nz, ix, iz);
        aos[3 * n] = u[q];
        aos[3 * n + 1] = u[N + q];
        aos[3 * n + 2] = u[2 * N + q];
    }
}

/* three force planes in the kernels' node order -> SoA planes [3][nz][nx] */
__global__ void blocks_to_soa_kernel(int nx, int nz, int bw, size_t nodes, const float *f3, float *out)
{
    const size_t N = (size_t)nx * nz;
    for (size_t q = blockIdx.x * (size_t)blockDim.x + threadIdx.x; q < N; q += (size_t)gridDim.x * blockDim.x) {
        const int iz = (int)(q / nx), ix = (int)(q - (size_t)iz * nx);
        const long n = node_at(bw, nz, ix, iz);
        out[q] = f3[n];
        out[N + q] = f3[nodes + n];
        out[2 * N + q] = f3[2 * nodes + n];
    }
}

/* scratch layout for a plane of `nodes` nodes: u xyz at 0, the three force
   planes after it (256-byte aligned); total bytes */
static size_t scratch_f_offset(size_t nodes) { return (nodes * 12 + 255) / 256 * 256; }
static size_t scratch_bytes(size_t nodes) { return scratch_f_offset(nodes) + nodes * 12; }

#define SEMSKS_MAX_DEVICES 64
static std::mutex g_scratch_mu;
static void *g_scratch_ptr[SEMSKS_MAX_DEVICES];

/* device address of this GPU's scratch array, looked up once per device
   with that device current (the engine's set-up) */
static hipError_t scratch_ptr(int device, void **out)
{
    *out = nullptr;
    if (device < 0 || device >= SEMSKS_MAX_DEVICES) return hipErrorInvalidDevice;
    std::lock_guard<std::mutex> lk(g_scratch_mu);
    if (!g_scratch_ptr[device]) {
        void *p = nullptr;
        const hipError_t err = hipGetSymbolAddress(&p, HIP_SYMBOL(scratch));
        if (err != hipSuccess) return err;
        g_scratch_ptr[device] = p;
    }
    *out = g_scratch_ptr[device];
    return hipSuccess;
}

/* ========================================================================= */
/* engine                                                                    */
/* ========================================================================= */
namespace {

struct Slot {
    bool used = false;
    int share = -1;           /* slot whose mesh/material this one uses */
    int refs = 0;             /* slots sharing this one's material */
    semsks_problem meta{};
    InstDesc d{};
    int nx = 0, nz = 0, bw = 0, nt_sem = 0, nlev = 0;
    size_t plane = 0;                 /* nodes per level plane (storage blocks, padded) */
    size_t bytes = 0;
    std::vector<void *> allocs;       /* owned device allocations */
    std::vector<uint16_t> emat;       /* host copy of the element-uniform record ids (owner) */
    uint64_t mat_hash = 0;            /* material_hash of the problem (owner) */
    /* semsks_stiff_dev (owner): its force-only descriptor (field and force
       in the scratch array), tile count and tiles, in one device block;
       null when the mesh does not fit the scratch */
    InstDesc *d_fdesc = nullptr;
    int *d_fcount = nullptr;
    int4 *d_ftiles = nullptr;
};

/* Tiling of one mesh.  A tile (slot, x0, z0, rows | mixed << 16) may run
   the single-record path when every element row it sweeps -- its own rows
   and the halo row below -- holds one material record over the wave's
   element columns (one halo column to the left).  Each column of tiles is
   cut along z so that rows that are not single-record (a layer cut inside
   an element row) get short per-node-material tiles of their own and the
   rest stay single-record tiles of up to tz rows.  A run of R
   single-record rows is cut into ceil(R / tz) tiles whose heights differ
   by at most one row (100 rows at tz = 8: 9 tiles of 8 and 4 of 7, not 12
   of 8 and a short one of 4), so no wave of the persistent grid is left
   with a tile much shorter than the others. */

/* A run of rows of one tile column: single-record (every row from its
   first, with the halo row below the first, holds one record) or
   per-node-material. */
struct ZRun { int x0, z, len, mixed; };

/* the runs of every tile column of one mesh (owner slot o), in x then z
   order; they do not depend on the tile height */
void tile_runs(const Slot &o, std::vector<ZRun> &out)
{
    const int nex = o.meta.nex, nez = o.meta.nez;
    const int xc = xchunk_of(o.meta.hpn), ncols = 64 / (o.meta.hpn + 1);
    std::vector<char> single(nez);
    for (int x0 = 0; x0 < nex; x0 += xc) {
        const int ex0 = std::max(x0 - 1, 0), ex1 = std::min(nex, x0 - 1 + ncols);
        for (int ez = 0; ez < nez; ++ez) {
            const uint16_t r = o.emat[(size_t)ez * nex + ex0];
            bool ok = r != 0xFFFF;
            for (int ex = ex0 + 1; ok && ex < ex1; ++ex) ok = o.emat[(size_t)ez * nex + ex] == r;
            single[ez] = ok;
        }
        /* a single-record tile may start at z when rows z - 1 and z are */
        auto starts = [&](int z) { return (z == 0 || single[z - 1]) && single[z]; };
        for (int z = 0; z < nez;) {
            int len = 0;
            if (starts(z)) {
                while (z + len < nez && single[z + len]) ++len;
                out.push_back({x0, z, len, 0});
            } else {
                while (z + len < nez && !starts(z + len)) ++len;
                out.push_back({x0, z, len, 1});
            }
            z += len;
        }
    }
}

/* cuts runs into tiles of up to tz rows: emit(x0, z0, rows, mixed) */
template <class F>
void cut_runs(const std::vector<ZRun> &runs, int tz, F emit)
{
    for (const ZRun &r : runs) {
        const int k = (r.len + tz - 1) / tz;
        for (int i = 0, z = r.z; i < k; ++i) {
            /* single-record runs in heights that differ by at most one
               row; per-node-material runs (rarely longer than a row or
               two) in rows of tz and the rest */
            const int h = r.mixed ? std::min(tz, r.z + r.len - z) : r.len / k + (i < r.len % k ? 1 : 0);
            emit(r.x0, z, h, r.mixed);
            z += h;
        }
    }
}

/* walks the tiling of one mesh (owner slot o) at tile height tz */
template <class F>
void walk_tiles(const Slot &o, int tz, F emit)
{
    std::vector<ZRun> runs;
    tile_runs(o, runs);
    cut_runs(runs, tz, emit);
}

void build_tiles(const Slot &o, int slot, int tz, std::vector<int4> &t)
{
    walk_tiles(o, tz, [&](int x0, int z0, int rows, int mixed) {
        t.push_back(make_int4(slot, x0, z0, rows | (mixed << 16)));
    });
}

template <class T>
hipError_t dalloc(Slot &s, T **p, size_t n)
{
    void *q = nullptr;
    hipError_t e = hipMalloc(&q, n * sizeof(T) + 256);
    if (e == hipSuccess) { s.allocs.push_back(q); s.bytes += n * sizeof(T); }
    *p = (T *)q;
    return e;
}

}  // namespace

struct semsks_engine {
    semsks_engine_opts opts{};
    hipStream_t stream = nullptr;
    int device = 0, n_cu = 0;
    size_t budget = 0, used = 0;
    std::vector<Slot> slots;
    InstDesc *d_desc = nullptr;
    int *d_levels = nullptr;
    int4 *d_tiles = nullptr;
    int *d_ntiles = nullptr;          /* [7]: per hpn 3..6 (index hpn) */
    int *d_qctr = nullptr;            /* tile-queue counters, SEMSKS_QCOUNTERS lines */
    int *d_pctr = nullptr;            /* post_kernel: finished blocks per slot */
    std::vector<int> h_levels;
    int tile_cap = 0;
    int tile_off[SEMSKS_MAXNP] = {0}; /* offset of each hpn's tiles */
    bool dirty = true;
    /* captured steps: graph[i] holds 16 >> i steps (16, 8, 4, 2, 1), all
       built for the same instance set (the graph_* fields below) */
    hipGraph_t graph[5] = {};
    hipGraphExec_t gexec[5] = {};
    unsigned hpn_mask = 0, graph_mask = 0;
    bool any_sponge = false, graph_sponge = false;
    /* post_kernel blocks per slot (upload_tables): one thread per interface
       node and receiver of the largest slot, 16 to 64 blocks, above 16 only
       while the grid stays within 16k blocks (it covers every slot, so that
       captured steps stay valid as slots are admitted; any count is correct,
       the kernel strides over its work).  One C2 instance: 0.2195 against
       0.2262 ms per step with 16 */
    int post_bx = 16;
    int graph_off[SEMSKS_MAXNP] = {0};  /* tile offsets the captured step was built with */
    hipEvent_t ev[4] = {nullptr, nullptr, nullptr, nullptr};
    std::vector<hipEvent_t> tev;      /* per-step timing events */
    int timed_steps = 0;
    double last_ms[2] = {0, 0};
    /* tile height in element rows: opts.tile_z when set, else the
       autotuner's choice for the current set of instances */
    int tz_cur = 8;
    bool tuning = false;               /* the autotuner is trying heights */
    bool tuned = false;
    std::map<uint64_t, int> tz_cache;   /* instance-set signature -> tile height */
    /* admissions run on their own stream, so that their device work (FP64
       setup, plane-wave records, boundary series, uploads) and the host
       waits for it overlap a step round already queued on `stream`; their
       temporaries live in one engine-owned scratch that only grows (no
       hipFree, which would wait for the running round) */
    hipStream_t setup = nullptr;
    unsigned char *sbuf = nullptr;    /* admission scratch (counted in used) */
    size_t scap = 0;
    /* device-resident plane-wave records waiting for the next admit: the
       first pend_bytes of the scratch */
    float *pend_psv = nullptr, *pend_sh = nullptr;
    int pend_nt = 0, pend_nx = 0, pend_nz = 0;
    size_t pend_bytes = 0;
    void *scratch = nullptr;          /* this device's scratch array (semsks_stiff_dev) */
};

static semsks_status alloc_failed(hipError_t err, const char *what);

static size_t scratch_growth(const semsks_engine *e, size_t bytes)
{
    return bytes > e->scap ? bytes - e->scap : 0;
}

/* makes the admission scratch at least `bytes` long, keeping its first
   `keep` bytes (pending plane-wave records); it only grows, and growing
   waits for the device (hipFree), which a serving loop meets only while
   the largest request so far is admitted */
static semsks_status setup_scratch(semsks_engine *e, size_t bytes, size_t keep)
{
    if (bytes <= e->scap) return SEMSKS_OK;
    size_t cap = std::max(bytes, e->scap + e->scap / 2);
    if (e->used - e->scap + cap > e->budget) cap = bytes;
    if (e->used - e->scap + cap > e->budget)
        return semsks_set_error(SEMSKS_ENOHBM, "admission scratch of %zu bytes does not fit the budget", cap);
    unsigned char *nb = nullptr;
    const hipError_t err = hipMalloc(&nb, cap);
    if (err != hipSuccess) return alloc_failed(err, "admission scratch");
    if (keep && e->sbuf) HIPCHK(hipMemcpyAsync(nb, e->sbuf, keep, hipMemcpyDeviceToDevice, e->setup));
    HIPCHK(hipStreamSynchronize(e->setup));
    if (e->sbuf) HIPCHK(hipFree(e->sbuf));
    if (e->pend_psv) {
        e->pend_sh = (float *)(nb + ((unsigned char *)e->pend_sh - e->sbuf));
        e->pend_psv = (float *)nb;
    }
    e->used += cap - e->scap;
    e->sbuf = nb;
    e->scap = cap;
    return SEMSKS_OK;
}

static int tile_rows(const semsks_engine *e, int nez)
{
    int tz = e->opts.tile_z > 0 ? e->opts.tile_z : e->tz_cur;
    return std::min(tz, std::max(1, nez));
}

static void drop_graph(semsks_engine *e);
static int default_tile_rows(const semsks_engine *e);

/* What the best tile height depends on: degree mix, how many instances
   share the step and how many element columns and rows they have (both
   rounded to powers of two). */
static uint64_t tune_signature(const semsks_engine *e)
{
    uint64_t nsl = 0, ncol = 0, nrow = 0;
    for (auto &s : e->slots)
        if (s.used) { ++nsl; ncol += s.meta.nex; nrow += s.meta.nez; }
    auto lg = [](uint64_t v) { uint64_t l = 0; while (v > 1) { v >>= 1; ++l; } return l; };
    return (uint64_t)e->hpn_mask | (lg(nsl) << 8) | (lg(ncol / std::max<uint64_t>(nsl, 1)) << 16) |
           (lg(nrow / std::max<uint64_t>(nsl, 1)) << 24);
}

static semsks_status upload_tables(semsks_engine *e)
{
    /* descriptors */
    std::vector<InstDesc> h(e->slots.size());
    for (size_t i = 0; i < e->slots.size(); ++i) {
        h[i] = e->slots[i].d;
        h[i].active = e->slots[i].used ? 1 : 0;
    }
    HIPCHK(hipMemcpyAsync(e->d_desc, h.data(), h.size() * sizeof(InstDesc), hipMemcpyHostToDevice, e->stream));
    e->any_sponge = false;
    for (auto &s : e->slots) e->any_sponge |= s.used && s.d.atten > 0;
    {
        long mw = 0;
        for (auto &s : e->slots)
            if (s.used) {
                const int nbl = s.meta.neh * s.meta.hpn;
                mw = std::max<long>(mw, (long)(s.nx - nbl) + (s.nz - nbl - 1) + s.meta.ntrace);
            }
        const long by_work = (mw + 255) / 256;
        const long by_grid = 16384 / (long)std::max<size_t>(e->slots.size(), 1);
        e->post_bx = (int)std::max<long>(16, std::min<long>(64, std::min(by_work, by_grid)));
    }
    /* tiles, grouped by hpn, mixed (per-node material) tiles first since
       they take longest, then by slot, x and z: the waves of one XCD, which
       take consecutive tiles (see step_kernel), sweep the tiles of one
       column of one instance -- one storage block, contiguous in every
       level plane -- and its neighbour columns at the same time, so that
       the open DRAM pages and the lines along shared edges are few.  On the
       64-wave 2000x500 batch this order ran the step kernel 2% faster than
       z-major (z, slot, x) order, the single instance 5% faster */
    std::vector<int4> t;
    int cnt[SEMSKS_MAXNP] = {0};
    e->hpn_mask = 0;
    const int tzg = e->opts.tile_z > 0 ? e->opts.tile_z : e->tz_cur;
    for (int n = 3; n <= 6; ++n) {
        e->tile_off[n] = (int)t.size();
        std::vector<int4> tn, own;
        std::map<int, std::vector<int4>> by_owner;   /* owner slot -> its tiling (slot field = owner) */
        for (size_t k = 0; k < e->slots.size(); ++k) {
            const Slot &s = e->slots[k];
            if (!s.used || s.meta.hpn != n) continue;
            const int ow = s.share >= 0 ? s.share : (int)k;
            auto it = by_owner.find(ow);
            if (it == by_owner.end()) {
                own.clear();
                build_tiles(e->slots[ow], ow, tzg, own);
                it = by_owner.emplace(ow, own).first;
            }
            for (int4 q : it->second) tn.push_back(make_int4((int)k, q.y, q.z, q.w));
        }
        std::stable_sort(tn.begin(), tn.end(), [](const int4 &a, const int4 &b) {
            const int ma = a.w >> 16, mb = b.w >> 16;
            if (ma != mb) return ma > mb;
            if (a.x != b.x) return a.x < b.x;
            if (a.y != b.y) return a.y < b.y;
            return a.z < b.z;
        });
        t.insert(t.end(), tn.begin(), tn.end());
        cnt[n] = (int)t.size() - e->tile_off[n];
        if (cnt[n]) e->hpn_mask |= 1u << n;
    }
    if (t.size() >= (1u << 28))       /* step_kernel's run bounds nt * q / NXCD are 32-bit */
        return semsks_set_error(SEMSKS_EINVAL, "tile list of %zu tiles is too long", t.size());
    if ((int)t.size() > e->tile_cap) {
        drop_graph(e);                /* a captured step holds the old list */
        HIPCHK(hipStreamSynchronize(e->stream));
        if (e->d_tiles) HIPCHK(hipFree(e->d_tiles));
        e->tile_cap = (int)t.size() + 1024;
        HIPCHK(hipMalloc(&e->d_tiles, e->tile_cap * sizeof(int4)));
    }
    if (!t.empty())
        HIPCHK(hipMemcpyAsync(e->d_tiles, t.data(), t.size() * sizeof(int4), hipMemcpyHostToDevice, e->stream));
    HIPCHK(hipMemcpyAsync(e->d_ntiles, cnt, sizeof cnt, hipMemcpyHostToDevice, e->stream));
    HIPCHK(hipStreamSynchronize(e->stream));
    e->dirty = false;
    if (e->opts.tile_z <= 0) {
        auto it = e->tz_cache.find(tune_signature(e));
        e->tuned = it != e->tz_cache.end();
        if (e->tuned && it->second != tzg) {
            e->tz_cur = it->second;
            return upload_tables(e);
        }
        if (!e->tuned && !e->tuning) {
            const int d = default_tile_rows(e);
            if (d != tzg) {
                e->tz_cur = d;
                return upload_tables(e);
            }
        }
    }
    return SEMSKS_OK;
}


template <bool FORCE_ONLY> const void *step_fn(int n)
{
    switch (n) {
    case 3: return reinterpret_cast<const void *>(&step_kernel<3, FORCE_ONLY>);
    case 4: return reinterpret_cast<const void *>(&step_kernel<4, FORCE_ONLY>);
    case 5: return reinterpret_cast<const void *>(&step_kernel<5, FORCE_ONLY>);
    default: return reinterpret_cast<const void *>(&step_kernel<6, FORCE_ONLY>);
    }
}

/* persistent grid: as many 256-thread blocks as the kernel keeps resident on
   every CU (or waves_per_cu / 4 per CU when the option is set) */
template <bool FORCE_ONLY>
static int persistent_blocks(const semsks_engine *e, int n)
{
    if (e->opts.waves_per_cu > 0) return std::max(1, e->n_cu * e->opts.waves_per_cu / 4);
    int per_cu = 0;
    if (hipOccupancyMaxActiveBlocksPerMultiprocessor(&per_cu, step_fn<FORCE_ONLY>(n), 256, 0) != hipSuccess || per_cu < 1)
        per_cu = 2;
    return e->n_cu * per_cu;
}

/* Predicted duration of one step at tile height tz, in element-row sweeps
   of one wave.  Each degree's launch deals its tiles to the earliest-free
   of its W waves (the persistent grid's queues); a tile costs its rows,
   0.7 for its halo row (no update), 0.5 for its start (the first fetch),
   and a per-node-material tile 1.5 times that.  Tiles are dealt longest
   first, whole rounds of W at a time, the rest one by one.  The model
   ranks heights for the autotuner and picks the height of calls too short
   to tune: short tiles keep the grid of a small mesh busy and its sweeps
   short, tall tiles read fewer halo rows, and a height whose tile count
   just fills a whole number of rounds of the grid leaves no wave idle at
   the end of the step. */
struct TileModel {
    struct Mesh { int hpn; long copies; std::vector<ZRun> runs; };
    std::vector<Mesh> meshes;       /* one per owner, with the slots that share it */
    long waves[SEMSKS_MAXNP] = {0}; /* persistent grid of each degree's launch */
    int maxz = 1;
};

static TileModel tile_model(const semsks_engine *e)
{
    TileModel m;
    std::map<int, size_t> at;
    for (size_t k = 0; k < e->slots.size(); ++k) {
        const Slot &s = e->slots[k];
        if (!s.used) continue;
        const int ow = s.share >= 0 ? s.share : (int)k;
        auto it = at.find(ow);
        if (it == at.end()) {
            TileModel::Mesh me{s.meta.hpn, 0, {}};
            tile_runs(e->slots[ow], me.runs);
            it = at.emplace(ow, m.meshes.size()).first;
            m.meshes.push_back(std::move(me));
        }
        ++m.meshes[it->second].copies;
        m.maxz = std::max(m.maxz, s.meta.nez);
        if (!m.waves[s.meta.hpn]) m.waves[s.meta.hpn] = 4L * persistent_blocks<false>(e, s.meta.hpn);
    }
    return m;
}

static double predict_step(const TileModel &m, int tz)
{
    double total = 0.0;
    std::vector<double> fin;
    for (int n = 3; n <= 6; ++n) {
        std::map<double, long> cnt;     /* tile cost -> tiles */
        for (const auto &me : m.meshes) {
            if (me.hpn != n) continue;
            cut_runs(me.runs, tz, [&](int, int z0, int rows, int mixed) {
                cnt[(rows + (z0 > 0 ? 0.7 : 0.0) + 0.5) * (mixed ? 1.5 : 1.0)] += me.copies;
            });
        }
        if (cnt.empty()) continue;
        const long W = m.waves[n];
        fin.assign((size_t)W, 0.0);
        for (auto kv = cnt.rbegin(); kv != cnt.rend(); ++kv) {
            const double c = kv->first;
            const long rounds = kv->second / W, rest = kv->second % W;
            if (rounds)
                for (double &f : fin) f += (double)rounds * c;
            /* fin is kept ascending: the rest go to the earliest-free waves */
            for (long i = 0; i < rest; ++i) fin[(size_t)i] += c;
            std::sort(fin.begin(), fin.end());
        }
        total += fin.back();
    }
    return total;
}

/* the model's best height in [lo, hi] (the taller one on a tie) */
static int model_tile_rows(const TileModel &m, int lo, int hi)
{
    hi = std::min(hi, m.maxz);
    lo = std::min(lo, hi);
    int best = hi;
    double best_t = 1e300;
    for (int tz = hi; tz >= lo; --tz) {
        const double t = predict_step(m, tz);
        if (t < best_t * (1.0 - 1e-9)) { best_t = t; best = tz; }
    }
    return best;
}

/* Tile height before (or without) the autotuner: the tallest of 32, 16,
   ..., 1 element rows that still gives every degree's launch at least four
   tiles per wave of its persistent grid.  Many short tiles keep the grid of
   a small mesh busy and its sweeps short (one 200x100-element instance:
   0.020 ms per step at 1 row, 0.036 at 4); tall tiles read fewer halo rows
   (64 instances on that mesh: 0.32 ms at 8-16 rows, 0.49 at 1).  In the
   serving loop the instance set changes with every admission and
   retirement; this rule moves between few heights, where the step model's
   best height would move at nearly every change and re-tile and re-capture
   the step graphs each time (the C5 window: 8.5e10 point-steps/s against
   7.2e10 with the model's height and 8.1e10 with the model restricted to
   powers of two).  Four tiles per wave rather than two keep the grid's
   tail short when the serving loop caps the active points (the C5 window
   with at most 64M active points: 8.07e10 at 16 rows, 7.71e10 at 32). */
static int default_tile_rows(const semsks_engine *e)
{
    int best = 32;
    for (int n = 3; n <= 6; ++n) {
        const long waves = 4L * persistent_blocks<false>(e, n);
        const int xc = xchunk_of(n);
        bool any = false;
        int tz = 1;
        for (int c = 32; c >= 1; c /= 2) {
            long tiles = 0;
            for (const Slot &sl : e->slots)
                if (sl.used && sl.meta.hpn == n) {
                    any = true;
                    tiles += (long)((sl.meta.nex + xc - 1) / xc) * ((sl.meta.nez + c - 1) / c);
                }
            if (tiles >= 4 * waves) { tz = c; break; }
        }
        if (any) best = std::min(best, tz);
    }
    return best;
}

/* one launch of the element-force kernel for the tiles of polynomial degree n */
template <bool FORCE_ONLY>
static void launch_tiles(const semsks_engine *e, int n, const InstDesc *desc, const int *levels, const int4 *tl,
                         const int *nt)
{
    const int nb = persistent_blocks<FORCE_ONLY>(e, n);
    hipStream_t st = e->stream;
    /* time steps take tiles from the queues (cleared by the post_kernel of
       the launch before, or at engine creation); the force-only launch
       deals them statically */
    int *q = FORCE_ONLY ? nullptr : e->d_qctr + n * SEMSKS_NQ * SEMSKS_QSTRIDE;
    switch (n) {
    case 3: hipLaunchKernelGGL((step_kernel<3, FORCE_ONLY>), dim3(nb), dim3(256), 0, st, desc, levels, tl, nt, q); break;
    case 4: hipLaunchKernelGGL((step_kernel<4, FORCE_ONLY>), dim3(nb), dim3(256), 0, st, desc, levels, tl, nt, q); break;
    case 5: hipLaunchKernelGGL((step_kernel<5, FORCE_ONLY>), dim3(nb), dim3(256), 0, st, desc, levels, tl, nt, q); break;
    case 6: hipLaunchKernelGGL((step_kernel<6, FORCE_ONLY>), dim3(nb), dim3(256), 0, st, desc, levels, tl, nt, q); break;
    }
}

/* one launch: steps every instance by one level */
static semsks_status launch_step(semsks_engine *e, hipEvent_t ta, hipEvent_t tb)
{
    const int nslots = (int)e->slots.size();
    if (ta) HIPCHK(hipEventRecord(ta, e->stream));
    for (int n = 3; n <= 6; ++n)
        if (e->hpn_mask & (1u << n))
            launch_tiles<false>(e, n, e->d_desc, e->d_levels, e->d_tiles + e->tile_off[n], e->d_ntiles + n);
    if (tb) HIPCHK(hipEventRecord(tb, e->stream));
    if (e->any_sponge) hipLaunchKernelGGL(sponge_kernel, dim3(64, nslots), dim3(256), 0, e->stream, e->d_desc, e->d_levels);
    hipLaunchKernelGGL(post_kernel, dim3(e->post_bx, nslots), dim3(256), 0, e->stream, e->d_desc, e->d_levels, 1, e->d_qctr,
                       e->d_pctr);
    HIPCHK(hipGetLastError());
    return SEMSKS_OK;
}

static void drop_graph(semsks_engine *e)
{
    for (int i = 0; i < 5; ++i) {
        if (e->gexec[i]) { hipGraphExecDestroy(e->gexec[i]); e->gexec[i] = nullptr; }
        if (e->graph[i]) { hipGraphDestroy(e->graph[i]); e->graph[i] = nullptr; }
    }
}

/* the captured steps still match the instance set (tiles, degrees, sponge,
   sweep mode)? */
static bool graphs_current(const semsks_engine *e)
{
    return e->graph_mask == e->hpn_mask && memcmp(e->graph_off, e->tile_off, sizeof e->graph_off) == 0 &&
           e->graph_sponge == e->any_sponge;
}

/* captures graph[i], 16 >> i steps (a multiple of the steps per launch) */
static semsks_status build_graph(semsks_engine *e, int i)
{
    bool any = false;
    for (int j = 0; j < 5; ++j) any = any || e->gexec[j];
    if (any && !graphs_current(e)) drop_graph(e);
    HIPCHK(hipStreamBeginCapture(e->stream, hipStreamCaptureModeThreadLocal));
    const int steps = 16 >> i;
    for (int s = 0; s < steps; ++s) {
        semsks_status st = launch_step(e, nullptr, nullptr);
        if (st != SEMSKS_OK) { hipGraph_t g; hipStreamEndCapture(e->stream, &g); if (g) hipGraphDestroy(g); return st; }
    }
    HIPCHK(hipStreamEndCapture(e->stream, &e->graph[i]));
    HIPCHK(hipGraphInstantiate(&e->gexec[i], e->graph[i], nullptr, nullptr, 0));
    e->graph_mask = e->hpn_mask;
    e->graph_sponge = e->any_sponge;
    memcpy(e->graph_off, e->tile_off, sizeof e->graph_off);
    return SEMSKS_OK;
}

extern "C" {

/* device, stream and the fixed tables of a new engine; on failure the
   caller destroys the partly built engine */
static semsks_status engine_init(semsks_engine *e)
{
    HIPCHK(hipGetDevice(&e->device));
    HIPCHK(scratch_ptr(e->device, &e->scratch));
    hipDeviceProp_t prop;
    HIPCHK(hipGetDeviceProperties(&prop, e->device));
    e->n_cu = prop.multiProcessorCount;
    size_t fr = 0, tot = 0;
    HIPCHK(hipMemGetInfo(&fr, &tot));
    e->budget = e->opts.hbm_budget ? e->opts.hbm_budget : (size_t)(0.9 * (double)fr);
    HIPCHK(hipStreamCreateWithFlags(&e->stream, hipStreamNonBlocking));
    HIPCHK(hipStreamCreateWithFlags(&e->setup, hipStreamNonBlocking));
    for (auto &v : e->ev) HIPCHK(hipEventCreate(&v));
    e->slots.resize(e->opts.max_slots);
    e->h_levels.assign(e->opts.max_slots, -1);
    HIPCHK(hipMalloc(&e->d_desc, e->slots.size() * sizeof(InstDesc)));
    HIPCHK(hipMalloc(&e->d_levels, e->slots.size() * sizeof(int)));
    HIPCHK(hipMalloc(&e->d_ntiles, SEMSKS_MAXNP * sizeof(int)));
    HIPCHK(hipMalloc(&e->d_qctr, SEMSKS_QCOUNTERS * SEMSKS_QSTRIDE * sizeof(int)));
    HIPCHK(hipMemset(e->d_qctr, 0, SEMSKS_QCOUNTERS * SEMSKS_QSTRIDE * sizeof(int)));
    HIPCHK(hipMalloc(&e->d_pctr, e->slots.size() * sizeof(int)));
    HIPCHK(hipMemset(e->d_pctr, 0, e->slots.size() * sizeof(int)));
    std::vector<int> z(e->slots.size(), 0);
    HIPCHK(hipMemcpy(e->d_levels, z.data(), z.size() * sizeof(int), hipMemcpyHostToDevice));
    return upload_tables(e);
}

semsks_status semsks_engine_create(const semsks_engine_opts *opts, semsks_engine **out)
{
    if (!out) return semsks_set_error(SEMSKS_EINVAL, "bad arguments");
    *out = nullptr;
    if (opts && (opts->steps_per_sweep < 0 || opts->steps_per_sweep > 1))
        return semsks_set_error(SEMSKS_EINVAL, "steps_per_sweep must be 0 or 1 (one time step per launch)");
    semsks_engine *e = new semsks_engine();
    if (opts) e->opts = *opts;
    if (e->opts.max_slots <= 0) e->opts.max_slots = 1;
    int ndev = 0;
    if (hipGetDeviceCount(&ndev) != hipSuccess || ndev < 1) {
        delete e;
        return semsks_set_error(SEMSKS_EDEVICE, "no HIP device");
    }
    semsks_status st = engine_init(e);
    if (st != SEMSKS_OK) { semsks_engine_destroy(e); return st; }
    *out = e;
    return SEMSKS_OK;
}

static void free_slot(semsks_engine *e, int k)
{
    Slot &s = e->slots[k];
    for (void *p : s.allocs) hipFree(p);
    e->used -= s.bytes;
    if (s.share >= 0) e->slots[s.share].refs--;
    s = Slot();
}

void semsks_engine_destroy(semsks_engine *e)
{
    if (!e) return;
    if (e->stream) hipStreamSynchronize(e->stream);
    drop_graph(e);
    for (size_t k = 0; k < e->slots.size(); ++k) if (e->slots[k].used) free_slot(e, (int)k);
    if (e->setup) hipStreamSynchronize(e->setup);
    if (e->sbuf) hipFree(e->sbuf);
    hipFree(e->d_desc); hipFree(e->d_levels); hipFree(e->d_tiles); hipFree(e->d_ntiles); hipFree(e->d_qctr); hipFree(e->d_pctr);
    for (auto &v : e->ev) if (v) hipEventDestroy(v);
    for (auto &v : e->tev) hipEventDestroy(v);
    if (e->stream) hipStreamDestroy(e->stream);
    if (e->setup) hipStreamDestroy(e->setup);
    delete e;
}

static int nx_of(const semsks_problem *p) { return p->nex * p->hpn + 1; }
static int nz_of(const semsks_problem *p) { return p->nez * p->hpn + 1; }
/* storage block width: the owned node columns of one step-kernel tile */
static int bw_of(const semsks_problem *p) { return xchunk_of(p->hpn) * p->hpn; }
/* nodes per level plane: whole blocks, rounded up to 64 so that every plane
   of a slot starts on a 256-byte boundary */
static size_t plane_of(const semsks_problem *p)
{
    const size_t bw = bw_of(p), nblk = (nx_of(p) + bw - 1) / bw;
    return (nblk * bw * nz_of(p) + 63) / 64 * 64;
}
static int nt_sem_of(const semsks_problem *p) { return (int)((float)((p->nt - 3) * p->dp / p->sdt)); }

size_t semsks_footprint(const semsks_problem *p, int shared)
{
    const size_t nodes = plane_of(p);
    /* three levels: xyz plane + compensation words (none for the plain update) */
    size_t b = nodes * 4 * 3 * ((p->flags & SEMSKS_F_NO_KAHAN) ? 3 : 4);
    if (!shared) b += nodes * 2 + (size_t)p->nex * p->nez * 2 +
                      (size_t)p->n_mat * (64 + 4 * (p->hpn + 1) * (p->hpn + 1)) +
                      (size_t)semsks_n_edge(nx_of(p), nz_of(p), p->hpn) * 24 +
                      2048 + (size_t)(p->nex / 2 + 1) * (p->nez + 1) * 16;   /* semsks_stiff_dev tiles (bound) */
    const int mt = 2 * (nx_of(p) + nz_of(p)), sht = nx_of(p) + nz_of(p);
    int nlev = std::min(std::max(nt_sem_of(p), 1), std::max(p->nst, 1));
    b += (size_t)nlev * (mt + sht) * 4;
    b += (size_t)p->nst * 3 * std::max(p->ntrace, 1) * 4 + (size_t)p->ntrace * 8;
    b += (size_t)(nx_of(p) + nz_of(p)) * 12;        /* pre_inj (bound) */
    return b;
}

/* host mirror of interp_const's slot ring (interpl.c:59-91) for one or two
   streams sharing it: which record sits in each tap and its weight */
static void build_schedule(int nlev, float rdt, int nstream, const int nrec[2], std::vector<SchedEntry> out[2])
{
    int ns[4] = {0, 1, 2, 3};
    int slot_rec[2][4] = {{-1, -1, -1, -1}, {-1, -1, -1, -1}};
    int nread[2] = {0, 0};
    out[0].resize(nlev);
    out[1].resize(nlev);
    for (int it = 0; it < nlev; ++it) {
        for (int s = 0; s < nstream; ++s) {
            const float nst0 = (float)it, t = nst0 * rdt;
            const int ns0 = (int)(t + 0.00001);
            if (ns0 > (int)((nst0 - 1) * rdt + 0.00001) || nst0 == 0.) {
                int r = ns[0];
                for (int j = 0; j < 3; ++j) ns[j] = ns[j + 1];
                ns[3] = r;
                slot_rec[s][ns[3]] = nread[s] < nrec[s] ? nread[s] : -1;
                nread[s]++;
            }
            const float d1 = t - (int)(t + 0.00001);
            const float dty1 = d1, dty2 = dty1 * dty1, dty3 = dty1 * dty2;
            SchedEntry &E = out[s][it];
            E.w[0] = (float)(-0.5 * dty3 + dty2 - 0.5 * dty1);
            E.w[1] = (float)(3.0 * dty3 - 2.5 * dty2 + 1.0);
            E.w[2] = (float)(-3.0 * dty3 + 2.0 * dty2 + 0.5 * dty1);
            E.w[3] = (float)(0.5 * (dty3 - dty2));
            for (int j = 0; j < 4; ++j) E.rec[j] = slot_rec[s][ns[j]];
        }
    }
}

/* device bytes a slot for p keeps */
static size_t admit_resident(const semsks_engine *, const semsks_problem *p, bool shared)
{
    return semsks_footprint(p, shared);
}

/* bytes of admission temporaries after the pending records in the
   scratch: descriptor, GLL weights and the Jacobian check (one 256-byte
   line each), the FP64 lumped mass and the small tables of the one-off
   setup (none for a slot that shares its material), the interpolation
   schedule, and the upload of one stream's host boundary records (none
   when they were generated on the device, pending) */
static size_t al256(size_t b) { return (b + 255) / 256 * 256; }
static size_t admit_temp(const semsks_problem *p, bool shared, bool pending)
{
    const size_t mt = 2 * (size_t)(nx_of(p) + nz_of(p));
    const int nlev = std::min(std::max(nt_sem_of(p), 1), std::max(p->nst, 1));
    size_t b = 3 * 256 + al256(sizeof(InstDesc));
    if (!shared) b += al256((size_t)nx_of(p) * nz_of(p) * 8);
    b += al256((size_t)nlev * sizeof(SchedEntry));
    if (!pending && p->nt > 0 && (p->psv || p->sh)) b += al256((size_t)p->nt * mt * 4);
    return b;
}

/* device bytes of the plane-wave records semsks_planewave(..., NULL, NULL)
   keeps for the next admission */
static size_t pending_bytes(int nt, int nx, int nz) { return al256((size_t)nt * 2 * (nx + nz) * 4) + al256((size_t)nt * (nx + nz) * 4); }

/* an allocation failure maps to ENOHBM (the serving loop then waits for
   memory) when the device is out of memory, to EDEVICE otherwise; the
   runtime's last-error state is cleared so the next check does not see it */
static semsks_status alloc_failed(hipError_t err, const char *what)
{
    (void)hipGetLastError();
    if (err == hipErrorOutOfMemory) return semsks_set_error(SEMSKS_ENOHBM, "device out of memory (%s)", what);
    return semsks_set_error(SEMSKS_EDEVICE, "%s: %s", what, hipGetErrorString(err));
}

size_t semsks_admission_bytes(const semsks_problem *p, int device_records)
{
    if (!p) return 0;
    const size_t pend = device_records ? pending_bytes(p->nt, nx_of(p), nz_of(p)) : 0;
    return pend + admit_temp(p, false, device_records != 0);
}

semsks_status semsks_room(semsks_engine *e, const semsks_problem *p, size_t pending)
{
    if (!e || !p) return semsks_set_error(SEMSKS_EINVAL, "null argument");
    bool free_slot_left = false;
    for (auto &s : e->slots) free_slot_left = free_slot_left || !s.used;
    if (!free_slot_left) return semsks_set_error(SEMSKS_ENOSLOT, "all %zu slots busy", e->slots.size());
    const size_t need = admit_resident(e, p, false) + scratch_growth(e, pending + admit_temp(p, false, pending > 0));
    if (e->used + need > e->budget)
        return semsks_set_error(SEMSKS_ENOHBM, "slot needs %zu bytes, %zu of %zu in use", need, e->used, e->budget);
    return SEMSKS_OK;
}

/* 64-bit hash of a problem's material (records and per-node ids): a slot
   may share another's material only when the two hash alike */
static uint64_t material_hash(const semsks_problem *p)
{
    uint64_t h = 0x9e3779b97f4a7c15ull;
    auto mix = [&h](const void *data, size_t n) {
        const unsigned char *c = (const unsigned char *)data;
        size_t i = 0;
        for (; i + 8 <= n; i += 8) {
            uint64_t v;
            memcpy(&v, c + i, 8);
            h = (h ^ v) * 0xff51afd7ed558ccdull;
            h ^= h >> 33;
        }
        for (; i < n; ++i) h = (h ^ c[i]) * 0x100000001b3ull;
    };
    mix(p->mat_table, (size_t)p->n_mat * SEMSKS_MAT_STRIDE * sizeof(float));
    mix(p->mat_id, (size_t)nx_of(p) * nz_of(p) * sizeof(uint16_t));
    return h;
}

semsks_status semsks_admit(semsks_engine *e, const semsks_problem *p, int share, int *slot_out)
{
    if (!e || !p || !slot_out) return semsks_set_error(SEMSKS_EINVAL, "null argument");
    if (p->hpn < 3 || p->hpn > 6 || p->nex <= p->neh || p->nez <= p->neh || p->nst < 1 || !(p->sdt > 0) ||
        !(p->eh > 0) || p->n_mat < 1 || !p->mat_table || !p->mat_id || p->ntrace < 0 || p->nt < 0 ||
        (p->solver < 0 || p->solver > 2))
        return semsks_set_error(SEMSKS_EINVAL, "invalid problem description");
    if (plane_of(p) >= ((size_t)1 << 31))   /* node offsets are 32-bit in the kernels */
        return semsks_set_error(SEMSKS_EINVAL, "mesh of %zu nodes per plane is too large", plane_of(p));
    int k = -1;
    for (size_t i = 0; i < e->slots.size(); ++i) if (!e->slots[i].used) { k = (int)i; break; }
    if (k < 0) return semsks_set_error(SEMSKS_ENOSLOT, "all %zu slots busy", e->slots.size());
    if (share >= 0) {
        if (share >= (int)e->slots.size() || !e->slots[share].used || e->slots[share].share >= 0)
            return semsks_set_error(SEMSKS_EINVAL, "share_slot %d is not a material owner", share);
        const semsks_problem &q = e->slots[share].meta;
        if (q.nex != p->nex || q.nez != p->nez || q.hpn != p->hpn || q.eh != p->eh || q.sdt != p->sdt ||
            q.absorb != p->absorb)
            return semsks_set_error(SEMSKS_EINVAL, "share_slot %d has a different mesh", share);
        if (e->slots[share].mat_hash != material_hash(p))
            return semsks_set_error(SEMSKS_EINVAL, "share_slot %d has a different material", share);
    }
    const bool pending = !p->psv && !p->sh && e->pend_psv && e->pend_nt == p->nt && e->pend_nx == nx_of(p) &&
                         e->pend_nz == nz_of(p);
    const size_t pendb = pending ? e->pend_bytes : 0;
    const size_t need = admit_resident(e, p, share >= 0) + scratch_growth(e, pendb + admit_temp(p, share >= 0, pending));
    if (e->used + need > e->budget)
        return semsks_set_error(SEMSKS_ENOHBM, "slot needs %zu bytes, %zu of %zu in use", need, e->used, e->budget);
    for (int i = 0; i < p->ntrace; ++i)
        if (p->rec_ix[i] < 0 || p->rec_ix[i] >= nx_of(p) || p->rec_iz[i] < 0 || p->rec_iz[i] >= nz_of(p))
            return semsks_set_error(SEMSKS_EINVAL, "receiver %d outside the mesh", i);
    for (long q = 0; q < (long)nx_of(p) * nz_of(p); ++q)
        if (p->mat_id[q] >= p->n_mat) return semsks_set_error(SEMSKS_EINVAL, "material id out of range at node %ld", q);

    Slot &s = e->slots[k];
    s = Slot();
    s.meta = *p;
    s.nx = nx_of(p); s.nz = nz_of(p); s.bw = bw_of(p); s.plane = plane_of(p);
    s.nt_sem = nt_sem_of(p);
    s.nlev = std::max(1, std::min(s.nt_sem, p->nst));
    s.share = share;
    s.mat_hash = share >= 0 ? e->slots[share].mat_hash : material_hash(p);
    InstDesc &d = s.d;
    memset(&d, 0, sizeof d);
    d.nex = p->nex; d.nez = p->nez; d.nx = s.nx; d.nz = s.nz; d.bw = s.bw;
    d.hpn = p->hpn; d.neh = p->neh; d.nst = p->nst; d.nt_sem = s.nt_sem; d.solver = p->solver;
    d.kahan = (p->flags & SEMSKS_F_NO_KAHAN) ? 0 : 1;
    d.ortho = 1;     /* every record orthotropic in the mesh axes: the short stress form */
    for (int i = 0; i < p->n_mat; ++i)
        if (!semsks_ortho_record(p->mat_table + (size_t)i * SEMSKS_MAT_STRIDE + 1)) d.ortho = 0;
    d.ntrace = p->ntrace;
    const size_t nodes = s.plane;
    const int mt = 2 * (s.nx + s.nz), sht = s.nx + s.nz;
    semsks_status st = SEMSKS_OK;
#define ACHK(x) do { const hipError_t ae_ = (x); if (ae_ != hipSuccess) { st = alloc_failed(ae_, #x); goto fail; } } while (0)
    /* temporaries: the scratch after the pending records */
    st = setup_scratch(e, pendb + admit_temp(p, share >= 0, pending), pendb);
    if (st != SEMSKS_OK) { s = Slot(); return st; }
    {
        float *st_all = nullptr;
        /* three levels, each an xyz plane and, with compensation, a
           compensation-word plane (the plain update never touches the
           words: no planes, comp[l] null); every plane starts on a 256-byte
           boundary (nodes is a multiple of 64) */
        d.nlev = 3;
        const int nplanes = d.nlev * (d.kahan ? 4 : 3);
        ACHK(dalloc(s, &st_all, nodes * nplanes));
        ACHK(hipMemsetAsync(st_all, 0, nodes * nplanes * sizeof(float), e->setup));
        for (int l = 0; l < d.nlev; ++l) {
            d.u[l] = st_all + 3 * l * nodes;
            d.comp[l] = d.kahan ? reinterpret_cast<uint32_t *>(st_all + 3 * d.nlev * nodes + l * nodes) : nullptr;
        }
    }
    if (p->atten > 0) {   /* atten_taper (sem2d_psvsh.c:1725-1738), double then float */
        const int nh = p->neh * p->hpn;
        if (nh < 1 || s.nx < 2 * nh || s.nz < 2 * nh) {
            st = semsks_set_error(SEMSKS_EINVAL, "atten: strips of %d nodes do not fit the mesh", nh);
            goto fail;
        }
        std::vector<float> tp(nh);
        const double alpha = 1.0 / nh * log10((double)p->ass) / log10(2.73);
        for (int i = 0; i < nh; ++i) tp[i] = (float)exp(alpha * i);
        float *dt = nullptr;
        ACHK(dalloc(s, &dt, (size_t)nh));
        ACHK(hipMemcpyAsync(dt, tp.data(), nh * sizeof(float), hipMemcpyHostToDevice, e->setup));
        ACHK(hipStreamSynchronize(e->setup));
        d.taper = dt;
        d.atten = p->atten;
        d.nh = nh;
    }
    if (share >= 0) {
        const InstDesc &o = e->slots[share].d;
        d.mat_id = o.mat_id; d.emat = o.emat; d.mat = o.mat; d.mclass = o.mclass; d.edge_a = o.edge_a; d.edge_b = o.edge_b;
        d.ortho = o.ortho;    /* the stress form follows the records the slot reads: the owner's */
        memcpy(d.Dg, o.Dg, sizeof d.Dg); memcpy(d.Dw, o.Dw, sizeof d.Dw); memcpy(d.cw, o.cw, sizeof d.cw);
        e->slots[share].refs++;
    } else {
        uint16_t *mid = nullptr, *emat = nullptr;
        float *mat = nullptr, *mcl = nullptr, *ea = nullptr, *eb = nullptr;
        const int ncls = (p->hpn + 1) * (p->hpn + 1);
        ACHK(dalloc(s, &mid, nodes));
        ACHK(dalloc(s, &emat, (size_t)p->nex * p->nez));
        ACHK(dalloc(s, &mat, (size_t)p->n_mat * SEMSKS_MAT_STRIDE));
        ACHK(dalloc(s, &mcl, (size_t)p->n_mat * ncls));
        const int ne = semsks_n_edge(s.nx, s.nz, p->hpn);
        ACHK(dalloc(s, &ea, (size_t)ne * 3));
        ACHK(dalloc(s, &eb, (size_t)ne * 3));
        {   /* record ids in the block order of the level planes */
            std::vector<uint16_t> ids(nodes, 0);
            for (int iz = 0; iz < s.nz; ++iz)
                for (int ix = 0; ix < s.nx; ++ix)
                    ids[node_at(s.bw, s.nz, ix, iz)] = p->mat_id[(size_t)iz * s.nx + ix];
            ACHK(hipMemcpyAsync(mid, ids.data(), nodes * 2, hipMemcpyHostToDevice, e->setup));
            ACHK(hipStreamSynchronize(e->setup));
        }
        ACHK(hipMemcpyAsync(mat, p->mat_table, (size_t)p->n_mat * SEMSKS_MAT_STRIDE * 4, hipMemcpyHostToDevice, e->setup));
        {   /* element-uniform record ids */
            std::vector<uint16_t> em((size_t)p->nex * p->nez);
            const int n = p->hpn;
            for (int ez = 0; ez < p->nez; ++ez)
                for (int ex = 0; ex < p->nex; ++ex) {
                    uint16_t r = p->mat_id[(size_t)(ez * n) * s.nx + ex * n];
                    for (int b = 0; b <= n && r != 0xFFFF; ++b)
                        for (int a = 0; a <= n; ++a)
                            if (p->mat_id[(size_t)(ez * n + b) * s.nx + ex * n + a] != r) { r = 0xFFFF; break; }
                    em[(size_t)ez * p->nex + ex] = r;
                }
            ACHK(hipMemcpyAsync(emat, em.data(), em.size() * 2, hipMemcpyHostToDevice, e->setup));
            ACHK(hipStreamSynchronize(e->setup));
            s.emat = std::move(em);
        }
        d.mat_id = mid; d.emat = emat; d.mat = mat; d.mclass = mcl; d.edge_a = ea; d.edge_b = eb;
        /* FP64 setup on the device */
        {
            unsigned char *tb = e->sbuf + pendb;
            InstDesc *dd = (InstDesc *)tb;
            double *w64 = (double *)(tb + al256(sizeof(InstDesc)));
            unsigned long long *dev = (unsigned long long *)(tb + al256(sizeof(InstDesc)) + 256);
            double *mass = (double *)(tb + al256(sizeof(InstDesc)) + 512);
            ACHK(hipMemsetAsync(dev, 0, sizeof(unsigned long long), e->setup));
            ACHK(hipMemcpyAsync(dd, &d, sizeof d, hipMemcpyHostToDevice, e->setup));
            hipLaunchKernelGGL(setup_tables_kernel, dim3(1), dim3(64), 0, e->setup, dd, w64, (double)p->eh);
            hipLaunchKernelGGL(jacobian_kernel, dim3(256), dim3(256), 0, e->setup, p->nex, p->nez, p->hpn, (double)p->eh, dev);
            const double eh = p->eh, je = (double)((double)eh * eh / 4.0), jb = eh / 2.0;
            hipLaunchKernelGGL(mass_kernel, dim3(1024), dim3(256), 0, e->setup, d, w64, je, p->sdt, mass);
            hipLaunchKernelGGL(mclass_kernel, dim3(64), dim3(256), 0, e->setup, mat, p->n_mat, p->hpn, w64, je, p->sdt, mcl);
            hipLaunchKernelGGL(absorb_kernel, dim3(64), dim3(256), 0, e->setup, d, w64, jb, p->sdt, p->absorb, mass,
                               (double *)nullptr, ea, eb);
            InstDesc back;
            ACHK(hipMemcpyAsync(&back, dd, sizeof back, hipMemcpyDeviceToHost, e->setup));
            unsigned long long devbits = 0;
            ACHK(hipMemcpyAsync(&devbits, dev, sizeof devbits, hipMemcpyDeviceToHost, e->setup));
            ACHK(hipStreamSynchronize(e->setup));
            memcpy(d.Dg, back.Dg, sizeof d.Dg); memcpy(d.Dw, back.Dw, sizeof d.Dw); memcpy(d.cw, back.cw, sizeof d.cw);
            double maxdev;
            memcpy(&maxdev, &devbits, sizeof maxdev);
            if (!(maxdev <= 1.e-2)) { st = semsks_set_error(SEMSKS_EINVAL, "non-uniform Jacobian (max dev %g)", maxdev); goto fail; }
        }
    }
    /* the force-only descriptor and tiles of semsks_stiff_dev (owners whose
       plane fits the scratch array) */
    if (share < 0 && scratch_bytes(nodes) <= SEMSKS_SCRATCH_BYTES) {
        std::vector<int4> ft;
        build_tiles(s, 0, std::min(FORCE_TZ, std::max(1, p->nez)), ft);
        InstDesc fd = d;
        fd.active = 1;
        unsigned char *base = (unsigned char *)e->scratch;
        fd.u[0] = (float *)base;
        for (int c = 0; c < 3; ++c) fd.fout[c] = (float *)(base + scratch_f_offset(nodes)) + (size_t)c * nodes;
        unsigned char *blk = nullptr;
        const int nft = (int)ft.size();
        const size_t dsz = (sizeof(InstDesc) + 255) / 256 * 256;     /* [desc | count | tiles] */
        ACHK(dalloc(s, &blk, dsz + 256 + ft.size() * sizeof(int4)));
        s.d_fdesc = (InstDesc *)blk;
        s.d_fcount = (int *)(blk + dsz);
        s.d_ftiles = (int4 *)(blk + dsz + 256);
        ACHK(hipMemcpyAsync(s.d_fdesc, &fd, sizeof fd, hipMemcpyHostToDevice, e->setup));
        ACHK(hipMemcpyAsync(s.d_fcount, &nft, sizeof nft, hipMemcpyHostToDevice, e->setup));
        if (nft) ACHK(hipMemcpyAsync(s.d_ftiles, ft.data(), ft.size() * sizeof(int4), hipMemcpyHostToDevice, e->setup));
        ACHK(hipStreamSynchronize(e->setup));
    }
    /* boundary series */
    {
        float *G = nullptr, *Gv = nullptr;
        ACHK(dalloc(s, &G, (size_t)s.nlev * mt));
        ACHK(dalloc(s, &Gv, (size_t)s.nlev * sht));
        ACHK(hipMemsetAsync(G, 0, (size_t)s.nlev * mt * 4, e->setup));
        ACHK(hipMemsetAsync(Gv, 0, (size_t)s.nlev * sht * 4, e->setup));
        const float rdt = p->sdt / p->dp;
        const int nstream = p->solver == SEMSKS_SOLVER_PSVSH ? 2 : 1;
        int nrec[2] = {(p->psv || pending) ? p->nt : 0, (p->sh || pending) ? p->nt : 0};
        std::vector<SchedEntry> sch[2];
        const int nlev_series = std::min(s.nt_sem, s.nlev);
        if (nlev_series > 0) {
            build_schedule(nlev_series, rdt, nstream, nrec, sch);
            /* after the setup temporaries (the mass is no longer needed) */
            unsigned char *tb = e->sbuf + pendb + 3 * 256 + al256(sizeof(InstDesc));
            SchedEntry *dsch = (SchedEntry *)tb;
            float *dup = (float *)(tb + al256((size_t)nlev_series * sizeof(SchedEntry)));
            for (int str = 0; str < nstream; ++str) {
                const int width = str ? sht : mt;
                const float *src = str ? p->sh : p->psv;
                float *drec = nullptr;
                bool own = false;
                if (pending) drec = str ? e->pend_sh : e->pend_psv;
                else if (src && p->nt > 0) {
                    drec = dup;
                    ACHK(hipMemcpyAsync(drec, src, (size_t)p->nt * width * 4, hipMemcpyHostToDevice, e->setup));
                    own = true;
                }
                if (!drec) continue;
                ACHK(hipMemcpyAsync(dsch, sch[str].data(), nlev_series * sizeof(SchedEntry), hipMemcpyHostToDevice, e->setup));
                hipLaunchKernelGGL(series_kernel, dim3(std::min(nlev_series, 65535)), dim3(256), 0, e->setup, drec,
                                   width, nrec[str], dsch, nlev_series, str ? Gv : G);
                ACHK(hipStreamSynchronize(e->setup));
                (void)own;
            }
        }
        if (pending) {              /* consumed: the scratch keeps its size */
            e->pend_psv = e->pend_sh = nullptr;
            e->pend_nt = 0;
            e->pend_bytes = 0;
        }
        d.G = G; d.Gv = Gv;
    }
    /* receivers and traces */
    {
        int *rec = nullptr;
        float *tr = nullptr;
        std::vector<int> off(std::max(p->ntrace, 1), 0);
        for (int i = 0; i < p->ntrace; ++i) off[i] = (int)node_at(s.bw, s.nz, p->rec_ix[i], p->rec_iz[i]);
        ACHK(dalloc(s, &rec, off.size()));
        ACHK(dalloc(s, &tr, (size_t)p->nst * 3 * std::max(p->ntrace, 1)));
        ACHK(hipMemcpyAsync(rec, off.data(), off.size() * 4, hipMemcpyHostToDevice, e->setup));
        ACHK(hipMemsetAsync(tr, 0, (size_t)p->nst * 3 * std::max(p->ntrace, 1) * 4, e->setup));
        d.rec = rec; d.traces = tr;
        /* receivers on interface nodes, recorded by post_kernel's thread for
           that node before it injects: per interface node the first such
           receiver, per receiver the next one at its node (-1 ends a list,
           -2 marks a receiver off the interface) */
        const int nbl = p->neh * p->hpn, nb = std::max(s.nx - nbl, 0), nl = std::max(s.nz - nbl - 1, 0);
        std::vector<int> head(std::max(nb + nl, 1), -1), next(std::max(p->ntrace, 1), -2);
        for (int r = p->ntrace - 1; r >= 0; --r) {
            const int ix = p->rec_ix[r], iz = p->rec_iz[r];
            int i = -1;
            if (iz == nbl && ix >= nbl && ix < s.nx) i = ix - nbl;
            else if (ix == nbl && iz > nbl && iz < s.nz) i = nb + (iz - nbl - 1);
            if (i < 0) continue;
            next[r] = head[i];
            head[i] = r;
        }
        int *dhead = nullptr, *dnext = nullptr;
        ACHK(dalloc(s, &dhead, head.size()));
        ACHK(dalloc(s, &dnext, next.size()));
        /* the interface nodes of the newest level before its injection
           (semsks_get_snapshot) */
        ACHK(dalloc(s, &d.pre_inj, (size_t)std::max(nb + nl, 1) * 3));
        ACHK(hipMemcpyAsync(dhead, head.data(), head.size() * 4, hipMemcpyHostToDevice, e->setup));
        ACHK(hipMemcpyAsync(dnext, next.data(), next.size() * 4, hipMemcpyHostToDevice, e->setup));
        d.inj_rec = dhead; d.rec_next = dnext;
    }
    /* level 0: its state, its record, then the injection of the first step */
    hipLaunchKernelGGL(init_state_kernel, dim3(64), dim3(256), 0, e->setup, d);
    if (p->ntrace > 0) hipLaunchKernelGGL(record0_kernel, dim3(1), dim3(256), 0, e->setup, d);
    hipLaunchKernelGGL(inject0_kernel, dim3(16), dim3(256), 0, e->setup, d);
    {
        int zero = 0;
        ACHK(hipMemcpyAsync(e->d_levels + k, &zero, sizeof zero, hipMemcpyHostToDevice, e->setup));
        ACHK(hipStreamSynchronize(e->setup));
    }
    if (hipGetLastError() != hipSuccess) { st = semsks_set_error(SEMSKS_EDEVICE, "setup kernel launch failed"); goto fail; }
#undef ACHK
    s.used = true;
    e->used += s.bytes;
    e->h_levels[k] = 0;
    e->dirty = true;
    *slot_out = k;
    return SEMSKS_OK;
fail:
    hipStreamSynchronize(e->setup);
    for (void *q : s.allocs) hipFree(q);
    if (share >= 0) e->slots[share].refs--;
    s = Slot();
    return st;
}

semsks_status semsks_retire(semsks_engine *e, int k)
{
    if (!e || k < 0 || k >= (int)e->slots.size() || !e->slots[k].used) return semsks_set_error(SEMSKS_EINVAL, "bad slot %d", k);
    if (e->slots[k].refs > 0) return semsks_set_error(SEMSKS_EINVAL, "slot %d still shares its material with %d slots", k, e->slots[k].refs);
    HIPCHK(hipStreamSynchronize(e->stream));
    free_slot(e, k);
    e->h_levels[k] = -1;
    e->dirty = true;
    return SEMSKS_OK;
}

semsks_status semsks_advance(semsks_engine *e, int nsteps, int sync)
{
    if (!e || nsteps < 0) return semsks_set_error(SEMSKS_EINVAL, "bad arguments");
    if (e->dirty) { semsks_status st = upload_tables(e); if (st) return st; }
    bool any = false;
    for (size_t k = 0; k < e->slots.size(); ++k)
        if (e->slots[k].used && e->h_levels[k] < e->slots[k].meta.nst - 1) any = true;
    if (!any || nsteps == 0) { e->last_ms[0] = e->last_ms[1] = 0; return SEMSKS_OK; }
    HIPCHK(hipEventRecord(e->ev[0], e->stream));
    double kern_ms = 0.0;
    e->timed_steps = 0;
    int done = 0;
    /* autotuner: with tile_z unset, the first call that has the steps to
       spare runs four real launches at each of six candidate tile heights
       and keeps the height whose fastest of the last three was fastest
       (the tiling does not change the arithmetic, so these steps are
       ordinary steps of the run).  The candidates are the step model's
       best height (predict_step) in each of the bands 1, 2-3, 4-7, 8-15,
       16-31 and 32-64 rows.  Until then, and in calls too short for it
       (the serving loop's rounds), default_tile_rows picks the height.
       Measured: one instance on the 200x100 mesh is fastest at 1 row; a
       64-instance batch on the 2000x500 mesh is nearly insensitive to the
       height */
    static const int band[][2] = {{1, 1}, {2, 3}, {4, 7}, {8, 15}, {16, 31}, {32, 64}};
    const int ncand = 6;
    if (e->opts.tile_z <= 0 && !e->tuned && nsteps >= 4 * ncand && sync != 2) {
        struct TuneFlag {
            semsks_engine *e;
            ~TuneFlag() { e->tuning = false; }
        } tune_flag{e};
        e->tuning = true;
        int cand[ncand];
        {
            const TileModel m = tile_model(e);
            for (int c = 0; c < ncand; ++c) cand[c] = model_tile_rows(m, band[c][0], band[c][1]);
        }
        while ((int)e->tev.size() < 8) {
            hipEvent_t v;
            HIPCHK(hipEventCreate(&v));
            e->tev.push_back(v);
        }
        int best = e->tz_cur;
        float best_ms = 1e30f;
        for (int c = 0; c < ncand; ++c) {
            if (std::find(cand, cand + c, cand[c]) != cand + c) continue;   /* bands that collapse on a low mesh */
            e->tz_cur = cand[c];
            semsks_status st = upload_tables(e);
            if (st) return st;
            /* four launches; the fastest of the last three counts */
            for (int r = 0; r < 4; ++r) {
                st = launch_step(e, e->tev[2 * r], e->tev[2 * r + 1]);
                if (st) return st;
            }
            HIPCHK(hipEventSynchronize(e->tev[7]));
            float ms = 1e30f;
            for (int r = 1; r < 4; ++r) {
                float m = 0.f;
                HIPCHK(hipEventElapsedTime(&m, e->tev[2 * r], e->tev[2 * r + 1]));
                ms = std::min(ms, m);
            }
            if (ms < best_ms) { best_ms = ms; best = cand[c]; }
            done += 4;
        }
        e->tz_cur = best;
        semsks_status st = upload_tables(e);
        if (st) return st;
        e->tz_cache[tune_signature(e)] = best;
        e->tuned = true;
        drop_graph(e);
    }
    if (e->opts.use_graph && sync != 2) {
        /* runs of 16 captured steps, the rest in captured runs of 8, 4, 2
           and 1 */
        if (!graphs_current(e)) drop_graph(e);
        while (done < nsteps) {
            const int rem = nsteps - done;
            int i = 0;
            while ((16 >> i) > rem) ++i;
            if (!e->gexec[i]) {
                semsks_status st = build_graph(e, i);
                if (st) return st;
            }
            HIPCHK(hipGraphLaunch(e->gexec[i], e->stream));
            done += 16 >> i;
        }
    } else {
        /* eager launches with an event pair around every step-kernel launch
           (the launches of the first 1024 steps of the call are timed) */
        const int nl = nsteps - done;
        const int ntimed = sync ? std::min(nl, 1024) : 0;
        while ((int)e->tev.size() < 2 * ntimed) {
            hipEvent_t v;
            HIPCHK(hipEventCreate(&v));
            e->tev.push_back(v);
        }
        int tsteps = 0;
        for (int j = 0; done < nsteps; ++j) {
            semsks_status st = j < ntimed ? launch_step(e, e->tev[2 * j], e->tev[2 * j + 1]) : launch_step(e, nullptr, nullptr);
            if (st) return st;
            if (j < ntimed) ++tsteps;
            ++done;
        }
        if (ntimed) {
            HIPCHK(hipEventSynchronize(e->tev[2 * ntimed - 1]));
            for (int i = 0; i < ntimed; ++i) {
                float ms = 0.f;
                HIPCHK(hipEventElapsedTime(&ms, e->tev[2 * i], e->tev[2 * i + 1]));
                kern_ms += ms;
            }
            e->timed_steps = tsteps;
        }
    }
    HIPCHK(hipEventRecord(e->ev[1], e->stream));
    for (size_t k = 0; k < e->slots.size(); ++k)
        if (e->slots[k].used) e->h_levels[k] = std::min(e->h_levels[k] + nsteps, e->slots[k].meta.nst - 1);
    if (sync) {
        HIPCHK(hipEventSynchronize(e->ev[1]));
        float ms = 0.f;
        HIPCHK(hipEventElapsedTime(&ms, e->ev[0], e->ev[1]));
        e->last_ms[0] = ms;
        e->last_ms[1] = kern_ms;
    }
    return SEMSKS_OK;
}

semsks_status semsks_prepare(semsks_engine *e, int nsteps)
{
    if (!e || nsteps < 0) return semsks_set_error(SEMSKS_EINVAL, "bad arguments");
    if (e->dirty) { semsks_status st = upload_tables(e); if (st) return st; }
    if (!e->opts.use_graph) return SEMSKS_OK;
    if (!graphs_current(e)) drop_graph(e);
    /* the runs of 16, 8, 4, 2 and 1 steps semsks_advance would replay */
    for (int done = 0; done < nsteps;) {
        const int rem = nsteps - done;
        int i = 0;
        while ((16 >> i) > rem) ++i;
        if (!e->gexec[i]) {
            semsks_status st = build_graph(e, i);
            if (st) return st;
        }
        done += 16 >> i;
    }
    return SEMSKS_OK;
}

semsks_status semsks_last_timing(semsks_engine *e, double out[2])
{
    if (!e || !out) return semsks_set_error(SEMSKS_EINVAL, "bad arguments");
    out[0] = e->last_ms[0];
    out[1] = e->last_ms[1];
    return SEMSKS_OK;
}

int semsks_last_timed_steps(semsks_engine *e) { return e ? e->timed_steps : 0; }

semsks_status semsks_sync(semsks_engine *e)
{
    if (!e) return semsks_set_error(SEMSKS_EINVAL, "bad arguments");
    HIPCHK(hipStreamSynchronize(e->stream));
    return SEMSKS_OK;
}

int semsks_sweep_steps(semsks_engine *e)
{
    return e ? 1 : 0;
}

int semsks_tile_rows(semsks_engine *e)
{
    if (!e) return 0;
    return e->opts.tile_z > 0 ? e->opts.tile_z : e->tz_cur;
}

semsks_status semsks_slot_status(semsks_engine *e, int k, int *level, int *done)
{
    if (!e || k < 0 || k >= (int)e->slots.size()) return semsks_set_error(SEMSKS_EINVAL, "bad slot %d", k);
    if (!e->slots[k].used) { *level = -1; *done = 0; return SEMSKS_OK; }
    *level = e->h_levels[k];
    *done = e->h_levels[k] >= e->slots[k].meta.nst - 1;
    return SEMSKS_OK;
}

semsks_status semsks_get_traces(semsks_engine *e, int k, int it0, int n, float *out)
{
    if (!e || k < 0 || k >= (int)e->slots.size() || !e->slots[k].used) return semsks_set_error(SEMSKS_EINVAL, "bad slot %d", k);
    const Slot &s = e->slots[k];
    if (it0 < 0 || n < 0 || it0 + n > s.meta.nst) return semsks_set_error(SEMSKS_EINVAL, "trace range");
    const size_t w = (size_t)3 * s.meta.ntrace;
    if (w && n) HIPCHK(hipMemcpyAsync(out, s.d.traces + (size_t)it0 * w, (size_t)n * w * 4, hipMemcpyDeviceToHost, e->stream));
    HIPCHK(hipStreamSynchronize(e->stream));
    /* a run that became unstable (the reference lets NaN/Inf propagate
       into seis.*) is reported; the traces are still copied */
    for (size_t i = 0; i < (size_t)n * w; ++i)
        if (!isfinite(out[i]))
            return semsks_set_error(SEMSKS_ENONFINITE, "slot %d: non-finite trace value at level %zu", k,
                                    (size_t)it0 + i / w);
    return SEMSKS_OK;
}

semsks_status semsks_get_field(semsks_engine *e, int k, float *out)
{
    if (!e || k < 0 || k >= (int)e->slots.size() || !e->slots[k].used) return semsks_set_error(SEMSKS_EINVAL, "bad slot %d", k);
    const Slot &s = e->slots[k];
    const int p = e->h_levels[k] % s.d.nlev;
    std::vector<float> aos(s.plane * 3);
    HIPCHK(hipMemcpyAsync(aos.data(), s.d.u[p], aos.size() * 4, hipMemcpyDeviceToHost, e->stream));
    HIPCHK(hipStreamSynchronize(e->stream));
    for (int c = 0; c < 3; ++c)
        for (int iz = 0; iz < s.nz; ++iz)
            for (int ix = 0; ix < s.nx; ++ix)
                out[((size_t)c * s.nz + iz) * s.nx + ix] = aos[node_at(s.bw, s.nz, ix, iz) * 3 + c];
    return SEMSKS_OK;
}

semsks_status semsks_get_snapshot(semsks_engine *e, int k, float *out)
{
    semsks_status st = semsks_get_field(e, k, out);
    if (st != SEMSKS_OK) return st;
    const Slot &s = e->slots[k];
    const int it = e->h_levels[k];
    /* injects(): the current level holds the next step's injection */
    if (!(it < s.meta.nst - 1 && it + 1 < s.nt_sem)) return SEMSKS_OK;
    const int nbl = s.meta.neh * s.meta.hpn, nb = s.nx - nbl, nl = s.nz - nbl - 1;
    std::vector<float> pre((size_t)(nb + nl) * 3);
    HIPCHK(hipMemcpyAsync(pre.data(), s.d.pre_inj, pre.size() * 4, hipMemcpyDeviceToHost, e->stream));
    HIPCHK(hipStreamSynchronize(e->stream));
    for (int i = 0; i < nb + nl; ++i) {
        const int ix = i < nb ? nbl + i : nbl, iz = i < nb ? nbl : nbl + 1 + (i - nb);
        for (int c = 0; c < 3; ++c) out[((size_t)c * s.nz + iz) * s.nx + ix] = pre[3 * i + c];
    }
    return SEMSKS_OK;
}

semsks_status semsks_stiff(semsks_engine *e, int k, const float *u, float *f)
{
    if (!e || k < 0 || k >= (int)e->slots.size() || !e->slots[k].used) return semsks_set_error(SEMSKS_EINVAL, "bad slot %d", k);
    const Slot &s = e->slots[k];
    const size_t nodes = s.plane;
    struct Scratch {           /* freed on every return (hipFree waits for the device) */
        void *p[4] = {nullptr, nullptr, nullptr, nullptr};
        ~Scratch() { for (void *q : p) if (q) hipFree(q); }
    } sc;
    float *buf = nullptr;
    HIPCHK(hipMalloc(&buf, nodes * 6 * sizeof(float) + 256));
    sc.p[0] = buf;
    HIPCHK(hipMemsetAsync(buf, 0, nodes * 6 * sizeof(float), e->stream));
    InstDesc d = s.d;
    std::vector<float> aos(nodes * 3, 0.f);
    for (int c = 0; c < 3; ++c)
        for (int iz = 0; iz < s.nz; ++iz)
            for (int ix = 0; ix < s.nx; ++ix)
                aos[node_at(s.bw, s.nz, ix, iz) * 3 + c] = u[((size_t)c * s.nz + iz) * s.nx + ix];
    d.u[0] = buf;
    for (int c = 0; c < 3; ++c) d.fout[c] = buf + (3 + c) * nodes;
    HIPCHK(hipMemcpyAsync(d.u[0], aos.data(), aos.size() * 4, hipMemcpyHostToDevice, e->stream));
    d.active = 1;
    InstDesc *dd = nullptr;
    int4 *tl = nullptr;
    int *nt = nullptr;
    std::vector<int4> t;
    build_tiles(s.share >= 0 ? e->slots[s.share] : s, 0, tile_rows(e, s.meta.nez), t);
    const int cnt = (int)t.size();
    HIPCHK(hipMalloc(&dd, sizeof d));
    sc.p[1] = dd;
    HIPCHK(hipMalloc(&tl, t.size() * sizeof(int4)));
    sc.p[2] = tl;
    HIPCHK(hipMalloc(&nt, sizeof cnt));
    sc.p[3] = nt;
    HIPCHK(hipMemcpyAsync(dd, &d, sizeof d, hipMemcpyHostToDevice, e->stream));
    HIPCHK(hipMemcpyAsync(tl, t.data(), t.size() * sizeof(int4), hipMemcpyHostToDevice, e->stream));
    HIPCHK(hipMemcpyAsync(nt, &cnt, sizeof cnt, hipMemcpyHostToDevice, e->stream));
    launch_tiles<true>(e, s.meta.hpn, dd, e->d_levels, tl, nt);
    HIPCHK(hipGetLastError());
    std::vector<float> fb(nodes * 3);
    HIPCHK(hipMemcpyAsync(fb.data(), d.fout[0], fb.size() * 4, hipMemcpyDeviceToHost, e->stream));
    HIPCHK(hipStreamSynchronize(e->stream));
    for (int c = 0; c < 3; ++c)
        for (int iz = 0; iz < s.nz; ++iz)
            for (int ix = 0; ix < s.nx; ++ix)
                f[((size_t)c * s.nz + iz) * s.nx + ix] = fb[c * nodes + node_at(s.bw, s.nz, ix, iz)];
    return SEMSKS_OK;
}

semsks_status semsks_stiff_dev(semsks_engine *e, int k, const float *u_dev, float *f_dev)
{
    if (!e || k < 0 || k >= (int)e->slots.size() || !e->slots[k].used) return semsks_set_error(SEMSKS_EINVAL, "bad slot %d", k);
    if (!u_dev || !f_dev) return semsks_set_error(SEMSKS_EINVAL, "null field or force pointer");
    const Slot &s = e->slots[k];
    const Slot &o = s.share >= 0 ? e->slots[s.share] : s;
    const size_t nodes = s.plane;
    if (!o.d_fdesc || !e->scratch)
        return semsks_set_error(SEMSKS_EINVAL, "mesh of %zu nodes per plane needs %zu bytes of scratch, limit %zu",
                                nodes, scratch_bytes(nodes), (size_t)SEMSKS_SCRATCH_BYTES);
    unsigned char *base = (unsigned char *)e->scratch;
    float *ub = (float *)base, *fb = (float *)(base + scratch_f_offset(nodes));
    /* the scratch keeps the previous call's contents: clear both regions
       (fill kernels, so the calls could be captured in a graph) */
    hipLaunchKernelGGL(fill_kernel, dim3(1024), dim3(256), 0, e->stream, ub, nodes * 3, 0.f);
    hipLaunchKernelGGL(fill_kernel, dim3(1024), dim3(256), 0, e->stream, fb, nodes * 3, 0.f);
    hipLaunchKernelGGL(soa_to_blocks_kernel, dim3(1024), dim3(256), 0, e->stream, s.nx, s.nz, s.bw, u_dev, ub);
    launch_tiles<true>(e, s.meta.hpn, o.d_fdesc, e->d_levels, o.d_ftiles, o.d_fcount);
    hipLaunchKernelGGL(blocks_to_soa_kernel, dim3(1024), dim3(256), 0, e->stream, s.nx, s.nz, s.bw, nodes, fb, f_dev);
    HIPCHK(hipGetLastError());
    HIPCHK(hipStreamSynchronize(e->stream));
    return SEMSKS_OK;
}

semsks_status semsks_get_setup(semsks_engine *e, int k, double *mass, float *mdt2, double *ab)
{
    if (!e || k < 0 || k >= (int)e->slots.size() || !e->slots[k].used) return semsks_set_error(SEMSKS_EINVAL, "bad slot %d", k);
    const Slot &s = e->slots[k];
    const Slot &o = s.share >= 0 ? e->slots[s.share] : s;
    const size_t N = (size_t)s.nx * s.nz;
    double *w64 = nullptr, *dm = nullptr, *dab = nullptr;
    float *md = nullptr, *ea = nullptr, *eb = nullptr;
    const int ne = semsks_n_edge(s.nx, s.nz, s.meta.hpn);
    HIPCHK(hipMalloc(&w64, SEMSKS_MAXNP * sizeof(double)));
    HIPCHK(hipMalloc(&dm, N * sizeof(double)));
    HIPCHK(hipMalloc(&dab, N * 3 * sizeof(double)));
    HIPCHK(hipMalloc(&md, N * sizeof(float)));
    HIPCHK(hipMalloc(&ea, ne * 3 * sizeof(float)));
    HIPCHK(hipMalloc(&eb, ne * 3 * sizeof(float)));
    HIPCHK(hipMemsetAsync(dab, 0, N * 3 * sizeof(double), e->stream));
    InstDesc d = o.d, *dd = nullptr;
    HIPCHK(hipMalloc(&dd, sizeof d));
    HIPCHK(hipMemcpyAsync(dd, &d, sizeof d, hipMemcpyHostToDevice, e->stream));
    const double eh = s.meta.eh;
    hipLaunchKernelGGL(setup_tables_kernel, dim3(1), dim3(64), 0, e->stream, dd, w64, eh);
    hipLaunchKernelGGL(mass_kernel, dim3(1024), dim3(256), 0, e->stream, d, w64, (double)(eh * eh / 4.0), s.meta.sdt, dm);
    hipLaunchKernelGGL(absorb_kernel, dim3(64), dim3(256), 0, e->stream, d, w64, eh / 2.0, s.meta.sdt, s.meta.absorb, dm, dab, ea, eb);
    hipLaunchKernelGGL(expand_mdt2_kernel, dim3(256), dim3(256), 0, e->stream, d, md);
    HIPCHK(hipGetLastError());
    if (mass) HIPCHK(hipMemcpyAsync(mass, dm, N * sizeof(double), hipMemcpyDeviceToHost, e->stream));
    if (ab) HIPCHK(hipMemcpyAsync(ab, dab, N * 3 * sizeof(double), hipMemcpyDeviceToHost, e->stream));
    if (mdt2) HIPCHK(hipMemcpyAsync(mdt2, md, N * 4, hipMemcpyDeviceToHost, e->stream));
    HIPCHK(hipStreamSynchronize(e->stream));
    hipFree(w64); hipFree(dm); hipFree(dab); hipFree(md); hipFree(ea); hipFree(eb); hipFree(dd);
    return SEMSKS_OK;
}

semsks_status semsks_get_boundary_series(semsks_engine *e, int k, int n, float *g, float *gv)
{
    if (!e || k < 0 || k >= (int)e->slots.size() || !e->slots[k].used) return semsks_set_error(SEMSKS_EINVAL, "bad slot %d", k);
    const Slot &s = e->slots[k];
    if (n < 0 || n > s.meta.nst) return semsks_set_error(SEMSKS_EINVAL, "n out of range");
    const int mt = 2 * (s.nx + s.nz), sht = s.nx + s.nz;
    /* levels at or beyond nt_sem keep the last interpolated vector */
    for (int i = 0; i < n; ++i) {
        int src = std::min(i, s.nlev - 1);
        if (s.nt_sem > 0) src = std::min(src, s.nt_sem - 1);
        HIPCHK(hipMemcpyAsync(g + (size_t)i * mt, s.d.G + (size_t)src * mt, mt * 4, hipMemcpyDeviceToHost, e->stream));
        HIPCHK(hipMemcpyAsync(gv + (size_t)i * sht, s.d.Gv + (size_t)src * sht, sht * 4, hipMemcpyDeviceToHost, e->stream));
    }
    HIPCHK(hipStreamSynchronize(e->stream));
    return SEMSKS_OK;
}

semsks_status semsks_planewave(semsks_engine *e, int nex, int nez, int hpn, int neh, float eh, int nt, float dp,
                               double p, double beta, double f0, double psi, float *psv, float *sh)
{
    if (!e || hpn < 3 || hpn > 6 || nex < 1 || nez < 1 || nt < 1 || !(p * beta >= 0 && p * beta < 1))
        return semsks_set_error(SEMSKS_EINVAL, "bad plane-wave arguments");
    const int nx = nex * hpn + 1, nz = nez * hpn + 1, mt = 2 * (nx + nz), sht = nx + nz;
    if (psv || sh) {              /* to host, through temporaries of their own */
        float *dpsv = nullptr, *dsh = nullptr;
        hipError_t ae = hipMalloc(&dpsv, (size_t)nt * mt * 4);
        if (ae != hipSuccess) return alloc_failed(ae, "plane-wave records");
        ae = hipMalloc(&dsh, (size_t)nt * sht * 4);
        if (ae != hipSuccess) { hipFree(dpsv); return alloc_failed(ae, "plane-wave records"); }
        hipLaunchKernelGGL(planewave_kernel, dim3(2048), dim3(256), 0, e->setup, nex, nez, hpn, neh, (double)eh, nt,
                           (double)dp, p, beta, f0, psi, dpsv, dsh);
        HIPCHK(hipGetLastError());
        if (psv) HIPCHK(hipMemcpyAsync(psv, dpsv, (size_t)nt * mt * 4, hipMemcpyDeviceToHost, e->setup));
        if (sh) HIPCHK(hipMemcpyAsync(sh, dsh, (size_t)nt * sht * 4, hipMemcpyDeviceToHost, e->setup));
        HIPCHK(hipStreamSynchronize(e->setup));
        hipFree(dpsv); hipFree(dsh);
        return SEMSKS_OK;
    }
    /* kept on the device for the next admission: the head of the admission
       scratch (replacing records an earlier call kept) */
    e->pend_psv = e->pend_sh = nullptr;
    e->pend_nt = 0;
    e->pend_bytes = 0;
    const size_t need = pending_bytes(nt, nx, nz);
    semsks_status st = setup_scratch(e, need, 0);
    if (st != SEMSKS_OK) return st;
    float *dpsv = (float *)e->sbuf, *dsh = (float *)(e->sbuf + al256((size_t)nt * mt * 4));
    hipLaunchKernelGGL(planewave_kernel, dim3(2048), dim3(256), 0, e->setup, nex, nez, hpn, neh, (double)eh, nt,
                       (double)dp, p, beta, f0, psi, dpsv, dsh);
    HIPCHK(hipGetLastError());
    e->pend_psv = dpsv; e->pend_sh = dsh; e->pend_nt = nt; e->pend_nx = nx; e->pend_nz = nz;
    e->pend_bytes = need;
    return SEMSKS_OK;
}

size_t semsks_planewave_bytes(int nex, int nez, int hpn, int nt)
{
    return pending_bytes(nt, nex * hpn + 1, nez * hpn + 1);
}

semsks_status semsks_run_problem(const semsks_problem *p, float *traces)
{
    semsks_engine_opts o{};
    o.max_slots = 1;
    o.use_graph = 1;
    semsks_engine *e = nullptr;
    semsks_status st = semsks_engine_create(&o, &e);
    if (st) return st;
    int slot = -1;
    st = semsks_admit(e, p, -1, &slot);
    if (!st) st = semsks_advance(e, p->nst - 1, 1);
    if (!st && traces) st = semsks_get_traces(e, slot, 0, p->nst, traces);
    semsks_engine_destroy(e);
    return st;
}

}  /* extern "C" */
