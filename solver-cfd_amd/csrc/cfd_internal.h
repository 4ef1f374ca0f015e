/* cfd_internal.h — interface between the C host code and the HIP kernel launchers.
 * Plain C: the host side (cfd_host.c) is C11, the launchers live in the .hip files. */
#ifndef CFD_AMD_INTERNAL_H
#define CFD_AMD_INTERNAL_H
#include <stdint.h>
#include <hip/hip_runtime_api.h>
#include "../../include/cfd.h"

#ifdef __cplusplus
extern "C" {
#endif

/* Device geometry (C-ABI sizes of c_cuda/comum.cu:133-164) plus the 1-D factors the fused
 * kernels precompute from it. */
typedef struct cfd_devgeo {
    double *x, *y, *xm, *ym;
    double *aun, *aus, *aue, *auw, *avn, *avs, *ave, *avw;
    double *eps, *lp;
    int32_t imax, jmax;
} cfd_devgeo;

/* ---- per-operator kernels (reference call structure) ---------------------------------- */
/* mode 0: the whole residual; 1: only the upwind strips (upwind_U_pair / upwind_V_pair). */
int cfdk_resu(const cfd_devgeo *g, const cfd_params *P, int mode, const double *um, const double *vm,
              const double *p, const double *T, double *ru, hipStream_t s);
int cfdk_resv(const cfd_devgeo *g, const cfd_params *P, int mode, const double *um, const double *vm,
              const double *p, const double *T, double *rv, hipStream_t s);
/* species = 0: RESZ (1/Pe, or k(T)/Pe with k_model CFD_K_POLY4 at the temperature kT),
 * 1: RESC (1/(Re Sc)); kT may be NULL (constant diffusivity) */
int cfdk_ressc(const cfd_devgeo *g, const cfd_params *P, int species, const double *um, const double *vm,
               const double *z, const double *kT, double *rz, hipStream_t s);
/* RK3 stage k (1,2,3) on the u region (i=3..imax-1, j=2..jmax-1), v region (i=2..imax-1,
 * j=3..jmax-1) or scalar region (i=2..imax-1, j=2..jmax-1):
 *   res = ((phys - tau) + r*dt)*dtau ; out = tau + res | 0.75 tau + 0.25 (prev + res) |
 *   1/3 tau + 2/3 (prev + res).   res may be NULL (scalars). */
int cfdk_stage(const cfd_devgeo *g, const cfd_params *P, int region, int stage,
               const double *phys, const double *tau, const double *r, const double *prev,
               double *out, double *res, hipStream_t s);
/* solve_P stage k: stage 1 computes rp and pi; 2 updates pi; 3 writes res and pn. */
int cfdk_pstage(const cfd_devgeo *g, const cfd_params *P, int stage, const double *p,
                const double *um_n, const double *vm_n, double *rp, double *pi, double *pn,
                double *res, hipStream_t s);
/* bcUV (either pointer may be NULL to skip that half), bcP (kind 0), bcZ (1), bcC (2). */
int cfdk_bc_uv(const cfd_devgeo *g, const cfd_params *P, double *u, double *v, hipStream_t s);
int cfdk_bc_s(const cfd_devgeo *g, const cfd_params *P, int kind, double *a, hipStream_t s);
/* max |a(i,j)| over i0..i1, j0..j1 (row stride ld) folded into *out (uint64 bit pattern of a
 * non-negative double; NaN propagates as the largest pattern). *out must be cleared first. */
int cfdk_zero_u64(uint64_t *out, int n, hipStream_t s);
/* One launch that clears n residual slots and, when l2 is not NULL, the 3 sum-of-squares
 * accumulators beside them. */
int cfdk_zero_slots(uint64_t *slot, int n, double *l2, hipStream_t s);
/* sum of a(i,j)^2 over i0..i1, j0..j1 (row stride ld) added to *out (atomically, in no fixed
 * order: the last bits depend on the schedule). */
int cfdk_sumsq(const double *a, int ld, int i0, int i1, int j0, int j1, double *out, hipStream_t s);
int cfdk_maxabs(const double *a, int ld, int i0, int i1, int j0, int j1, uint64_t *out, hipStream_t s);
int cfdk_maxabs_linear(const double *a, long n, uint64_t *out, hipStream_t s);
int cfdk_copy(double *dst, const double *src, long n, hipStream_t s);
int cfdk_fill(double *dst, double v, long n, hipStream_t s);
/* Layout conversion between the padded C layout and Fortran column-major (n x m). */
int cfdk_from_fortran(double *dst_c, const double *src_f, int n, int m, int ld, hipStream_t s);
int cfdk_to_fortran(double *dst_f, const double *src_c, int n, int m, int ld, hipStream_t s);

/* ---- fused pseudo-step ------------------------------------------------------------- */
typedef struct cfd_fused_args {
    /* fields in the context precision (double, or float when prec == CFD_PREC_FP32);
     * pressure is double in both */
    const void *um, *vm, *um_tau, *vm_tau, *T, *T_tau, *C, *C_tau;
    const double *p;
    void *um_n, *vm_n, *T_n, *C_n;
    double *pn;
    const void *epsr, *lpr;        /* porosity fields in the field precision */
    /* residual accumulators: res_max[0..2] (uint64 bit patterns of max|r3| for u, v, p) and,
     * when res_l2 is not NULL, the sums of r3^2 over the same entries (3 doubles) */
    uint64_t *res_max;
    double *res_l2;
    /* if not NULL, the u/v/p kernel writes zeros to res_next[0..2]: the residual slots of the
     * next pseudo-step (never res_max itself), so that no separate clearing launch precedes
     * that step */
    uint64_t *res_next;
    int32_t prec;
    /* row-block heights of K1 / K2 (0: built-in heuristic) and the register-budget variants
     * (occ1 in {1, 2} waves per SIMD or 3 = split u/v chains, occ2 in {1, 4} waves per SIMD or
     * 2 = split T/C chains);
     * chosen by the autotuner */
    int32_t hb1, hb2, occ1, occ2;
    /* built-in heuristic for a lone grid (1) or a batch case (0): a lone grid that fills
     * fewer blocks than there are CUs takes row blocks down to 4 rows, since its step time
     * is the row count of one block (plus the stage lags) times the per-row latency; batch
     * tiles keep 24 (K1) and 16 (K2) rows, where halo rows cost throughput */
    int32_t lone;
    /* optional events recorded before K1, after K1, after K2 and after the boundary pass */
    hipEvent_t ev[4];
    /* optional second stream: K2 and the T/C boundary cells run on s2 after ev_k1 (recorded
     * on the main stream after K1), ev_k2_rec is recorded on s2 at the end, and K1 first
     * waits for ev_k2_wait (the K2 two steps back, whose input buffer K1 overwrites) */
    hipStream_t s2;
    hipEvent_t ev_k1, ev_k2_wait, ev_k2_rec;
} cfd_fused_args;

/* Precomputed 1-D geometry factors for the fused kernels (allocated by the host). */
typedef struct cfd_devfactors {
    double *rdx;     /* 1/(x_i - x_{i-1})          i = 2..imax   */
    double *rdy;     /* 1/(y_j - y_{j-1})          j = 2..jmax   */
    double *rdxm;    /* 1/(xm_i - xm_{i-1})        i = 2..imax+1 */
    double *rdym;    /* 1/(ym_j - ym_{j-1})        j = 2..jmax+1 */
    /* packed copy for the fused kernels, 4*(imax+2) + 4*(jmax+2) doubles: per row i the
     * four values aun avn rdx rdxm side by side (element 4i+q), so that every row factor a
     * kernel reads in one row iteration sits at a fixed offset from one address; then the
     * columns [aue | ave | rdy | rdym], one array after the other */
    double *pack;
} cfd_devfactors;

/* 1 when this build has the fused kernels. */
int cfdk_fused_available(void);
int cfdk_factors(const cfd_devgeo *g, cfd_devfactors *f, hipStream_t s);
/* One pseudo-step of the fused path on BC-consistent iterates (kernel 0 = per-level). */
int cfdk_fused_step(const cfd_devgeo *g, const cfd_devfactors *f, const cfd_params *P,
                    const cfd_fused_args *a, int kernel, hipStream_t s);
/* Write the boundary cells of the five step outputs (the fused path reads boundaries
 * through the BC formulas and never stores them). vm_tau_in supplies the pre-BC column
 * imax the reference's corner (imax, jmax+1) depends on. */
int cfdk_materialize_bc(const cfd_devgeo *g, const cfd_params *P, double *um_n, double *vm_n,
                        double *pn, double *T_n, double *C_n, const double *vm_prev, hipStream_t s);
/* ---- batched persistent path --------------------------------------------------------- */
/* One case of a batch: its parameters, 1-D device geometry and the fused-step arguments for
 * both tau parities (parity 1 has the tau and n buffers of parity 0 swapped). */
typedef struct cfd_batch_case {
    cfd_params P;
    const double *y;        /* device y[0..jmax] */
    const double *pack;     /* device packed 1-D factors (cfd_devfactors.pack layout) */
    cfd_fused_args a[2];
} cfd_batch_case;

/* Host-side sizes of one entry of the device argument tables (per precision). */
int cfdk_batch_arg_sizes(int32_t prec, size_t *k1, size_t *k2, size_t *mat);
/* Fill the three host tables (2 entries per case, parity-major within a case) and the tile
 * geometry: rows per tile hb1/hb2 per case (0: built-in rule). */
int cfdk_batch_build_args(int32_t prec, int32_t ncases, const cfd_batch_case *cases, const int32_t *hb1,
                          const int32_t *hb2, void *k1, void *k2, void *mat, int32_t span);
/* Column span (lanes) of the batch's workgroups for these case widths: 64 or 256. */
int32_t cfdk_batch_span(int32_t ncases, const int32_t *jmax);
/* Tiles of one case: count and fill (c, bi, bj) triples into out (4 int32 per tile). */
int cfdk_batch_case_tiles(int kernel, int32_t span, int32_t imax, int32_t jmax, int32_t hb, int32_t c,
                          int32_t *out);

typedef struct cfd_batch_launch {
    int32_t prec, steady, mupoly, occ1, occ2, ncases, max_mat;
    int32_t span;                    /* columns per workgroup of the batch kernels (64 or 256) */
    const void *k1, *k2, *mat;       /* device argument tables */
    const void *tiles1, *tiles2;     /* device tile lists (4 int32 per tile) */
    int32_t ntiles1, ntiles2, grid1, grid2;
    int32_t *ctl;                    /* per case: bit 0 active, bit 1 tau parity */
    uint64_t *res;                   /* 3 per case */
    /* convergence control (device) */
    double tol, diverge;
    int64_t itc_max;
    int32_t *status;                 /* 0 running, 1 converged, 2 non-finite, 3 diverged, 4 itc_max */
    int64_t *itc;                    /* pseudo-steps taken */
    double *best, *last;             /* best max residual; last residuals (3 per case) */
    int32_t *nactive;                /* written by cfdk_batch_count */
} cfd_batch_launch;

/* Blocks per CU the batch kernels can keep resident (for the persistent grid size). */
int cfdk_batch_occupancy(int32_t prec, int32_t steady, int32_t mupoly, int32_t occ1, int32_t occ2, int32_t span,
                         int *k1_blocks_per_cu, int *k2_blocks_per_cu);
/* One batched pseudo-step: K1, K2, boundary cells, convergence control (which clears the
 * residual slots of the cases that go on, for their next step). */
int cfdk_batch_step(const cfd_batch_launch *L, hipStream_t s);
/* Clears the residual slots of the active cases: once before the first step of a run. */
int cfdk_batch_clear(const cfd_batch_launch *L, hipStream_t s);
/* Active-case count into *L->nactive. */
int cfdk_batch_count(const cfd_batch_launch *L, hipStream_t s);
/* Boundary cells of every case's tau-level iterates (makes an uploaded state BC-consistent). */
int cfdk_batch_apply_bc(const cfd_batch_launch *L, hipStream_t s);

/* Precision conversion of n contiguous elements. */
int cfdk_d2f(float *dst, const double *src, long n, hipStream_t s);
int cfdk_f2d(double *dst, const float *src, long n, hipStream_t s);

#ifdef __cplusplus
}
#endif
#endif
