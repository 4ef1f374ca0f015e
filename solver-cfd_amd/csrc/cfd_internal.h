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
/* species = 0: RESZ (1/Pe), 1: RESC (1/(Re Sc)) */
int cfdk_ressc(const cfd_devgeo *g, const cfd_params *P, int species, const double *um, const double *vm,
               const double *z, double *rz, hipStream_t s);
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
int cfdk_maxabs(const double *a, int ld, int i0, int i1, int j0, int j1, uint64_t *out, hipStream_t s);
int cfdk_maxabs_linear(const double *a, long n, uint64_t *out, hipStream_t s);
int cfdk_copy(double *dst, const double *src, long n, hipStream_t s);
int cfdk_fill(double *dst, double v, long n, hipStream_t s);
/* Layout conversion between the padded C layout and Fortran column-major (n x m). */
int cfdk_from_fortran(double *dst_c, const double *src_f, int n, int m, int ld, hipStream_t s);
int cfdk_to_fortran(double *dst_f, const double *src_c, int n, int m, int ld, hipStream_t s);

/* ---- fused pseudo-step ------------------------------------------------------------- */
typedef struct cfd_fused_args {
    /* inputs (C layout, FP64) */
    const double *um, *vm, *um_tau, *vm_tau, *p, *T, *T_tau, *C, *C_tau;
    /* outputs */
    double *um_n, *vm_n, *pn, *T_n, *C_n;
    /* scratch for stage values (u/v layouts), per-level path only */
    double *u1, *u2, *v1, *v2, *z1, *z2, *c1, *c2;
    /* residual accumulators: res_max[0..2] (uint64 bit patterns of max|r3| for u, v, p) */
    uint64_t *res_max;
    /* row-block heights of K1 / K2 (0: built-in heuristic) and the register-budget variants
     * (occ1 in {1, 2}, occ2 in {1, 4} waves per SIMD); chosen by the autotuner */
    int32_t hb1, hb2, occ1, occ2;
    /* optional events recorded before K1, after K1, after K2 and after the boundary pass */
    hipEvent_t ev[4];
} cfd_fused_args;

/* Precomputed 1-D geometry factors for the fused kernels (allocated by the host). */
typedef struct cfd_devfactors {
    double *rdx;     /* 1/(x_i - x_{i-1})          i = 2..imax   */
    double *rdy;     /* 1/(y_j - y_{j-1})          j = 2..jmax   */
    double *rdxm;    /* 1/(xm_i - xm_{i-1})        i = 2..imax+1 */
    double *rdym;    /* 1/(ym_j - ym_{j-1})        j = 2..jmax+1 */
    /* packed copy for the fused kernels, 4*(imax+2) + 4*(jmax+2) doubles:
     * rows [aun | avn | rdx | rdxm] then columns [aue | ave | rdy | rdym] */
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

#ifdef __cplusplus
}
#endif
#endif
