/*
 * cfd_legacy.h — the reference's hot-path entry points, same names and signatures.
 *
 * Each declaration replaces the function of the same name in the reference's CUDA C
 * version (prototypes at c_cuda/functions.h:7-21,36; bodies in c_cuda/resu.cu, resv.cu,
 * resz.cu, resc.cu, upwind_U_pair.cu, upwind_V_pair.cu, solve_U.cu, solve_V.cu, solve_P.cu,
 * solve_Z.cu, solve_C.cu, bcUV.cu, bcP.cu, bcZ.cu, bcC.cu, max_reduce.cu, comum.cu:131-212).
 * The arithmetic follows OpenMp_Multicore/equations.f90 and boundary.f90 (the Fortran
 * semantics; set cfd_params.compat = CFD_COMPAT_CUDA for the c_cuda differences: V upwind
 * strip at j=jmax-1, solve_P stage 2 from j=3, residuals over the first n linear entries).
 *
 * Arrays are DEVICE pointers in the reference C layout (see cfd.h), owned by the caller,
 * as in c_cuda/main.cu:59-89. The reference's implicit state (imax, jmax, geometry arrays,
 * Re, Da, Cf, g, dtau, dt, beta, b_art ... of c_cuda/comum.h:17-45) lives in a default
 * context set up once with cfd_legacy_init() and cfd_legacy_set_mesh(). All work runs on
 * the default (NULL) stream, ordered with the caller's own launches; the functions that
 * return a double (solve_U, solve_V, solve_P, max_reduce) synchronise to read it back.
 * Errors are reported through cfd_last_error(); a failed call returns NaN where it returns
 * a double.
 */
#ifndef CFD_AMD_LEGACY_H
#define CFD_AMD_LEGACY_H
#include "cfd.h"

#ifdef __cplusplus
extern "C" {
#endif

/* Default context: replaces the globals of c_cuda/comum.cu:6-110 set by calcular(), init()
 * and mesh(). x[1..imax], y[1..jmax] node coordinates (host); eps, lp host arrays in the p
 * layout (NULL: eps = 1, lp = 0). */
int  cfd_legacy_init(const cfd_params *params);
int  cfd_legacy_set_params(const cfd_params *params);
int  cfd_legacy_set_mesh(const double *x, const double *y, const double *eps, const double *lp);
cfd_ctx *cfd_legacy_context(void);

/* c_cuda/comum.cu:131-174 / 176-212. alocar_globais() creates the default context with the
 * reference defaults (51 x 84) if cfd_legacy_init() was not called; desalocar_globais()
 * releases it. */
void alocar_globais(void);
void desalocar_globais(void);

/* c_cuda/functions.h:9-12 (equations.f90:2-116, 305-420, 723-765, 823-867) */
void RESU(double *dev_um, double *dev_vm, double *dev_p, double *dev_ru);
void RESV(double *dev_um, double *dev_vm, double *dev_p, double *dev_t, double *dev_rv);
void RESZ(double *dev_um_n, double *dev_vm_n, double *dev_z, double *dev_rz);
void RESC(double *dev_um_n, double *dev_vm_n, double *dev_c, double *dev_rc);

/* c_cuda/functions.h:7-8: the first-order upwind strips of RESU / RESV only
 * (equations.f90:109-113, 413-417). */
void upwind_U_pair(double *dev_um, double *dev_vm, double *dev_p, double *dev_ru);
void upwind_V_pair(double *dev_um, double *dev_vm, double *dev_p, double *dev_t, double *dev_rv);

/* c_cuda/functions.h:13-17 (equations.f90:244-301, 546-603, 607-665, 669-719, 769-819).
 * solve_U / solve_V write dev_um_n_tau / dev_vm_n_tau and rewrite the boundary of
 * dev_vm_tau / dev_um_tau, as the reference's bcUV calls do. dev_um_n, dev_vm_n, and the
 * other-velocity physical level are unused, as in the reference. */
void   solve_Z(double *dev_um_n, double *dev_vm_n, double *dev_z, double *dev_z_n_tau, double *dev_z_tau, double *dev_rz);
void   solve_C(double *dev_um_n, double *dev_vm_n, double *dev_c, double *dev_c_n_tau, double *dev_c_tau, double *dev_rc);
double solve_U(double *dev_um, double *dev_vm, double *dev_um_n, double *dev_um_tau, double *dev_vm_tau,
               double *dev_um_n_tau, double *dev_p, double *dev_ui, double *dev_ru, double *dev_res_u);
double solve_V(double *dev_um, double *dev_vm, double *dev_vm_n, double *dev_um_tau, double *dev_vm_tau,
               double *dev_vm_n_tau, double *dev_p, double *dev_t, double *dev_vi, double *dev_rv, double *dev_res_v);
double solve_P(double *dev_p, double *dev_um_n, double *dev_vm_n, double *dev_pn);

/* c_cuda/functions.h:18-21 (boundary.f90) */
void bcUV(double *dev_um, double *dev_vm);
void bcP(double *dev_pn);
void bcZ(double *dev_zt);
void bcC(double *dev_cn);

/* c_cuda/functions.h:36 (max_reduce.cu:14-41): max |a| over the first imax*jmax entries;
 * a NaN entry makes the result NaN (the reference's fmax would drop it). */
double max_reduce(double *dev_matriz_linearizada, int imax, int jmax);

/* No counterpart in the reference: max_reduce for callers that cannot allocate device memory
 * (code under HIP graph capture, serving loops, harnesses that forbid hipMalloc). The same
 * max |a| over the first n_i*n_j linear entries (NaN wins, 0 for no entries), written to the
 * caller's device double dev_out on `stream` (a hipStream_t; NULL: the default stream),
 * without synchronising and without any allocation: the partial maxima live in a static
 * 8 KiB device array of the library (one per GPU, kept until the process ends). Needs no
 * default context. n_i*n_j at most 2^36 entries. Calls that overlap on one GPU share that
 * array and race. Returns CFD_OK or an error code (cfd_last_error()). */
int max_reduce_noalloc(const double *dev_a, int n_i, int n_j, double *dev_out, void *stream);

#ifdef __cplusplus
}
#endif
#endif
