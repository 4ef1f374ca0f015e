/*
 * cfd_oracle.h — CPU oracle of the pseudo-time iteration (TEST INFRASTRUCTURE ONLY).
 *
 * A plain C, FP64 restatement of the reference's hot path, written to check the HIP
 * library. Only tests/, __graft_entry__.smoke() and bench.py's cpu_baseline leg use it,
 * and only as the checker or the timed CPU baseline; the product never links it.
 *
 * Every routine follows OpenMp_Multicore/equations.f90 and boundary.f90 statement by
 * statement (same loop ranges, same operation order), on the 1-based padded C layout of
 * c_cuda/main.cu:59-89. Build flags: -O2 -ffp-contract=off, so that each operation is one
 * IEEE FP64 rounding, as in the reference built without contraction.
 *
 * Parity pins (see tests/): the flang-built reference (oracle/ref/build_ref.sh) on seeded
 * inputs (tests/golden/*.bin), closed forms (uniform flow, linear pressure, telescoping
 * continuity, RK3 algebra) and brute-force tiny grids.
 */
#ifndef CFD_ORACLE_H
#define CFD_ORACLE_H
#include "../include/cfd.h"

#ifdef __cplusplus
extern "C" {
#endif

typedef struct orc_ctx orc_ctx;

/* Geometry from node coordinates (x[1..imax], y[1..jmax]) + porosity fields (p layout). */
orc_ctx *orc_create(const cfd_params *prm, const double *x, const double *y,
                    const double *eps, const double *lp);
/* Same, with every staggered array given explicitly (for feeding the reference's own geometry). */
orc_ctx *orc_create_full(const cfd_params *prm, const double *x, const double *y,
                         const double *xm, const double *ym,
                         const double *aun, const double *aus, const double *aue, const double *auw,
                         const double *avn, const double *avs, const double *ave, const double *avw,
                         const double *eps, const double *lp);
void orc_destroy(orc_ctx *o);
void orc_set_params(orc_ctx *o, const cfd_params *prm);
void orc_get_geometry(const orc_ctx *o, double *xm, double *ym, double *aun, double *aue,
                      double *avn, double *ave);

/* Residual operators (equations.f90:2-116, 305-420, 723-765, 823-867). T is used by the
 * v buoyancy and by the mu(T) model. */
void orc_resu(orc_ctx *o, const double *um, const double *vm, const double *p, const double *T, double *ru);
void orc_resv(orc_ctx *o, const double *um, const double *vm, const double *p, const double *T, double *rv);
void orc_resz(orc_ctx *o, const double *um, const double *vm, const double *z, double *rz);
/* RESZ with the k(T) diffusivity (k_model CFD_K_POLY4) evaluated at the temperature T. */
void orc_resz_kt(orc_ctx *o, const double *um, const double *vm, const double *z, const double *T, double *rz);
void orc_resc(orc_ctx *o, const double *um, const double *vm, const double *c, double *rc);
/* Upwind rows/columns (equations.f90:120-240, 423-542). */
void orc_upwind_ui(orc_ctx *o, const double *um, const double *vm, const double *p, const double *T, double *ru, int j);
void orc_upwind_uj(orc_ctx *o, const double *um, const double *vm, const double *p, const double *T, double *ru, int i);
void orc_upwind_vi(orc_ctx *o, const double *um, const double *vm, const double *p, const double *T, double *rv, int j);
void orc_upwind_vj(orc_ctx *o, const double *um, const double *vm, const double *p, const double *T, double *rv, int i);

/* Boundary conditions (boundary.f90), in place. */
void orc_bcuv(orc_ctx *o, double *um, double *vm);
void orc_bcp(orc_ctx *o, double *p);
void orc_bcz(orc_ctx *o, double *z);
void orc_bcc(orc_ctx *o, double *c);

/* RK3 solves (equations.f90:244-301, 546-603, 607-665, 669-719, 769-819). um_tau/vm_tau
 * are modified on their boundaries exactly as the reference's bcUV side effects do. */
double orc_solve_u(orc_ctx *o, const double *um, double *um_tau, double *vm_tau, double *um_n_tau,
                   const double *p, const double *T);
double orc_solve_v(orc_ctx *o, const double *vm, double *um_tau, double *vm_tau, double *vm_n_tau,
                   const double *p, const double *T);
double orc_solve_p(orc_ctx *o, const double *p, const double *um_n, const double *vm_n, double *pn);
void   orc_solve_z(orc_ctx *o, const double *um_n, const double *vm_n, const double *z,
                   double *z_n_tau, const double *z_tau);
void   orc_solve_c(orc_ctx *o, const double *um_n, const double *vm_n, const double *c,
                   double *c_n_tau, const double *c_tau);

/* Full state of the pseudo-time loop (main.f90:15-20), C layout. */
typedef struct orc_state {
    double *um, *vm, *um_tau, *vm_tau, *um_n_tau, *vm_n_tau;
    double *p, *pn, *T, *T_tau, *T_n_tau, *C, *C_tau, *C_n_tau;
} orc_state;

/* One pseudo-step of main.f90:120-151 (U, V, P, Z(T), C, residual, tau swap unless
 * `converged_exit` is returned). `steady` params alias the physical level to the iterate.
 * res[3] = (res_u, res_v, res_p). Returns the error max(res_u,res_v,res_p). */
double orc_pseudo_step(orc_ctx *o, orc_state *s, double res[3], int do_swap);
/* n pseudo-steps with swaps; history (3n doubles) optional. */
void   orc_pseudo_steps(orc_ctx *o, orc_state *s, int n, double *history);
/* One physical step (main.f90:116-157); first != 0 reproduces the itc=1 start. Returns
 * the pseudo-steps taken. */
long   orc_physical_step(orc_ctx *o, orc_state *s, double eps, long itc_max, int first, double res[3]);

/* Reference mesh formulas (nonsymetric_mesh.f90:7-44) for a centred x-domain of width L:
 * x_i = (i-1)*L/(n-1) - L/2, then the two-sided stretch about node n/2+1. */
void orc_mesh_ref_x(int n, double L, double P, double Q, double *x);

/* Max-abs over the Fortran update region of an array (i0..i1, j0..j1), row stride ld. */
double orc_maxabs(const double *a, int ld, int i0, int i1, int j0, int j1);
/* L2 norms of the last solve_U / solve_V / solve_P residuals over their update regions. */
void orc_last_res_l2(const orc_ctx *o, double l2[3]);

#ifdef __cplusplus
}
#endif
#endif
