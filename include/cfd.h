/*
 * cfd.h — C-ABI of the MI355X pseudo-time iteration library (libcfd_amd.so).
 *
 * The library runs the artificial-compressibility pseudo-time loop of the
 * staggered-grid finite-volume solver on one AMD Instinct MI355X (gfx950):
 * SSP-RK3 stages of u, v (QUICK or first-order upwind convection, central
 * diffusion, Darcy-Forchheimer drag, buoyancy), the pressure update, and the
 * energy (T) and species (C) transport, with their boundary conditions.
 *
 * Reference interface this replaces:
 *   c_cuda/functions.h:7-21,36 (RESU..bcC, max_reduce) -> include/cfd_legacy.h
 *   c_cuda/main.cu:203-277 (pseudo-time and physical-time loops) -> cfd_pseudo_steps,
 *     cfd_run_to_tol, cfd_physical_step below
 *   OpenMp_Multicore/main.f90:116-157 (same loops, Fortran)
 *
 * Array layout (identical to the reference C layout, c_cuda/main.cu:59-89):
 *   1-based, index 0 is padding, j (y direction) is the contiguous axis.
 *     p-like (p, pn, T, C, eps, L):  element (i,j) at i*(jmax+1)+j, i in [1,imax],   j in [1,jmax]
 *                                    allocation (imax+1)*(jmax+1) doubles
 *     u (um, um_tau, um_n_tau):      element (i,j) at i*(jmax+1)+j, i in [1,imax+1], j in [1,jmax]
 *                                    allocation (imax+2)*(jmax+1) doubles
 *     v (vm, vm_tau, vm_n_tau):      element (i,j) at i*(jmax+2)+j, i in [1,imax],   j in [1,jmax+1]
 *                                    allocation (imax+1)*(jmax+2) doubles
 *   Fortran callers (column-major, i fastest, no padding) use CFD_LAYOUT_FORTRAN in the
 *   upload/download calls; the library converts at the boundary.
 *
 * Errors: every cfd_* function returns a status code (CFD_OK == 0). A text for the last
 * failure on the calling thread is available from cfd_last_error(). A non-finite residual
 * is a failure (CFD_ERR_NONFINITE), never a converged state.
 *
 * Threading: one context per host thread; a context owns one HIP stream. Contexts on the
 * same device are independent.
 */
#ifndef CFD_AMD_CFD_H
#define CFD_AMD_CFD_H

#include <stdint.h>
#include <stddef.h>

#ifdef __cplusplus
extern "C" {
#endif

/* ---- status codes ---------------------------------------------------------------- */
#define CFD_OK               0
#define CFD_ERR_ARG         -1   /* invalid argument / size / layout */
#define CFD_ERR_HIP         -2   /* HIP runtime error (allocation, launch) */
#define CFD_ERR_NONFINITE   -3   /* a residual became NaN or Inf */
#define CFD_ERR_NOCONV      -4   /* iteration limit reached before the tolerance */
#define CFD_ERR_DIVERGED    -5   /* residual grew past the divergence guard */
#define CFD_ERR_STATE       -6   /* call out of order (no geometry, no context) */
#define CFD_ERR_UNSUPPORTED -7   /* option not available in this build */

/* ---- model switches --------------------------------------------------------------- */
/* Convection scheme of the momentum residuals. */
#define CFD_SCHEME_QUICK     0   /* QUICK interior + first-order upwind strips (equations.f90:11-113) */
#define CFD_SCHEME_UPWIND    1   /* first-order upwind everywhere (equations.f90:120-240 on every row) */

/* Boundary-condition set. */
#define CFD_BC_REFERENCE     0   /* bottom inflow v=v_i, top extrapolation, slip side walls (boundary.f90) */
#define CFD_BC_CHANNEL       1   /* left inflow u=u_in, right zero-gradient outflow (p=0), no-slip top/bottom,
                                    T Dirichlet t_bottom/t_top, species slot on the left wall */

/* Buoyancy term of the v-momentum residual. */
#define CFD_BUOY_REFERENCE   0   /* invfr2 * (1 - 1/Tbar)  (equations.f90:405) */
#define CFD_BUOY_BOUSSINESQ  1   /* ri * (Tbar - t_ref) */

/* Semantics where the two reference implementations differ. */
#define CFD_COMPAT_FORTRAN   0   /* OpenMp_Multicore semantics (default) */
#define CFD_COMPAT_CUDA      1   /* c_cuda: V strips at j=2,jmax-1; solve_P stage 2 skips j=2;
                                    residuals over the first n linear entries */

/* Viscosity model in the momentum diffusion coefficients. */
#define CFD_MU_CONST         0   /* eps/Re */
#define CFD_MU_POLY4         1   /* eps*mu(T)/Re, mu(T) = sum_k mu_coef[k] T^k, T of the physical level */
#define CFD_K_CONST          0   /* energy diffusivity 1/Pe */
#define CFD_K_POLY4          1   /* k(T)/Pe, k(T) = sum_k k_coef[k] T^k (conductivity over its reference
                                    value), T of the physical level */

/* Storage precision of the fused pseudo-step. */
#define CFD_PREC_FP64        0   /* FP64 fields and arithmetic (the drop-in) */
#define CFD_PREC_FP32        1   /* FP32 u,v,T,C storage + FP32 stencil; FP64 pressure, FP64 residual norms */

/* Field identifiers for upload / download / device pointers. */
#define CFD_F_UM        0   /* u, physical level            (u layout) */
#define CFD_F_VM        1   /* v, physical level            (v layout) */
#define CFD_F_UM_TAU    2   /* u, pseudo-time iterate       (u layout) */
#define CFD_F_VM_TAU    3   /* v, pseudo-time iterate       (v layout) */
#define CFD_F_P         4   /* p, current pressure          (p layout) */
#define CFD_F_T         5   /* T, physical level            (p layout) */
#define CFD_F_T_TAU     6   /* T, iterate                   (p layout) */
#define CFD_F_C         7   /* C, physical level            (p layout) */
#define CFD_F_C_TAU     8   /* C, iterate                   (p layout) */
#define CFD_F_EPS       9   /* porosity epsilon1            (p layout) */
#define CFD_F_LP       10   /* porous-zone switch liga_poros (p layout) */
#define CFD_NFIELDS    11

#define CFD_LAYOUT_C        0   /* the 1-based padded C layout above */
#define CFD_LAYOUT_FORTRAN  1   /* column-major (1:n, 1:m), i fastest, no padding */

typedef struct cfd_params {
    int32_t imax, jmax;         /* grid nodes in x and y (>= 6 each) */
    int32_t scheme;             /* CFD_SCHEME_* */
    int32_t bc;                 /* CFD_BC_* */
    int32_t buoyancy;           /* CFD_BUOY_* */
    int32_t compat;             /* CFD_COMPAT_* */
    int32_t steady;             /* 1: the physical level follows the iterate (phi^n == phi^tau) */
    int32_t mu_model;           /* CFD_MU_* */
    double re;                  /* Reynolds number */
    double pe;                  /* Peclet number (energy diffusivity 1/Pe) */
    double sc;                  /* Schmidt number (species diffusivity 1/(Re Sc)) */
    double darcy;               /* Darcy number */
    double cf;                  /* Forchheimer coefficient */
    double g;                   /* body force -g*eps in x-momentum */
    double invfr2;              /* 1/Fr^2 for CFD_BUOY_REFERENCE */
    double ri;                  /* Richardson number for CFD_BUOY_BOUSSINESQ */
    double t_ref;               /* reference temperature for CFD_BUOY_BOUSSINESQ */
    double v_i;                 /* reference BC set: bottom inflow velocity */
    double tinf;                /* reference BC set: bottom temperature */
    double dtau, dt;            /* pseudo and physical time steps */
    double beta;                /* artificial compressibility */
    double b_art;               /* bulk artificial viscosity */
    double u_in;                /* channel BC set: inflow velocity */
    double t_bottom, t_top;     /* channel BC set: wall temperatures */
    double slot_lo, slot_hi;    /* channel BC set: y-range of the species slot on the left wall */
    double slot_value;          /* channel BC set: species value inside the slot (0 outside) */
    double mu_coef[5];          /* CFD_MU_POLY4 coefficients c0..c4 */
    double k_coef[5];           /* CFD_K_POLY4 coefficients c0..c4 */
    int32_t k_model;            /* CFD_K_* */
    int32_t reserved;           /* 0 */
} cfd_params;

/* Reference defaults (OpenMp_Multicore/comum.f90:29-115, main.f90:105-106, input/iterations.dat):
 * Re=Pe=Sc=1, Da=1e-2, Cf=1.75/sqrt(150*0.5^3), g=9.80665, Fr=1, v_i=0.5, Tinf=1,
 * dtau=0.05, dt=0.005, beta=1300, b_art=0, QUICK, reference BCs. */
void cfd_params_default(cfd_params *p, int32_t imax, int32_t jmax);

/* Measured HBM bandwidth of this device: a plain copy kernel over `bytes` (two buffers of
 * that size), `reps` timed launches; *gbs = bytes read + bytes written per second / 1e9. The
 * roofline figures of the benchmark report it beside the nominal 8 TB/s. */
int cfd_hbm_copy_bandwidth(int64_t bytes, int32_t reps, double *gbs);

/* ---- geometry ---------------------------------------------------------------------- */
/* Node coordinates x[1..imax], y[1..jmax] (arrays of imax+1 / jmax+1 doubles, index 0 unused)
 * give the staggered midpoints and face areas by the reference formulas
 * (nonsymetric_mesh.f90:48-101). Output arrays follow the C-ABI sizes of comum.cu:133-147. */
typedef struct cfd_geometry {
    double *x, *y, *xm, *ym;
    double *areau_n, *areau_s, *areau_e, *areau_w;
    double *areav_n, *areav_s, *areav_e, *areav_w;
} cfd_geometry;

/* Host-side mesh generator: nodes on [a,b] with n points; P=1 is uniform, P<1 clusters
 * toward both ends with the tanh stretch s(eta) = P*eta + (1-P)*(1 - tanh(Q(1-eta))/tanh(Q))
 * applied on each half (nonsymetric_mesh.f90:17-29). out[1..n] written, out[0] untouched. */
int cfd_mesh_1d(int32_t n, double a, double b, double P, double Q, double *out);
/* Midpoints and areas from x,y (host arrays, C-ABI sizes). */
int cfd_geometry_derive(int32_t imax, int32_t jmax, const double *x, const double *y,
                        double *xm, double *ym,
                        double *areau_n, double *areau_s, double *areau_e, double *areau_w,
                        double *areav_n, double *areav_s, double *areav_e, double *areav_w);

/* ---- context API -------------------------------------------------------------------- */
typedef struct cfd_ctx cfd_ctx;

int  cfd_ctx_create(const cfd_params *params, int32_t precision, cfd_ctx **out);
int  cfd_ctx_destroy(cfd_ctx *ctx);
int  cfd_ctx_get_params(const cfd_ctx *ctx, cfd_params *out);
/* Replace the scalar parameters (sizes, precision and switches must not change). */
int  cfd_ctx_set_params(cfd_ctx *ctx, const cfd_params *params);
/* Host node coordinates (x[1..imax], y[1..jmax]); derives and uploads the staggered geometry. */
int  cfd_ctx_set_mesh(cfd_ctx *ctx, const double *x, const double *y);
/* Copy a host field into the context (and convert from `layout`). */
int  cfd_ctx_upload(cfd_ctx *ctx, int32_t field, const double *host, int32_t layout);
/* Copy a context field to the host. Fields CFD_F_* plus CFD_F_PN (last pressure update). */
int  cfd_ctx_download(cfd_ctx *ctx, int32_t field, double *host, int32_t layout);
/* Device pointer of a field (FP64 contexts only; the pointer changes after each step,
 * because the iterate and the new level swap by pointer). */
int  cfd_ctx_device_ptr(cfd_ctx *ctx, int32_t field, void **out);
/* Reset pn := p, T_tau := T, C_tau := C, um_tau := um, vm_tau := vm on the device (the
 * reference leaves pn, T_tau, C_tau unset before the first pseudo-step, main.f90:105-109). */
int  cfd_ctx_sync_levels(cfd_ctx *ctx);

/* Run `nsteps` pseudo-steps (U, V, P, T, C, then the tau-level swap), recording the
 * residuals max|r3| of u, v, p after the last step into res[0..2] (may be NULL). With
 * history != NULL, history[3*k+0..2] receives the residuals of step k. No host
 * synchronisation happens between steps unless history is requested. */
int  cfd_pseudo_steps(cfd_ctx *ctx, int32_t nsteps, double res[3], double *history);

/* Iterate until max(res_u,res_v,res_p) < tol (test skipped on the first step, as
 * main.f90:142), checking every `check_every` steps, at most itc_max steps. A residual that
 * exceeds `diverge_factor` times the smallest residual seen returns CFD_ERR_DIVERGED
 * (diverge_factor <= 0 disables the guard). */
int  cfd_run_to_tol(cfd_ctx *ctx, double tol, int64_t itc_max, int32_t check_every,
                    double diverge_factor, int64_t *iters, double res[3]);

/* One physical time step: pseudo-iterate to `eps` (at most itc_max steps, first-step
 * quirk of main.f90:77,182 reproduced when first != 0), then um=um_n_tau, vm=vm_n_tau,
 * T=T_n_tau, C=C_n_tau (main.f90:154-157). */
int  cfd_physical_step(cfd_ctx *ctx, double eps, int64_t itc_max, int32_t first,
                       int64_t *iters, double res[3]);

/* Capture `steps_per_graph` fused pseudo-steps into a hipGraph and replay it `replays`
 * times (steps_per_graph must be even so the ping-pong buffers return to their slots). */
int  cfd_pseudo_steps_graph(cfd_ctx *ctx, int32_t steps_per_graph, int32_t replays, double res[3]);

/* Mean kernel time per pseudo-step of the last cfd_pseudo_steps / _graph call, measured
 * with HIP events on the context's stream (milliseconds). */
int  cfd_ctx_last_timing(const cfd_ctx *ctx, double *ms_per_step);

/* Which kernel implementation the fused step uses (0 = per-level kernels, 1 = row-streaming
 * single-pass kernel). Mainly for tests that pin both against each other. */
int  cfd_ctx_set_kernel(cfd_ctx *ctx, int32_t kernel);

/* Two-stream pipeline of the fused path (default on for contexts from cfd_ctx_create):
 * for unsteady problems the T/C kernel of pseudo-step n runs on a second stream while the
 * u/v/p kernel of step n+1 runs, since the latter reads only u, v, p and the physical-level
 * T. Results are identical either way; every call that steps joins the streams before it
 * returns. Off = one stream. Steady problems always use one stream. */
int  cfd_ctx_set_pipeline(cfd_ctx *ctx, int32_t on);

/* Shapes of the fused kernels. Zeros select the built-in choice (one row block per CU,
 * no register-budget override). */
typedef struct cfd_tiles {
    int32_t rows_uvp;    /* row-block height of the u/v/p kernel, 0 or 1..imax            */
    int32_t rows_tc;     /* row-block height of the T/C kernel, 0 or 1..imax              */
    int32_t waves_uvp;   /* register budget of the u/v/p kernel: 0|1 = one wave per SIMD,
                            2 = two waves per SIMD (excess registers spilled to scratch),
                            3 = u and v chains in separate waves of a 512-lane workgroup */
    int32_t waves_tc;    /* register budget of the T/C kernel: 0|1 = as compiled (3 waves
                            per SIMD), 4 = four waves per SIMD (with spills),
                            2 = T and C chains in separate waves of a 512-lane workgroup */
} cfd_tiles;

int  cfd_ctx_set_tiles(cfd_ctx *ctx, const cfd_tiles *tiles);
int  cfd_ctx_get_tiles(const cfd_ctx *ctx, cfd_tiles *tiles);

/* Autotuner: times `steps` fused pseudo-steps for each candidate shape on a snapshot of the
 * current state, restores the state (contents, step counters and first-step semantics),
 * and keeps the fastest shape. The result is cached per (imax, jmax, precision, model
 * switches) for the process, so later contexts of the same grid class start tuned.
 * out (may be NULL) receives the shape, out_ms (may be NULL) its mean time per step. */
int  cfd_ctx_autotune(cfd_ctx *ctx, int32_t steps, cfd_tiles *out, double *out_ms);

/* Pseudo-steps taken by each path since the context was created: fused kernels and the
 * exact per-operator path (the latter runs the first step after new inputs, and every step
 * with b_art != 0 or CFD_COMPAT_CUDA). */
/* L2 residual norms (no counterpart in the reference, whose equations.f90:298,600,662 take
 * MAXVAL(ABS(res_*)) only). With on = 1 every later pseudo-step also folds, in the same pass
 * as the maxima, the sums of squares of the final-stage increments r3 of u, v and p over the
 * update regions (the entries whose maximum is the step's residual); FP64 in both
 * precisions, summed in no fixed order (the last bits may vary from run to run). Off by
 * default: the folding costs the two-wave u/v/p kernel register spills.
 * cfd_ctx_last_residual_l2 returns sqrt of the sums of the last step (CFD_ERR_STATE when
 * off). */
int  cfd_ctx_set_residual_l2(cfd_ctx *c, int32_t on);
int  cfd_ctx_last_residual_l2(cfd_ctx *c, double l2[3]);
int  cfd_ctx_step_counts(const cfd_ctx *ctx, int64_t *fused, int64_t *exact);

/* `nsteps` pseudo-steps like cfd_pseudo_steps, with events around each launch of the fused
 * path: ms[0] = u/v/p kernel, ms[1] = T/C kernel, ms[2] = boundary pass, ms[3] = whole step,
 * each the mean over the fused steps (steps that take the exact path are not timed).
 * The steps run back to back on one stream (no two-stream pipeline) and the host waits
 * once at the end; nsteps <= 100000. For measurement, not production runs. */
int  cfd_pseudo_steps_timed(cfd_ctx *ctx, int32_t nsteps, double res[3], double ms[4]);

/* ---- mesh, porous flags, nodal fields and state files (either side of the iteration) -- */
/* The reference mesh of nonsymetric_mesh.f90:7-44: x over [-Lhori/2, Lhori/2] and y from
 * -y_down, both with spacing dx_c = Lhori/(imax-1), then tanh-stretched about the centre
 * (x, parameters Px, Q) and about the row int(y_down/dx_c)+1 (y, parameters Py, Q), in the
 * reference's update order. x[1..imax], y[1..jmax] (index 0 unused). */
int cfd_mesh_reference(int32_t imax, int32_t jmax, double lhori, double y_down, double px, double py, double q,
                       double *x, double *y);
/* Porous-block flags of nonsymetric_mesh.f90:165-218 for the block x <= x_edge, y >= y_edge:
 * interior, boundary and boundary-side cells are porous (eps = porosity, L = 1), the rest
 * fluid (eps = 1, L = 0). eps, lp (and flags, may be NULL: 0 fluid, 1 boundary, 2 interior,
 * 3 boundary side) are p-layout host arrays. */
int cfd_porous_flags_reference(int32_t imax, int32_t jmax, const double *x, const double *y, double x_edge,
                               double y_edge, double porosity, double *eps, double *lp, int32_t *flags);
/* comp_mean.f90:1-25: nodal u, v (p layout) from the staggered um, vm (C layouts), host. */
int cfd_nodal_velocity(int32_t imax, int32_t jmax, const double *um, const double *vm, double *u, double *v);

/* The final-state table of output.f90:12-18 (one record per node, i outer, j inner:
 * X Y u v p T C; u, v nodal, all p-layout host arrays). The reader accepts list-directed
 * files whose records wrap over lines (as the reference's compilers write them); any
 * output pointer may be NULL. */
int cfd_write_state_table(const char *path, int32_t imax, int32_t jmax, const double *x, const double *y,
                          const double *u, const double *v, const double *p, const double *T, const double *C);
int cfd_read_state_table(const char *path, int32_t imax, int32_t jmax, double *x, double *y, double *u, double *v,
                         double *p, double *T, double *C);

/* Restart files restartU.dat, restartV.dat, restartPTC.dat in directory `dir`, written as
 * output.f90:20-43 does (i outer). initial.f90:52-87 reads them back with j outer:
 * CFD_RESTART_ORDER_REFERENCE reproduces that read, CFD_RESTART_ORDER_WRITTEN reads in the
 * order of the writes (exact round trip). */
#define CFD_RESTART_ORDER_REFERENCE 0
#define CFD_RESTART_ORDER_WRITTEN   1
int cfd_write_restart(const char *dir, int32_t imax, int32_t jmax, const double *um, const double *vm,
                      const double *p, const double *T, const double *C);
int cfd_read_restart(const char *dir, int32_t imax, int32_t jmax, int32_t order, double *um, double *vm, double *p,
                     double *T, double *C);

/* The ParaView file of transient.f90:55-103 (data/paraview_output*.vtk): structured grid of
 * the nodes, point data |u|, T, P and the vector (u, v, 0), nodal p-layout host arrays.
 * title NULL writes the reference's title line. */
int cfd_write_vtk(const char *path, const char *title, int32_t imax, int32_t jmax, const double *x, const double *y,
                  const double *u, const double *v, const double *p, const double *T);

/* ---- batched solves ------------------------------------------------------------------
 * Many independent cases (each its own grid size, parameters, mesh and fields) advance
 * together: one persistent launch per kernel walks the tiles of every case, the per-case
 * exit test of the pseudo-time loop runs on the device, and a run replays a captured
 * hipGraph of `steps_per_graph` pseudo-steps until every case has stopped. Fields are
 * stored in the batch precision (CFD_PREC_FP64, or CFD_PREC_FP32 with FP64 pressure and
 * residual maxima). Every case must share `steady` and `mu_model`, k_model must be
 * CFD_K_CONST; b_art must be 0 and
 * compat CFD_COMPAT_FORTRAN (batches run the fused kernels only). Arrays use the C layout
 * of one case (see above); uploads and downloads are host arrays of doubles. */
typedef struct cfd_batch cfd_batch;

int  cfd_batch_create(int32_t ncases, const cfd_params *params, int32_t precision, cfd_batch **out);
int  cfd_batch_destroy(cfd_batch *b);
/* node coordinates x[1..imax], y[1..jmax] of one case (index 0 unused) */
int  cfd_batch_set_mesh(cfd_batch *b, int32_t c, const double *x, const double *y);
/* Fields of one case. Uploads write the current pseudo-time iterate (um_tau, ...) or the
 * physical level (um, ...; for steady batches the two are the same array), eps and L.
 * Downloads of um_tau/vm_tau/p/T_tau/C_tau return the case's latest iterate (the step's
 * output when it stopped on convergence, as the reference exits without the update). */
int  cfd_batch_upload(cfd_batch *b, int32_t c, int32_t field, const double *host);
int  cfd_batch_download(cfd_batch *b, int32_t c, int32_t field, double *host);
/* Writes the boundary cells of every case's current iterates from their interior values
 * (channel BC set), so that a raw initial state (zeros, a warm start) is BC-consistent. */
int  cfd_batch_apply_bc(cfd_batch *b);
int  cfd_batch_set_tiles(cfd_batch *b, const cfd_tiles *tiles);
int  cfd_batch_get_tiles(const cfd_batch *b, cfd_tiles *tiles);
/* Times candidate tile shapes (row heights and register variants of both kernels) with
 * `steps`-step runs on the current state, restores it, and keeps the fastest. */
int  cfd_batch_autotune(cfd_batch *b, int32_t steps, cfd_tiles *out, double *ms_per_step);
/* Runs every case until it stops: converged (itc != 1 and max(res_u, res_v, res_p) < tol,
 * status 1), non-finite residual (2), residual above diverge_factor x its best so far (3,
 * diverge_factor <= 0 disables), or itc_max pseudo-steps (4). tol <= 0 runs exactly
 * itc_max steps. iters[c], res[3c..3c+2] and status[c] (each may be NULL) receive the
 * per-case outcome. Returns CFD_ERR_NONFINITE if a case failed, CFD_ERR_NOCONV if a case
 * did not converge (tol > 0), CFD_OK otherwise. A later run continues from the stopped
 * states. */
int  cfd_batch_run(cfd_batch *b, double tol, int64_t itc_max, int32_t steps_per_graph, double diverge_factor,
                   int64_t *iters, double *res, int32_t *status);
/* Continues the last run: cases it stopped at itc_max go on (their step counts keep
 * growing, itc_max is the new total), the others stay stopped and leave the kernels' tile
 * lists, so the workgroups go to the cases still running (same results as one run). Lets a
 * caller run a long solve in slices, and change the row blocks between slices
 * (cfd_batch_set_tiles) as cases drop out. */
int  cfd_batch_resume(cfd_batch *b, double tol, int64_t itc_max, int32_t steps_per_graph, double diverge_factor,
                      int64_t *iters, double *res, int32_t *status);
/* Copies every case's current iterates and porosity from src into dst (same cases, same
 * grid sizes), converting precision on the device: e.g. an FP32 batch that has reached the
 * residual its storage can resolve hands over to an FP64 batch for the last decades. */
int  cfd_batch_copy_state(cfd_batch *dst, cfd_batch *src);
/* use_graph 1 (default): a run captures steps_per_graph pseudo-steps in a hipGraph once and
 * replays it; 0: the same launches are issued directly each round (same kernels, same
 * results; for profilers and callers that do not use graphs). */
int  cfd_batch_set_graph(cfd_batch *b, int32_t use_graph);
/* Stops case c between runs with status `status` (5 or above; 5 is used for a case the caller
 * judged stalled): a later cfd_batch_resume leaves it stopped, a new cfd_batch_run restarts
 * it. Lets a caller apply its own stop tests between the slices of a long solve. */
int  cfd_batch_stop_case(cfd_batch *b, int32_t c, int32_t status);
/* Wall time of the last run (HIP events on the batch stream), graph steps replayed, and
 * cell-updates done (sum over cases of pseudo-steps x imax x jmax). */
int  cfd_batch_last_timing(const cfd_batch *b, double *ms, int64_t *steps, int64_t *cell_updates);

/* ---- learned warm start --------------------------------------------------------------
 * A 4-level convolutional U-Net (channels 32/64/128/256, 3x3 kernels, 2x2 max pooling,
 * nearest upsampling with skip concatenation) trained and evaluated with hand-written HIP
 * kernels: BF16 activations, BF16 MFMA convolutions with FP32 accumulation, FP32 master
 * weights and Adam. Inputs [B][H][W][5] (mask L, porosity, log Re, log Da, Ri per cell),
 * outputs [B][H][W][4] (u, v, p, T), host float arrays; H and W multiples of 8. */
typedef struct cfd_unet cfd_unet;
int  cfd_unet_create(int32_t height, int32_t width, int32_t max_batch, uint64_t seed, cfd_unet **out);
int  cfd_unet_destroy(cfd_unet *net);
int  cfd_unet_num_params(const cfd_unet *net, int64_t *n);
/* Flat FP32 parameters: per convolution, weights [Cout][3][3][Cin] then biases [Cout], in
 * network order (encoder, bottleneck, decoder, head). set_params also resets Adam. */
int  cfd_unet_get_params(cfd_unet *net, float *host);
int  cfd_unet_set_params(cfd_unet *net, const float *host);
int  cfd_unet_predict(cfd_unet *net, const float *x, int32_t batch, float *y);
/* Mean squared error over all outputs and its gradient w.r.t. the parameters (grad may be
 * NULL: it then stays on the device for the next update). */
int  cfd_unet_loss_grad(cfd_unet *net, const float *x, const float *target, int32_t batch, double *loss, float *grad);
/* One Adam step (beta 0.9, 0.999, eps 1e-8) on the batch; *loss before the update. */
int  cfd_unet_train_step(cfd_unet *net, const float *x, const float *target, int32_t batch, float lr, double *loss);
/* Device time of the last cfd_unet_train_step from HIP events on the NULL stream:
 * ms[0] forward (with the loss), ms[1] backward, ms[2] Adam update and weight refresh
 * (host-to-device input copies excluded). Waits for that step. */
int  cfd_unet_last_timing(cfd_unet *net, double ms[3]);

/* The U-Net's building blocks on device pointers (BF16 as uint16 bit patterns, NHWC,
 * NULL stream): 3x3 convolution (stride 1, zero padding 1; weights [Cout][9 Cin]) with
 * optional bias and ReLU into BF16 y and/or FP32 y32; its weight gradient accumulated into
 * FP32 dw [Cout][9 Cin] and, when db is not NULL, its bias gradient (the sum of dy over all
 * pixels) accumulated into FP32 db [Cout] in the same pass; the flipped weights that turn the forward kernel into the input
 * gradient; 2x2 max pooling and its backward (gradient to the first maximal input, added
 * into FP32 gx); nearest x2 upsampling into channels cofs.. of a Cy-channel tensor, and its
 * backward (sum of the 2x2 children, added into FP32 gx). */
int  cfd_nn_conv3x3(const uint16_t *x, int32_t B, int32_t H, int32_t W, int32_t Cin, const uint16_t *w,
                    const float *bias, int32_t Cout, int32_t relu, uint16_t *y, float *y32);
int  cfd_nn_conv3x3_wgrad(const uint16_t *x, const uint16_t *dy, int32_t B, int32_t H, int32_t W, int32_t Cin,
                          int32_t Cout, float *dw, float *db);
int  cfd_nn_flip_weights(const float *w, int32_t Cin, int32_t Cout, uint16_t *wf);
int  cfd_nn_maxpool2(const uint16_t *x, int32_t B, int32_t H, int32_t W, int32_t C, uint16_t *y);
int  cfd_nn_maxpool2_bwd(const uint16_t *x, const uint16_t *y, const uint16_t *gy, int32_t B, int32_t H, int32_t W,
                         int32_t C, float *gx);
int  cfd_nn_upsample2(const uint16_t *x, int32_t B, int32_t H, int32_t W, int32_t C, uint16_t *y, int32_t Cy,
                      int32_t cofs);
int  cfd_nn_upsample2_bwd(const uint16_t *gy, int32_t B, int32_t H, int32_t W, int32_t C, int32_t Cy, int32_t cofs,
                          float *gx);

const char *cfd_last_error(void);
const char *cfd_version(void);

#define CFD_F_PN 11   /* download only: the pressure written by the last step */

#ifdef __cplusplus
}
#endif
#endif /* CFD_AMD_CFD_H */
