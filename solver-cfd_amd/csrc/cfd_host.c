/*
 * cfd_host.c — host side of libcfd_amd: contexts, drivers and the drop-in entry points.
 *
 * Plain C11 over the HIP runtime API. Kernels are launched through the C launchers of
 * cfd_ops.hip / cfd_fused.hip (cfd_internal.h).
 *
 * Two step implementations share one state:
 *   exact  — the reference call structure (solve_U, solve_V, solve_P, solve_Z, solve_C of
 *            OpenMp_Multicore/equations.f90 with the bcUV side effects on um_tau/vm_tau),
 *            one kernel per operator; used for the first step after new inputs (when the
 *            iterates may violate the boundary conditions) and by the legacy names;
 *   fused  — cfdk_fused_step on boundary-consistent iterates.
 * The tau-level update of main.f90:145-149 is a pointer swap.
 */
#include "cfd_internal.h"
#include "../../include/cfd_legacy.h"

#include <math.h>
#include <pthread.h>
#include <stdarg.h>
#include <stdio.h>
#include <stdlib.h>
#include <string.h>

/* ------------------------------------------------------------------------------------ */
/* errors                                                                                */
static _Thread_local char g_err[512] = "";

static int fail(int code, const char *fmt, ...)
{
    va_list ap;
    va_start(ap, fmt);
    vsnprintf(g_err, sizeof g_err, fmt, ap);
    va_end(ap);
    return code;
}

#define HIPCHK(call)                                                                   \
    do {                                                                               \
        hipError_t e_ = (call);                                                        \
        if (e_ != hipSuccess)                                                          \
            return fail(CFD_ERR_HIP, "%s failed: %s", #call, hipGetErrorString(e_));   \
    } while (0)

#define CHK(call)                                                                      \
    do {                                                                               \
        int r_ = (call);                                                               \
        if (r_ != CFD_OK) {                                                            \
            if (r_ == CFD_ERR_HIP && !g_err[0])                                        \
                fail(r_, "kernel launch failed in %s: %s", #call,                      \
                     hipGetErrorString(hipGetLastError()));                            \
            return r_;                                                                 \
        }                                                                              \
    } while (0)

/* the last-error channel for the other host sources of the library */
int cfd__fail(int code, const char *fmt, ...)
{
    va_list ap;
    va_start(ap, fmt);
    vsnprintf(g_err, sizeof g_err, fmt, ap);
    va_end(ap);
    return code;
}

const char *cfd_last_error(void) { return g_err; }
const char *cfd_version(void) { return "cfd_amd 0.1 (gfx950)"; }

/* ------------------------------------------------------------------------------------ */
/* parameters and mesh                                                                  */
void cfd_params_default(cfd_params *p, int32_t imax, int32_t jmax)
{
    memset(p, 0, sizeof *p);
    p->imax = imax; p->jmax = jmax;
    p->scheme = CFD_SCHEME_QUICK; p->bc = CFD_BC_REFERENCE; p->buoyancy = CFD_BUOY_REFERENCE;
    p->compat = CFD_COMPAT_FORTRAN; p->steady = 0; p->mu_model = CFD_MU_CONST;
    p->re = 1.0; p->pe = 1.0; p->sc = 1.0; p->darcy = 1.0e-2;
    p->cf = 1.75 / sqrt(150.0 * 0.5 * 0.5 * 0.5);   /* comum.f90:31, porosidade = 0.5 */
    p->g = 9.80665; p->invfr2 = 1.0; p->ri = 0.0; p->t_ref = 0.0;
    p->v_i = 0.5; p->tinf = 1.0; p->dtau = 5.0e-2; p->dt = 0.5e-2;
    p->beta = 1300.0; p->b_art = 0.0;
    p->u_in = 1.0; p->t_bottom = 1.0; p->t_top = 0.0;
    p->slot_lo = 0.0; p->slot_hi = 0.0; p->slot_value = 1.0;
    p->mu_coef[0] = 1.0;
    p->k_coef[0] = 1.0; p->k_model = CFD_K_CONST;
}

static int check_params(const cfd_params *p)
{
    if (!p) return fail(CFD_ERR_ARG, "params is NULL");
    if (p->imax < 6 || p->jmax < 6) return fail(CFD_ERR_ARG, "grid %dx%d too small (need >= 6x6)", p->imax, p->jmax);
    if ((int64_t)p->imax * p->jmax > (int64_t)1 << 31) return fail(CFD_ERR_ARG, "grid too large");
    if (p->scheme < 0 || p->scheme > 1) return fail(CFD_ERR_ARG, "unknown scheme %d", p->scheme);
    if (p->bc < 0 || p->bc > 1) return fail(CFD_ERR_ARG, "unknown bc set %d", p->bc);
    if (p->buoyancy < 0 || p->buoyancy > 1) return fail(CFD_ERR_ARG, "unknown buoyancy model %d", p->buoyancy);
    if (p->compat < 0 || p->compat > 1) return fail(CFD_ERR_ARG, "unknown compat %d", p->compat);
    if (p->mu_model < 0 || p->mu_model > 1) return fail(CFD_ERR_ARG, "unknown mu model %d", p->mu_model);
    if (p->k_model < 0 || p->k_model > 1) return fail(CFD_ERR_ARG, "unknown conductivity model %d", p->k_model);
    if (!(p->re > 0) || !(p->pe > 0) || !(p->sc > 0) || !(p->darcy > 0))
        return fail(CFD_ERR_ARG, "Re, Pe, Sc and Da must be positive");
    return CFD_OK;
}

/* Two-sided tanh stretch of nonsymetric_mesh.f90:17-29 on [a, b]. */
int cfd_mesh_1d(int32_t n, double a, double b, double P, double Q, double *x)
{
    if (n < 3 || !x) return fail(CFD_ERR_ARG, "cfd_mesh_1d: n=%d", n);
    if (P != 1.0 && !(Q > 0)) return fail(CFD_ERR_ARG, "cfd_mesh_1d: Q must be positive");
    const double h = (b - a) / (n - 1);
    for (int i = 1; i <= n; ++i) x[i] = a + (double)(i - 1) * h;
    if (P == 1.0) return CFD_OK;
    const int c = n / 2 + 1;
    const double xn = x[n], x0 = x[1], tq = tanh(Q);
    for (int i = c; i <= n; ++i) {
        const double eta = (x[i] - xn) / (x[c] - xn);
        const double s = P * eta + (1.0 - P) * (1.0 - tanh(Q * (1.0 - eta)) / tq);
        x[i] = xn - s * (xn - x[c]);
    }
    for (int i = c; i >= 1; --i) {
        const double eta = (x[i] - x0) / (x[c] - x0);
        const double s = P * eta + (1.0 - P) * (1.0 - tanh(Q * (1.0 - eta)) / tq);
        x[i] = x0 + s * (x[c] - x0);
    }
    return CFD_OK;
}

/* nonsymetric_mesh.f90:48-101 */
int cfd_geometry_derive(int32_t imax, int32_t jmax, const double *x, const double *y,
                        double *xm, double *ym,
                        double *aun, double *aus, double *aue, double *auw,
                        double *avn, double *avs, double *ave, double *avw)
{
    if (imax < 3 || jmax < 3 || !x || !y) return fail(CFD_ERR_ARG, "cfd_geometry_derive: bad arguments");
    xm[1] = -(x[2] + x[1]) * 0.5;
    xm[imax + 1] = (x[imax] + x[imax - 1]) * 0.5 + (x[imax] - x[imax - 1]);
    for (int i = 2; i <= imax; ++i) xm[i] = (x[i] + x[i - 1]) * 0.5;
    ym[1] = y[1] - (y[2] - y[1]) * 0.5;
    ym[jmax + 1] = (y[jmax] + y[jmax - 1]) * 0.5 + (y[jmax] - y[jmax - 1]);
    for (int j = 2; j <= jmax; ++j) ym[j] = (y[j] + y[j - 1]) * 0.5;
    for (int i = 2; i <= imax; ++i) { aun[i] = x[i] - x[i - 1]; aus[i] = aun[i]; }
    for (int j = 1; j <= jmax; ++j) { aue[j] = ym[j + 1] - ym[j]; auw[j] = aue[j]; }
    for (int i = 1; i <= imax; ++i) { avn[i] = xm[i + 1] - xm[i]; avs[i] = avn[i]; }
    for (int j = 2; j <= jmax; ++j) { ave[j] = y[j] - y[j - 1]; avw[j] = ave[j]; }
    return CFD_OK;
}

/* ------------------------------------------------------------------------------------ */
/* context                                                                               */
enum { S_UM, S_VM, S_UMT, S_VMT, S_UMN, S_VMN, S_P, S_PN, S_T, S_TT, S_TN, S_C, S_CT, S_CN, S_NSTATE };
enum { X_RU, X_RV, X_RZ, X_RC, X_UI, X_VI, X_ZI, X_CI, X_PI, X_RP, X_RESU, X_RESV, X_RESP, X_NSCR };

struct cfd_ctx {
    cfd_params P;
    int32_t prec, kernel, device;
    int owns_stream;
    hipStream_t s;
    cfd_devgeo g;
    cfd_devfactors fac;
    int have_mesh, bc_clean, uv_reads_pn, first_step, have_fused, pending_in_tau;
    size_t np, nu, nv;
    double *st[S_NSTATE];
    double *scr[X_NSCR];
    uint64_t *d_hist;      /* 3 uint64 per step */
    int64_t hist_cap;
    uint64_t *h_hist;      /* pinned mirror */
    double *d_l2;          /* sums of r3^2 of the last step (u, v, p) */
    int l2_on;             /* fold them (cfd_ctx_set_residual_l2) */
    hipEvent_t ev0, ev1;
    double last_ms_per_step;
    int64_t n_fused, n_exact;
    cfd_tiles tiles;       /* fused kernel shapes (zeros: built-in choice) */
    /* FP32 storage (precision CFD_PREC_FP32): float copies of every state buffer except
     * p and pn, and of eps and L. live32 = the float copies hold the current iterates
     * (the fused steps run on them); the FP64 buffers are brought up to date before any
     * other use (exact steps, downloads, level copies). */
    float *s32[S_NSTATE];
    float *eps32, *lp32;
    int live32, geo32_stale;
    int tiles_set;         /* chosen by the caller or the autotuner */
    /* two-stream pipeline of the fused path (unsteady problems, context stream): K2 of step
     * n on s2 overlaps K1 of step n+1 on s. ek2[n % 2] marks the end of K2(n); K1(n+2)
     * waits for it (it overwrites K2(n)'s u, v input). Every call that steps joins s2 back
     * into s before it returns, so other calls only ever see the main stream. */
    int pipe;              /* enabled (cfd_ctx_set_pipeline) */
    hipStream_t s2;
    hipEvent_t ek1, ek2[2];
    int ek2_live[2], ek2_last;
    int64_t pipe_n;
};

static void tiles_from_cache(cfd_ctx *c);
static int materialize(cfd_ctx *c);
static int to32(cfd_ctx *c);

static size_t kind_count(const cfd_ctx *c, char kind)
{
    return kind == 'u' ? c->nu : (kind == 'v' ? c->nv : c->np);
}

static int dalloc(double **p, size_t n)
{
    HIPCHK(hipMalloc((void **)p, n * sizeof(double)));
    HIPCHK(hipMemset(*p, 0, n * sizeof(double)));
    return CFD_OK;
}

static int ensure_hist(cfd_ctx *c, int64_t n)
{
    if (n <= c->hist_cap) return CFD_OK;
    if (c->d_hist) { hipFree(c->d_hist); c->d_hist = NULL; }
    if (c->h_hist) { hipHostFree(c->h_hist); c->h_hist = NULL; }
    int64_t cap = n < 64 ? 64 : n;
    HIPCHK(hipMalloc((void **)&c->d_hist, (size_t)cap * 3 * sizeof(uint64_t)));
    HIPCHK(hipHostMalloc((void **)&c->h_hist, (size_t)cap * 3 * sizeof(uint64_t), hipHostMallocDefault));
    c->hist_cap = cap;
    return CFD_OK;
}

static int ctx_alloc(cfd_ctx *c)
{
    const int im = c->P.imax, jm = c->P.jmax;
    c->np = (size_t)(im + 1) * (size_t)(jm + 1);
    c->nu = (size_t)(im + 2) * (size_t)(jm + 1);
    c->nv = (size_t)(im + 1) * (size_t)(jm + 2);
    cfd_devgeo *g = &c->g;
    g->imax = im; g->jmax = jm;
    CHK(dalloc(&g->x, im + 1)); CHK(dalloc(&g->y, jm + 1));
    CHK(dalloc(&g->xm, im + 2)); CHK(dalloc(&g->ym, jm + 2));
    CHK(dalloc(&g->aun, im + 2)); CHK(dalloc(&g->aus, im + 2));
    CHK(dalloc(&g->aue, jm + 1)); CHK(dalloc(&g->auw, jm + 1));
    CHK(dalloc(&g->avn, im + 1)); CHK(dalloc(&g->avs, im + 1));
    CHK(dalloc(&g->ave, jm + 2)); CHK(dalloc(&g->avw, jm + 2));
    CHK(dalloc(&g->eps, c->np)); CHK(dalloc(&g->lp, c->np));
    CHK(dalloc(&c->fac.rdx, im + 2)); CHK(dalloc(&c->fac.rdy, jm + 2));
    CHK(dalloc(&c->fac.rdxm, im + 2)); CHK(dalloc(&c->fac.rdym, jm + 2));
    CHK(dalloc(&c->fac.pack, (size_t)4 * (im + 2) + (size_t)4 * (jm + 2)));
    static const char kinds[S_NSTATE] = {'u', 'v', 'u', 'v', 'u', 'v', 'p', 'p', 'p', 'p', 'p', 'p', 'p', 'p'};
    for (int k = 0; k < S_NSTATE; ++k) CHK(dalloc(&c->st[k], kind_count(c, kinds[k])));
    static const char skinds[X_NSCR] = {'u', 'v', 'p', 'p', 'u', 'v', 'p', 'p', 'p', 'p', 'u', 'v', 'p'};
    for (int k = 0; k < X_NSCR; ++k) CHK(dalloc(&c->scr[k], kind_count(c, skinds[k])));
    if (c->prec == CFD_PREC_FP32) {
        for (int k = 0; k < S_NSTATE; ++k) {
            if (k == S_P || k == S_PN) continue;
            HIPCHK(hipMalloc((void **)&c->s32[k], kind_count(c, kinds[k]) * sizeof(float)));
            HIPCHK(hipMemset(c->s32[k], 0, kind_count(c, kinds[k]) * sizeof(float)));
        }
        HIPCHK(hipMalloc((void **)&c->eps32, c->np * sizeof(float)));
        HIPCHK(hipMalloc((void **)&c->lp32, c->np * sizeof(float)));
        c->geo32_stale = 1;
    }
    CHK(ensure_hist(c, 64));
    CHK(dalloc(&c->d_l2, 3));
    HIPCHK(hipEventCreate(&c->ev0));
    HIPCHK(hipEventCreate(&c->ev1));
    if (c->owns_stream) {
        HIPCHK(hipStreamCreateWithFlags(&c->s2, hipStreamNonBlocking));
        HIPCHK(hipEventCreateWithFlags(&c->ek1, hipEventDisableTiming));
        HIPCHK(hipEventCreateWithFlags(&c->ek2[0], hipEventDisableTiming));
        HIPCHK(hipEventCreateWithFlags(&c->ek2[1], hipEventDisableTiming));
        c->pipe = 1;
    }
    /* eps = 1 everywhere until a porosity field is uploaded */
    CHK(cfdk_fill(g->eps, 1.0, (long)c->np, c->s));
    return CFD_OK;
}

static void ctx_free(cfd_ctx *c)
{
    cfd_devgeo *g = &c->g;
    double *geo[] = {g->x, g->y, g->xm, g->ym, g->aun, g->aus, g->aue, g->auw, g->avn, g->avs,
                     g->ave, g->avw, g->eps, g->lp, c->fac.rdx, c->fac.rdy, c->fac.rdxm, c->fac.rdym,
                     c->fac.pack};
    for (size_t k = 0; k < sizeof geo / sizeof geo[0]; ++k) if (geo[k]) hipFree(geo[k]);
    for (int k = 0; k < S_NSTATE; ++k) if (c->st[k]) hipFree(c->st[k]);
    for (int k = 0; k < X_NSCR; ++k) if (c->scr[k]) hipFree(c->scr[k]);
    for (int k = 0; k < S_NSTATE; ++k) if (c->s32[k]) hipFree(c->s32[k]);
    if (c->eps32) hipFree(c->eps32);
    if (c->lp32) hipFree(c->lp32);
    if (c->d_hist) hipFree(c->d_hist);
    if (c->h_hist) hipHostFree(c->h_hist);
    if (c->d_l2) hipFree(c->d_l2);
    if (c->ev0) hipEventDestroy(c->ev0);
    if (c->ev1) hipEventDestroy(c->ev1);
    if (c->ek1) hipEventDestroy(c->ek1);
    for (int k = 0; k < 2; ++k) if (c->ek2[k]) hipEventDestroy(c->ek2[k]);
    if (c->s2) hipStreamDestroy(c->s2);
    if (c->owns_stream && c->s) hipStreamDestroy(c->s);
}

static int ctx_init(cfd_ctx *c, const cfd_params *params, int32_t precision, int own_stream)
{
    CHK(check_params(params));
    if (precision != CFD_PREC_FP64 && precision != CFD_PREC_FP32)
        return fail(CFD_ERR_ARG, "unknown precision %d", precision);
    c->P = *params;
    c->prec = precision;
    HIPCHK(hipGetDevice(&c->device));
    if (own_stream) {
        HIPCHK(hipStreamCreateWithFlags(&c->s, hipStreamNonBlocking));
        c->owns_stream = 1;
    } else {
        c->s = 0;   /* legacy default stream, ordered with the caller's own launches */
    }
    c->first_step = 1;
    c->have_fused = cfdk_fused_available();
    c->kernel = 1;
    tiles_from_cache(c);
    return ctx_alloc(c);
}

int cfd_ctx_create(const cfd_params *params, int32_t precision, cfd_ctx **out)
{
    if (!out) return fail(CFD_ERR_ARG, "out is NULL");
    *out = NULL;
    cfd_ctx *c = (cfd_ctx *)calloc(1, sizeof *c);
    if (!c) return fail(CFD_ERR_ARG, "out of host memory");
    int r = ctx_init(c, params, precision, 1);
    if (r != CFD_OK) { ctx_free(c); free(c); return r; }
    *out = c;
    return CFD_OK;
}

int cfd_ctx_destroy(cfd_ctx *c)
{
    if (!c) return CFD_OK;
    if (c->s || c->owns_stream) hipStreamSynchronize(c->s);
    if (c->s2) hipStreamSynchronize(c->s2);
    ctx_free(c);
    free(c);
    return CFD_OK;
}

int cfd_ctx_get_params(const cfd_ctx *c, cfd_params *out)
{
    if (!c || !out) return fail(CFD_ERR_ARG, "NULL argument");
    *out = c->P;
    return CFD_OK;
}

int cfd_ctx_set_params(cfd_ctx *c, const cfd_params *p)
{
    if (!c) return fail(CFD_ERR_STATE, "no context");
    CHK(check_params(p));
    if (p->imax != c->P.imax || p->jmax != c->P.jmax)
        return fail(CFD_ERR_ARG, "grid size cannot change (%dx%d -> %dx%d)", c->P.imax, c->P.jmax, p->imax, p->jmax);
    c->P = *p;
    return CFD_OK;
}

int cfd_ctx_set_pipeline(cfd_ctx *c, int32_t on)
{
    if (!c) return fail(CFD_ERR_STATE, "no context");
    if (on && !c->s2) return fail(CFD_ERR_UNSUPPORTED, "the pipeline needs a context stream (not the legacy default stream)");
    c->pipe = on ? 1 : 0;
    return CFD_OK;
}

int cfd_ctx_set_kernel(cfd_ctx *c, int32_t kernel)
{
    if (!c) return fail(CFD_ERR_STATE, "no context");
    if (kernel < 0 || kernel > 1) return fail(CFD_ERR_ARG, "unknown kernel %d", kernel);
    c->kernel = kernel;
    return CFD_OK;
}

static int upload_geometry(cfd_ctx *c, const double *x, const double *y)
{
    const int im = c->P.imax, jm = c->P.jmax;
    double *h = (double *)calloc((size_t)(12 * (im + 3) + 12 * (jm + 3)), sizeof(double));
    if (!h) return fail(CFD_ERR_ARG, "out of host memory");
    double *xm = h, *aun = xm + im + 3, *aus = aun + im + 3, *avn = aus + im + 3, *avs = avn + im + 3;
    double *ym = avs + im + 3, *aue = ym + jm + 3, *auw = aue + jm + 3, *ave = auw + jm + 3, *avw = ave + jm + 3;
    int r = cfd_geometry_derive(im, jm, x, y, xm, ym, aun, aus, aue, auw, avn, avs, ave, avw);
    if (r == CFD_OK) {
        struct { double *d; const double *h; size_t n; } cp[] = {
            {c->g.x, x, im + 1}, {c->g.y, y, jm + 1}, {c->g.xm, xm, im + 2}, {c->g.ym, ym, jm + 2},
            {c->g.aun, aun, im + 2}, {c->g.aus, aus, im + 2}, {c->g.aue, aue, jm + 1}, {c->g.auw, auw, jm + 1},
            {c->g.avn, avn, im + 1}, {c->g.avs, avs, im + 1}, {c->g.ave, ave, jm + 2}, {c->g.avw, avw, jm + 2}};
        for (size_t k = 0; k < sizeof cp / sizeof cp[0] && r == CFD_OK; ++k) {
            hipError_t e = hipMemcpy(cp[k].d, cp[k].h, cp[k].n * sizeof(double), hipMemcpyHostToDevice);
            if (e != hipSuccess) r = fail(CFD_ERR_HIP, "geometry upload: %s", hipGetErrorString(e));
        }
    }
    free(h);
    if (r == CFD_OK) r = cfdk_factors(&c->g, &c->fac, c->s);
    return r;
}

int cfd_ctx_set_mesh(cfd_ctx *c, const double *x, const double *y)
{
    if (!c) return fail(CFD_ERR_STATE, "no context");
    if (!x || !y) return fail(CFD_ERR_ARG, "x or y is NULL");
    for (int i = 2; i <= c->P.imax; ++i)
        if (!(x[i] > x[i - 1])) return fail(CFD_ERR_ARG, "x must increase strictly (x[%d] <= x[%d])", i, i - 1);
    for (int j = 2; j <= c->P.jmax; ++j)
        if (!(y[j] > y[j - 1])) return fail(CFD_ERR_ARG, "y must increase strictly (y[%d] <= y[%d])", j, j - 1);
    CHK(upload_geometry(c, x, y));
    c->have_mesh = 1;
    return CFD_OK;
}

static double *field_ptr(cfd_ctx *c, int32_t f, char *kind)
{
    switch (f) {
    case CFD_F_UM: *kind = 'u'; return c->st[S_UM];
    case CFD_F_VM: *kind = 'v'; return c->st[S_VM];
    case CFD_F_UM_TAU: *kind = 'u'; return c->st[S_UMT];
    case CFD_F_VM_TAU: *kind = 'v'; return c->st[S_VMT];
    case CFD_F_P: *kind = 'p'; return c->st[S_P];
    case CFD_F_T: *kind = 'p'; return c->st[S_T];
    case CFD_F_T_TAU: *kind = 'p'; return c->st[S_TT];
    case CFD_F_C: *kind = 'p'; return c->st[S_C];
    case CFD_F_C_TAU: *kind = 'p'; return c->st[S_CT];
    case CFD_F_EPS: *kind = 'p'; return c->g.eps;
    case CFD_F_LP: *kind = 'p'; return c->g.lp;
    case CFD_F_PN: *kind = 'p'; return c->uv_reads_pn ? c->st[S_PN] : c->st[S_P];
    default: return NULL;
    }
}

/* Fortran shape (n, m) of a layout kind; C row stride ld. */
static void kind_dims(const cfd_ctx *c, char kind, int *n, int *m, int *ld)
{
    const int im = c->P.imax, jm = c->P.jmax;
    if (kind == 'u') { *n = im + 1; *m = jm; *ld = jm + 1; }
    else if (kind == 'v') { *n = im; *m = jm + 1; *ld = jm + 2; }
    else { *n = im; *m = jm; *ld = jm + 1; }
}

int cfd_ctx_upload(cfd_ctx *c, int32_t field, const double *host, int32_t layout)
{
    if (!c) return fail(CFD_ERR_STATE, "no context");
    if (!host) return fail(CFD_ERR_ARG, "host pointer is NULL");
    if (field == CFD_F_PN) return fail(CFD_ERR_ARG, "CFD_F_PN is download-only");
    CHK(materialize(c));
    char kind;
    double *d = field_ptr(c, field, &kind);
    if (!d) return fail(CFD_ERR_ARG, "unknown field %d", field);
    if (field == CFD_F_EPS || field == CFD_F_LP) c->geo32_stale = 1;
    int n, m, ld;
    kind_dims(c, kind, &n, &m, &ld);
    if (layout == CFD_LAYOUT_C) {
        HIPCHK(hipMemcpyAsync(d, host, kind_count(c, kind) * sizeof(double), hipMemcpyHostToDevice, c->s));
    } else if (layout == CFD_LAYOUT_FORTRAN) {
        double *tmp = NULL;
        HIPCHK(hipMalloc((void **)&tmp, (size_t)n * m * sizeof(double)));
        hipError_t e = hipMemcpyAsync(tmp, host, (size_t)n * m * sizeof(double), hipMemcpyHostToDevice, c->s);
        int r = e == hipSuccess ? cfdk_from_fortran(d, tmp, n, m, ld, c->s) : CFD_ERR_HIP;
        hipStreamSynchronize(c->s);
        hipFree(tmp);
        if (r != CFD_OK) return fail(CFD_ERR_HIP, "Fortran-layout upload failed");
    } else {
        return fail(CFD_ERR_ARG, "unknown layout %d", layout);
    }
    HIPCHK(hipStreamSynchronize(c->s));
    if (field == CFD_F_P) c->uv_reads_pn = 0;
    if (field != CFD_F_EPS && field != CFD_F_LP) { c->bc_clean = 0; c->first_step = 1; }
    return CFD_OK;
}


int cfd_ctx_download(cfd_ctx *c, int32_t field, double *host, int32_t layout)
{
    if (!c) return fail(CFD_ERR_STATE, "no context");
    if (!host) return fail(CFD_ERR_ARG, "host pointer is NULL");
    CHK(materialize(c));
    char kind;
    double *d = field_ptr(c, field, &kind);
    if (!d) return fail(CFD_ERR_ARG, "unknown field %d", field);
    int n, m, ld;
    kind_dims(c, kind, &n, &m, &ld);
    if (layout == CFD_LAYOUT_C) {
        HIPCHK(hipMemcpyAsync(host, d, kind_count(c, kind) * sizeof(double), hipMemcpyDeviceToHost, c->s));
        HIPCHK(hipStreamSynchronize(c->s));
    } else if (layout == CFD_LAYOUT_FORTRAN) {
        double *tmp = NULL;
        HIPCHK(hipMalloc((void **)&tmp, (size_t)n * m * sizeof(double)));
        int r = cfdk_to_fortran(tmp, d, n, m, ld, c->s);
        hipError_t e = r == CFD_OK ? hipMemcpyAsync(host, tmp, (size_t)n * m * sizeof(double), hipMemcpyDeviceToHost, c->s)
                                   : hipErrorUnknown;
        hipStreamSynchronize(c->s);
        hipFree(tmp);
        if (e != hipSuccess) return fail(CFD_ERR_HIP, "Fortran-layout download failed");
    } else {
        return fail(CFD_ERR_ARG, "unknown layout %d", layout);
    }
    return CFD_OK;
}

int cfd_ctx_device_ptr(cfd_ctx *c, int32_t field, void **out)
{
    if (!c || !out) return fail(CFD_ERR_ARG, "NULL argument");
    CHK(materialize(c));
    char kind;
    double *d = field_ptr(c, field, &kind);
    if (!d) return fail(CFD_ERR_ARG, "unknown field %d", field);
    *out = d;
    return CFD_OK;
}

int cfd_ctx_sync_levels(cfd_ctx *c)
{
    if (!c) return fail(CFD_ERR_STATE, "no context");
    CHK(materialize(c));
    CHK(cfdk_copy(c->st[S_UMT], c->st[S_UM], (long)c->nu, c->s));
    CHK(cfdk_copy(c->st[S_VMT], c->st[S_VM], (long)c->nv, c->s));
    CHK(cfdk_copy(c->st[S_TT], c->st[S_T], (long)c->np, c->s));
    CHK(cfdk_copy(c->st[S_CT], c->st[S_C], (long)c->np, c->s));
    HIPCHK(hipStreamSynchronize(c->s));
    c->uv_reads_pn = 0;
    c->bc_clean = 0;
    c->first_step = 1;
    return CFD_OK;
}

/* ------------------------------------------------------------------------------------ */
/* exact step: the reference call structure                                             */
typedef struct {
    const cfd_devgeo *g;
    const cfd_params *P;
    hipStream_t s;
    double *l2;            /* sums of squares of the residuals (u, v, p), or NULL */
} xenv;

static int x_residual(const xenv *e, int region, const double *res, uint64_t *slot)
{
    const int im = e->g->imax, jm = e->g->jmax;
    if (e->P->compat == CFD_COMPAT_CUDA) {
        /* max_reduce(dev_res, n, m) over the first n*m linear entries (max_reduce.cu:14-41),
         * called with (imax-1, jmax-2) for u (solve_U.cu:58), (imax-1, jmax) for v
         * (solve_V.cu:52) and (imax, jmax) for p (solve_P.cu:50). */
        long n = region == 0 ? (long)(im - 1) * (jm - 2) : (region == 1 ? (long)(im - 1) * jm : (long)im * jm);
        if (e->l2) {
            static const int lo_i[3] = {3, 2, 2}, lo_j[3] = {2, 3, 2};
            CHK(cfdk_sumsq(res, region == 1 ? jm + 2 : jm + 1, lo_i[region], im - 1, lo_j[region], jm - 1,
                           e->l2 + region, e->s));
        }
        return cfdk_maxabs_linear(res, n, slot, e->s);
    }
    /* the update regions of equations.f90:298,600,662; the sums of squares (no counterpart in
     * the reference) run over the same entries */
    static const int lo_i[3] = {3, 2, 2}, lo_j[3] = {2, 3, 2};
    const int ld = region == 1 ? jm + 2 : jm + 1;
    if (e->l2) CHK(cfdk_sumsq(res, ld, lo_i[region], im - 1, lo_j[region], jm - 1, e->l2 + region, e->s));
    return cfdk_maxabs(res, ld, lo_i[region], im - 1, lo_j[region], jm - 1, slot, e->s);
}

/* equations.f90:244-301; T (may be NULL) feeds the mu(T) model only. */
static int x_solve_u(const xenv *e, const double *um, double *um_tau, double *vm_tau, double *um_n_tau,
                     const double *p, const double *T, double *ui, double *ru, double *res, uint64_t *slot)
{
    const cfd_devgeo *g = e->g; const cfd_params *P = e->P;
    CHK(cfdk_resu(g, P, 0, um_tau, vm_tau, p, T, ru, e->s));
    CHK(cfdk_stage(g, P, 0, 1, um, um_tau, ru, NULL, ui, res, e->s));
    CHK(cfdk_bc_uv(g, P, ui, vm_tau, e->s));
    CHK(cfdk_resu(g, P, 0, ui, vm_tau, p, T, ru, e->s));
    CHK(cfdk_stage(g, P, 0, 2, um, um_tau, ru, ui, ui, res, e->s));
    CHK(cfdk_bc_uv(g, P, ui, vm_tau, e->s));
    CHK(cfdk_resu(g, P, 0, ui, vm_tau, p, T, ru, e->s));
    CHK(cfdk_stage(g, P, 0, 3, um, um_tau, ru, ui, um_n_tau, res, e->s));
    CHK(cfdk_bc_uv(g, P, um_n_tau, vm_tau, e->s));
    return slot ? x_residual(e, 0, res, slot) : CFD_OK;
}

/* equations.f90:546-603 */
static int x_solve_v(const xenv *e, const double *vm, double *um_tau, double *vm_tau, double *vm_n_tau,
                     const double *p, const double *T, double *vi, double *rv, double *res, uint64_t *slot)
{
    const cfd_devgeo *g = e->g; const cfd_params *P = e->P;
    CHK(cfdk_resv(g, P, 0, um_tau, vm_tau, p, T, rv, e->s));
    CHK(cfdk_stage(g, P, 1, 1, vm, vm_tau, rv, NULL, vi, res, e->s));
    CHK(cfdk_bc_uv(g, P, um_tau, vi, e->s));
    CHK(cfdk_resv(g, P, 0, um_tau, vi, p, T, rv, e->s));
    CHK(cfdk_stage(g, P, 1, 2, vm, vm_tau, rv, vi, vi, res, e->s));
    CHK(cfdk_bc_uv(g, P, um_tau, vi, e->s));
    CHK(cfdk_resv(g, P, 0, um_tau, vi, p, T, rv, e->s));
    CHK(cfdk_stage(g, P, 1, 3, vm, vm_tau, rv, vi, vm_n_tau, res, e->s));
    CHK(cfdk_bc_uv(g, P, um_tau, vm_n_tau, e->s));
    return slot ? x_residual(e, 1, res, slot) : CFD_OK;
}

/* equations.f90:607-665 */
static int x_solve_p(const xenv *e, const double *p, const double *um_n, const double *vm_n, double *pn,
                     double *rp, double *pi, double *res, uint64_t *slot)
{
    const cfd_devgeo *g = e->g; const cfd_params *P = e->P;
    CHK(cfdk_pstage(g, P, 1, p, um_n, vm_n, rp, pi, pn, res, e->s));
    CHK(cfdk_bc_s(g, P, 0, pi, e->s));
    CHK(cfdk_pstage(g, P, 2, p, um_n, vm_n, rp, pi, pn, res, e->s));
    CHK(cfdk_bc_s(g, P, 0, pi, e->s));
    CHK(cfdk_pstage(g, P, 3, p, um_n, vm_n, rp, pi, pn, res, e->s));
    CHK(cfdk_bc_s(g, P, 0, pn, e->s));
    return slot ? x_residual(e, 2, res, slot) : CFD_OK;
}

/* equations.f90:669-719 (species = 0) and 769-819 (species = 1) */
static int x_solve_s(const xenv *e, int species, const double *um_n, const double *vm_n, const double *z,
                     double *z_n_tau, const double *z_tau, double *zi, double *rz)
{
    const cfd_devgeo *g = e->g; const cfd_params *P = e->P;
    const int kind = species ? 2 : 1;
    CHK(cfdk_ressc(g, P, species, um_n, vm_n, z_tau, z, rz, e->s));
    CHK(cfdk_stage(g, P, 2, 1, z, z_tau, rz, NULL, zi, NULL, e->s));
    CHK(cfdk_bc_s(g, P, kind, zi, e->s));
    CHK(cfdk_ressc(g, P, species, um_n, vm_n, zi, z, rz, e->s));
    CHK(cfdk_stage(g, P, 2, 2, z, z_tau, rz, zi, zi, NULL, e->s));
    CHK(cfdk_bc_s(g, P, kind, zi, e->s));
    CHK(cfdk_ressc(g, P, species, um_n, vm_n, zi, z, rz, e->s));
    CHK(cfdk_stage(g, P, 2, 3, z, z_tau, rz, zi, z_n_tau, NULL, e->s));
    CHK(cfdk_bc_s(g, P, kind, z_n_tau, e->s));
    return CFD_OK;
}

static void swap_ptr(double **a, double **b) { double *t = *a; *a = *b; *b = t; }

static void swap_f(float **a, float **b) { float *t = *a; *a = *b; *b = t; }

static void tau_swap(cfd_ctx *c)
{
    swap_f(&c->s32[S_UMT], &c->s32[S_UMN]);
    swap_f(&c->s32[S_VMT], &c->s32[S_VMN]);
    swap_f(&c->s32[S_TT], &c->s32[S_TN]);
    swap_f(&c->s32[S_CT], &c->s32[S_CN]);
    swap_ptr(&c->st[S_UMT], &c->st[S_UMN]);
    swap_ptr(&c->st[S_VMT], &c->st[S_VMN]);
    swap_ptr(&c->st[S_P], &c->st[S_PN]);
    swap_ptr(&c->st[S_TT], &c->st[S_TN]);
    swap_ptr(&c->st[S_CT], &c->st[S_CN]);
    c->uv_reads_pn = 0;
    c->pending_in_tau = 1;
}

/* One pseudo-step (main.f90:122-142) through the per-operator kernels; residuals into slot[0..2].
 * `next` (or NULL): the next step's residual slots, cleared here too. */
static int step_exact(cfd_ctx *c, uint64_t *slot, uint64_t *next)
{
    xenv e = {&c->g, &c->P, c->s, c->l2_on ? c->d_l2 : NULL};
    const int st = c->P.steady;
    double **S = c->st, **X = c->scr;
    const double *um = st ? S[S_UMT] : S[S_UM];
    const double *vm = st ? S[S_VMT] : S[S_VM];
    const double *T = st ? S[S_TT] : S[S_T];
    const double *C = st ? S[S_CT] : S[S_C];
    const double *p_uv = c->uv_reads_pn ? S[S_PN] : S[S_P];
    if (c->P.bc == CFD_BC_REFERENCE && !c->uv_reads_pn) {
        /* The reference's vm_n_tau keeps its previous values where the interior update does
         * not write; the corner BC (imax, jmax+1) reads two of them. With ping-pong buffers
         * the previous step's output is vm_tau, so those two cells are seeded from it (zero
         * before the first step, as a freshly allocated array). After an exit without the
         * tau update the buffer already holds the previous output. */
        const int im = c->P.imax, jm = c->P.jmax;
        const size_t k = (size_t)im * (jm + 2) + (jm - 2);
        if (c->first_step) {
            CHK(cfdk_fill(S[S_VMN] + k, 0.0, 2, c->s));
        } else {
            CHK(cfdk_copy(S[S_VMN] + k, S[S_VMT] + k, 2, c->s));
        }
    }
    CHK(cfdk_zero_slots(slot, 3, e.l2, c->s));
    if (next) CHK(cfdk_zero_slots(next, 3, NULL, c->s));
    CHK(x_solve_u(&e, um, S[S_UMT], S[S_VMT], S[S_UMN], p_uv, T, X[X_UI], X[X_RU], X[X_RESU], slot + 0));
    CHK(x_solve_v(&e, vm, S[S_UMT], S[S_VMT], S[S_VMN], p_uv, T, X[X_VI], X[X_RV], X[X_RESV], slot + 1));
    CHK(x_solve_p(&e, S[S_P], S[S_UMN], S[S_VMN], S[S_PN], X[X_RP], X[X_PI], X[X_RESP], slot + 2));
    CHK(x_solve_s(&e, 0, S[S_UMN], S[S_VMN], T, S[S_TN], S[S_TT], X[X_ZI], X[X_RZ]));
    CHK(x_solve_s(&e, 1, S[S_UMN], S[S_VMN], C, S[S_CN], S[S_CT], X[X_CI], X[X_RC]));
    c->bc_clean = 1;
    c->first_step = 0;
    c->n_exact++;
    return CFD_OK;
}

/* Joins the pipeline's second stream into the main stream: work queued on s after this
 * sees every K2 launched so far. */
static int pipe_join(cfd_ctx *c)
{
    if (c->ek2_live[0] || c->ek2_live[1]) {
        HIPCHK(hipStreamWaitEvent(c->s, c->ek2[c->ek2_last], 0));
        c->ek2_live[0] = c->ek2_live[1] = 0;
    }
    return CFD_OK;
}

/* Fused step: boundary-consistent iterates in, outputs with boundaries written last.
 * With `ev` (four events), they bracket the step's three launches. */
static int step_fused_ev(cfd_ctx *c, uint64_t *slot, uint64_t *next, int pre, const hipEvent_t *ev)
{
    const int timed = ev != NULL;
    const int st = c->P.steady;
    double **S = c->st;
    cfd_fused_args a;
    memset(&a, 0, sizeof a);
    void *F[S_NSTATE];
    if (c->prec == CFD_PREC_FP32) {
        CHK(to32(c));
        for (int k = 0; k < S_NSTATE; ++k) F[k] = c->s32[k];
        a.epsr = c->eps32; a.lpr = c->lp32;
    } else {
        for (int k = 0; k < S_NSTATE; ++k) F[k] = S[k];
        a.epsr = c->g.eps; a.lpr = c->g.lp;
    }
    a.prec = c->prec;
    a.um = st ? F[S_UMT] : F[S_UM];
    a.vm = st ? F[S_VMT] : F[S_VM];
    a.T = st ? F[S_TT] : F[S_T];
    a.C = st ? F[S_CT] : F[S_C];
    a.um_tau = F[S_UMT]; a.vm_tau = F[S_VMT];
    a.p = c->uv_reads_pn ? S[S_PN] : S[S_P];
    a.T_tau = F[S_TT]; a.C_tau = F[S_CT];
    a.um_n = F[S_UMN]; a.vm_n = F[S_VMN]; a.pn = S[S_PN]; a.T_n = F[S_TN]; a.C_n = F[S_CN];
    a.res_max = slot;
    a.res_l2 = c->l2_on ? c->d_l2 : NULL;
    a.hb1 = c->tiles.rows_uvp; a.hb2 = c->tiles.rows_tc;
    a.occ1 = c->tiles.waves_uvp; a.occ2 = c->tiles.waves_tc;
    a.lone = 1;
    if (timed) for (int k = 0; k < 4; ++k) a.ev[k] = ev[k];
    /* unsteady problems only: in steady mode K1 reads the T iterate that K2 writes */
    const int two = c->pipe && c->s2 && !timed && !st;
    if (two) {
        const int q = (int)(c->pipe_n & 1);
        a.s2 = c->s2; a.ev_k1 = c->ek1;
        a.ev_k2_wait = c->ek2_live[q] ? c->ek2[q] : NULL;
        a.ev_k2_rec = c->ek2[q];
    } else {
        CHK(pipe_join(c));
    }
    /* Residual slots: with `pre` the caller (or the step before) has cleared `slot`, and the
     * u/v/p kernel clears `next` (if any) for the step after this one: no separate launch.
     * Otherwise (the L2 sums, timed steps, callers without `pre`) a clearing launch precedes
     * the step, and clears `next` too. */
    if (pre && !a.res_l2 && !timed) {
        a.res_next = next;
    } else {
        CHK(cfdk_zero_slots(slot, 3, a.res_l2, c->s));
        if (next) CHK(cfdk_zero_slots(next, 3, NULL, c->s));
    }
    CHK(cfdk_fused_step(&c->g, &c->fac, &c->P, &a, c->kernel, c->s));
    if (two) {
        const int q = (int)(c->pipe_n & 1);
        c->ek2_live[q] = 1; c->ek2_last = q;
        c->pipe_n++;
    }
    c->bc_clean = 1;          /* cfdk_fused_step writes the boundary cells last */
    c->first_step = 0;
    c->n_fused++;
    return CFD_OK;
}

static int step_fused(cfd_ctx *c, uint64_t *slot, uint64_t *next, int pre)
{
    return step_fused_ev(c, slot, next, pre, NULL);
}

/* Float copies of the iterates the fused FP32 step reads (the physical level, the tau
 * level), and of the porosity fields when they changed. */
static int to32(cfd_ctx *c)
{
    static const int idx[] = {S_UM, S_VM, S_UMT, S_VMT, S_T, S_TT, S_C, S_CT};
    if (!c->live32) {
        for (size_t q = 0; q < sizeof idx / sizeof idx[0]; ++q) {
            const int k = idx[q];
            const size_t n = (k == S_UM || k == S_UMT) ? c->nu : ((k == S_VM || k == S_VMT) ? c->nv : c->np);
            CHK(cfdk_d2f(c->s32[k], c->st[k], (long)n, c->s));
        }
        c->live32 = 1;
    }
    if (c->geo32_stale) {
        CHK(cfdk_d2f(c->eps32, c->g.eps, (long)c->np, c->s));
        CHK(cfdk_d2f(c->lp32, c->g.lp, (long)c->np, c->s));
        c->geo32_stale = 0;
    }
    return CFD_OK;
}

/* Brings the FP64 buffers up to date (they are the only copy in FP64 contexts): after
 * fused FP32 steps the float buffers hold the state. */
static int materialize(cfd_ctx *c)
{
    if (!c->live32) return CFD_OK;
    for (int k = 0; k < S_NSTATE; ++k) {
        if (k == S_P || k == S_PN) continue;
        const size_t n = (k == S_UM || k == S_UMT || k == S_UMN) ? c->nu
                       : ((k == S_VM || k == S_VMT || k == S_VMN) ? c->nv : c->np);
        CHK(cfdk_f2d(c->st[k], c->s32[k], (long)n, c->s));
    }
    HIPCHK(hipStreamSynchronize(c->s));
    c->live32 = 0;
    return CFD_OK;
}

/* The fused kernels cover the reference's Fortran semantics without bulk viscosity
 * (b_art = 0 in every shipped input, OpenMp_Multicore/input/iterations.dat:8) on iterates
 * that already satisfy the boundary conditions; anything else takes the exact path. */
static int fused_ok(const cfd_ctx *c)
{
    return c->kernel == 1 && c->have_fused && c->bc_clean && !c->uv_reads_pn
        && c->P.b_art == 0.0 && c->P.compat == CFD_COMPAT_FORTRAN;
}

/* One pseudo-step with residuals into slot[0..2]. pre: the caller has cleared `slot`;
 * next (with pre): the slots of the step after this one, which this step clears. Without
 * pre, the step clears `slot` itself. */
static int one_step(cfd_ctx *c, uint64_t *slot, uint64_t *next, int pre)
{
    if (fused_ok(c))
        return step_fused(c, slot, next, pre);
    CHK(pipe_join(c));
    CHK(materialize(c));
    return step_exact(c, slot, next);
}

static double u64_to_res(uint64_t b)
{
    double d;
    memcpy(&d, &b, sizeof d);
    return d;
}

static int check_ready(cfd_ctx *c)
{
    if (!c) return fail(CFD_ERR_STATE, "no context");
    if (!c->have_mesh) return fail(CFD_ERR_STATE, "no mesh: call cfd_ctx_set_mesh first");
    return CFD_OK;
}

int cfd_pseudo_steps(cfd_ctx *c, int32_t nsteps, double res[3], double *history)
{
    CHK(check_ready(c));
    if (nsteps < 0) return fail(CFD_ERR_ARG, "nsteps < 0");
    if (nsteps == 0) return CFD_OK;
    if (history && nsteps > INT32_MAX / 3) return fail(CFD_ERR_ARG, "nsteps too large for a residual history");
    /* with a history every step has its own slots; otherwise two sets alternate, each step
     * clearing the other set for the step after it */
    const int64_t nslots = history ? nsteps : 2;
    CHK(ensure_hist(c, nslots));
    CHK(cfdk_zero_slots(c->d_hist, (int)(3 * nslots), NULL, c->s));
    HIPCHK(hipEventRecord(c->ev0, c->s));
    for (int32_t k = 0; k < nsteps; ++k) {
        uint64_t *slot = c->d_hist + 3 * (history ? k : (k & 1));
        uint64_t *next = history ? NULL : c->d_hist + 3 * ((k + 1) & 1);
        CHK(one_step(c, slot, next, 1));
        tau_swap(c);
    }
    CHK(pipe_join(c));
    HIPCHK(hipEventRecord(c->ev1, c->s));
    HIPCHK(hipMemcpyAsync(c->h_hist, c->d_hist, (size_t)nslots * 3 * sizeof(uint64_t), hipMemcpyDeviceToHost, c->s));
    HIPCHK(hipStreamSynchronize(c->s));
    float ms = 0.f;
    HIPCHK(hipEventElapsedTime(&ms, c->ev0, c->ev1));
    c->last_ms_per_step = (double)ms / nsteps;
    const uint64_t *last = c->h_hist + 3 * (history ? nslots - 1 : ((nsteps - 1) & 1));
    if (res) for (int q = 0; q < 3; ++q) res[q] = u64_to_res(last[q]);
    if (history) for (int64_t k = 0; k < 3 * nslots; ++k) history[k] = u64_to_res(c->h_hist[k]);
    for (int q = 0; q < 3; ++q)
        if (!isfinite(u64_to_res(last[q]))) return fail(CFD_ERR_NONFINITE, "residual %d is not finite", q);
    return CFD_OK;
}

int cfd_run_to_tol(cfd_ctx *c, double tol, int64_t itc_max, int32_t check_every,
                   double diverge_factor, int64_t *iters, double res[3])
{
    CHK(check_ready(c));
    if (check_every < 1) check_every = 1;
    CHK(ensure_hist(c, check_every));
    int64_t done = 0;
    double best = INFINITY, r[3] = {0, 0, 0};
    HIPCHK(hipEventRecord(c->ev0, c->s));
    while (done < itc_max) {
        int32_t n = (int32_t)((itc_max - done) < check_every ? (itc_max - done) : check_every);
        CHK(cfdk_zero_slots(c->d_hist, 3 * n, NULL, c->s));
        for (int32_t k = 0; k < n; ++k) {
            CHK(one_step(c, c->d_hist + 3 * k, NULL, 1));
            tau_swap(c);
        }
        CHK(pipe_join(c));
        HIPCHK(hipMemcpyAsync(c->h_hist, c->d_hist, (size_t)n * 3 * sizeof(uint64_t), hipMemcpyDeviceToHost, c->s));
        HIPCHK(hipStreamSynchronize(c->s));
        for (int32_t k = 0; k < n; ++k) {
            for (int q = 0; q < 3; ++q) r[q] = u64_to_res(c->h_hist[3 * k + q]);
            const double err = fmax(fmax(r[0], r[1]), r[2]);
            const int64_t itc = done + k + 1;
            if (!isfinite(r[0]) || !isfinite(r[1]) || !isfinite(r[2])) {
                if (iters) *iters = itc;
                if (res) memcpy(res, r, sizeof r);
                return fail(CFD_ERR_NONFINITE, "non-finite residual at pseudo-step %lld", (long long)itc);
            }
            if (itc != 1 && err < tol) {
                done += n;
                HIPCHK(hipEventRecord(c->ev1, c->s));
                HIPCHK(hipEventSynchronize(c->ev1));
                float ms = 0.f;
                hipEventElapsedTime(&ms, c->ev0, c->ev1);
                c->last_ms_per_step = (double)ms / (double)done;
                if (iters) *iters = itc;
                if (res) memcpy(res, r, sizeof r);
                return CFD_OK;
            }
            if (err < best) best = err;
            if (diverge_factor > 0 && err > diverge_factor * best) {
                if (iters) *iters = itc;
                if (res) memcpy(res, r, sizeof r);
                return fail(CFD_ERR_DIVERGED, "residual %.3e exceeds %.1f x best %.3e at step %lld",
                            err, diverge_factor, best, (long long)itc);
            }
        }
        done += n;
    }
    HIPCHK(hipEventRecord(c->ev1, c->s));
    HIPCHK(hipEventSynchronize(c->ev1));
    float ms = 0.f;
    hipEventElapsedTime(&ms, c->ev0, c->ev1);
    c->last_ms_per_step = done ? (double)ms / (double)done : 0.0;
    if (iters) *iters = done;
    if (res) memcpy(res, r, sizeof r);
    return fail(CFD_ERR_NOCONV, "tolerance %.3e not reached in %lld pseudo-steps (last %.3e)", tol,
                (long long)done, fmax(fmax(r[0], r[1]), r[2]));
}

/* main.f90:116-157, one pass of the physical-time loop body. */
int cfd_physical_step(cfd_ctx *c, double eps, int64_t itc_max, int32_t first, int64_t *iters, double res[3])
{
    CHK(check_ready(c));
    CHK(ensure_hist(c, 1));
    int64_t itc = first ? 1 : 0, taken = 0;
    double r[3] = {0, 0, 0};
    while (itc < itc_max) {
        CHK(one_step(c, c->d_hist, NULL, 0));
        ++taken;
        CHK(pipe_join(c));
        HIPCHK(hipMemcpyAsync(c->h_hist, c->d_hist, 3 * sizeof(uint64_t), hipMemcpyDeviceToHost, c->s));
        HIPCHK(hipStreamSynchronize(c->s));
        for (int q = 0; q < 3; ++q) r[q] = u64_to_res(c->h_hist[q]);
        const double err = fmax(fmax(r[0], r[1]), r[2]);
        if (!isfinite(r[0]) || !isfinite(r[1]) || !isfinite(r[2])) {
            if (iters) *iters = taken;
            return fail(CFD_ERR_NONFINITE, "non-finite residual in physical step");
        }
        if (itc != 1 && err < eps) {
            /* exit without the tau update: U/V now read pn, P reads the old p */
            CHK(materialize(c));
            c->uv_reads_pn = 1;
            double **S = c->st;
            CHK(cfdk_copy(S[S_UM], S[S_UMN], (long)c->nu, c->s));
            CHK(cfdk_copy(S[S_VM], S[S_VMN], (long)c->nv, c->s));
            CHK(cfdk_copy(S[S_T], S[S_TN], (long)c->np, c->s));
            CHK(cfdk_copy(S[S_C], S[S_CN], (long)c->np, c->s));
            HIPCHK(hipStreamSynchronize(c->s));
            if (iters) *iters = taken;
            if (res) memcpy(res, r, sizeof r);
            return CFD_OK;
        }
        tau_swap(c);
        ++itc;
    }
    CHK(materialize(c));
    double **S = c->st;
    CHK(cfdk_copy(S[S_UM], S[S_UMT], (long)c->nu, c->s));
    CHK(cfdk_copy(S[S_VM], S[S_VMT], (long)c->nv, c->s));
    CHK(cfdk_copy(S[S_T], S[S_TT], (long)c->np, c->s));
    CHK(cfdk_copy(S[S_C], S[S_CT], (long)c->np, c->s));
    HIPCHK(hipStreamSynchronize(c->s));
    if (iters) *iters = taken;
    if (res) memcpy(res, r, sizeof r);
    return CFD_OK;
}

int cfd_pseudo_steps_graph(cfd_ctx *c, int32_t steps_per_graph, int32_t replays, double res[3])
{
    CHK(check_ready(c));
    if (steps_per_graph < 2 || (steps_per_graph & 1)) return fail(CFD_ERR_ARG, "steps_per_graph must be even and >= 2");
    if (!c->owns_stream) return fail(CFD_ERR_STATE, "graph capture needs a context stream (not the legacy default stream)");
    if (replays < 1) return fail(CFD_ERR_ARG, "replays must be >= 1");
    CHK(ensure_hist(c, 2));
    /* A replay repeats the captured launches with the pointers fixed at capture, so every
     * captured step must be one that the next step would repeat as it is. Steps that are
     * not run here, outside the graph: the first step after new inputs (exact path, with
     * the reference's side effects and the vm_n corner seeding), and a step that reads pn
     * after a physical step stopped without the tau update. One exact step leaves the
     * iterates boundary-consistent, reading p, with the first-step seeding done. */
    if (c->first_step || c->uv_reads_pn || !c->bc_clean) {
        CHK(one_step(c, c->d_hist, NULL, 0));
        tau_swap(c);
    }
    if (c->first_step || c->uv_reads_pn || !c->bc_clean)
        return fail(CFD_ERR_STATE, "the iterates are not in a state that a graph replay can repeat");
    const int fused = fused_ok(c);
    if (c->kernel == 1 && c->have_fused && c->P.b_art == 0.0 && c->P.compat == CFD_COMPAT_FORTRAN && !fused)
        return fail(CFD_ERR_STATE, "fused steps are not available for the graph");
    /* precision conversions happen outside the graph: it replays fused steps only */
    if (fused && c->prec == CFD_PREC_FP32) CHK(to32(c));
    hipGraph_t graph = NULL;
    hipGraphExec_t exec = NULL;
    CHK(pipe_join(c));            /* no wait inside the graph on work from before it */
    /* buffer roles and step counts before the capture: a failed capture executes nothing */
    double *st0[S_NSTATE];
    float *s320[S_NSTATE];
    memcpy(st0, c->st, sizeof st0);
    memcpy(s320, c->s32, sizeof s320);
    const int64_t nf0 = c->n_fused, ne0 = c->n_exact, pipe0 = c->pipe_n;
    const int pend0 = c->pending_in_tau;
    /* two slot sets alternate inside the graph (an even number of steps), each step clearing
     * the other set for the step after it: the last step of a replay clears set 0 for the
     * first step of the next; both are cleared here, outside the graph, for the first replay */
    CHK(cfdk_zero_slots(c->d_hist, 6, NULL, c->s));
    HIPCHK(hipStreamBeginCapture(c->s, hipStreamCaptureModeThreadLocal));
    int r = CFD_OK;
    for (int32_t k = 0; k < steps_per_graph && r == CFD_OK; ++k) {
        r = one_step(c, c->d_hist + 3 * (k & 1), c->d_hist + 3 * ((k + 1) & 1), 1);
        if (r == CFD_OK) tau_swap(c);
    }
    if (r == CFD_OK) r = pipe_join(c);    /* the second stream rejoins before the capture ends */
    c->ek2_live[0] = c->ek2_live[1] = 0;
    hipError_t e = hipStreamEndCapture(c->s, &graph);
    if (r == CFD_OK && e == hipSuccess) e = hipGraphInstantiate(&exec, graph, NULL, NULL, 0);
    if (r != CFD_OK || e != hipSuccess) {
        if (graph) hipGraphDestroy(graph);
        memcpy(c->st, st0, sizeof st0);
        memcpy(c->s32, s320, sizeof s320);
        c->n_fused = nf0; c->n_exact = ne0; c->pipe_n = pipe0; c->pending_in_tau = pend0;
        if (r != CFD_OK) return r;
        return fail(CFD_ERR_HIP, "graph capture failed: %s", hipGetErrorString(e));
    }
    /* each replay runs the captured steps once more */
    c->n_fused = nf0 + (c->n_fused - nf0) * replays;
    c->n_exact = ne0 + (c->n_exact - ne0) * replays;
    HIPCHK(hipEventRecord(c->ev0, c->s));
    for (int32_t q = 0; q < replays && e == hipSuccess; ++q) e = hipGraphLaunch(exec, c->s);
    hipEventRecord(c->ev1, c->s);
    /* the last captured step (odd) wrote set 1 */
    hipMemcpyAsync(c->h_hist, c->d_hist + 3, 3 * sizeof(uint64_t), hipMemcpyDeviceToHost, c->s);
    hipError_t e2 = hipStreamSynchronize(c->s);
    hipGraphExecDestroy(exec);
    hipGraphDestroy(graph);
    if (e != hipSuccess || e2 != hipSuccess)
        return fail(CFD_ERR_HIP, "graph replay failed: %s", hipGetErrorString(e != hipSuccess ? e : e2));
    float ms = 0.f;
    HIPCHK(hipEventElapsedTime(&ms, c->ev0, c->ev1));
    c->last_ms_per_step = (double)ms / ((double)steps_per_graph * replays);
    double rr[3];
    for (int q = 0; q < 3; ++q) rr[q] = u64_to_res(c->h_hist[q]);
    if (res) memcpy(res, rr, sizeof rr);
    for (int q = 0; q < 3; ++q)
        if (!isfinite(rr[q])) return fail(CFD_ERR_NONFINITE, "residual %d is not finite", q);
    return CFD_OK;
}

/* ---- tile shapes and the autotuner ------------------------------------------------- */
/* Tuned shapes, kept per grid class for the life of the process. */
typedef struct tile_entry {
    int32_t imax, jmax, prec, steady, mu_model, scheme, bc, buoyancy;
    cfd_tiles t;
} tile_entry;
enum { TILE_CACHE = 32 };
static tile_entry g_tiles[TILE_CACHE];
static int g_ntiles;
/* contexts on different host threads share the cache */
static pthread_mutex_t g_tiles_lock = PTHREAD_MUTEX_INITIALIZER;

static int tile_key_eq(const tile_entry *e, const cfd_ctx *c)
{
    const cfd_params *P = &c->P;
    return e->imax == P->imax && e->jmax == P->jmax && e->prec == c->prec && e->steady == P->steady
        && e->mu_model == P->mu_model && e->scheme == P->scheme && e->bc == P->bc && e->buoyancy == P->buoyancy;
}

static void tiles_from_cache(cfd_ctx *c)
{
    pthread_mutex_lock(&g_tiles_lock);
    for (int k = 0; k < g_ntiles; ++k)
        if (tile_key_eq(&g_tiles[k], c)) { c->tiles = g_tiles[k].t; c->tiles_set = 1; break; }
    pthread_mutex_unlock(&g_tiles_lock);
}

static void tiles_to_cache(const cfd_ctx *c)
{
    pthread_mutex_lock(&g_tiles_lock);
    int k = 0;
    while (k < g_ntiles && !tile_key_eq(&g_tiles[k], c)) ++k;
    if (k == g_ntiles) {
        if (g_ntiles < TILE_CACHE) ++g_ntiles;
        else k = TILE_CACHE - 1;          /* full: overwrite the last entry */
    }
    const cfd_params *P = &c->P;
    tile_entry e = {P->imax, P->jmax, c->prec, P->steady, P->mu_model, P->scheme, P->bc, P->buoyancy, c->tiles};
    g_tiles[k] = e;
    pthread_mutex_unlock(&g_tiles_lock);
}

static int tiles_valid(const cfd_ctx *c, const cfd_tiles *t)
{
    if (t->rows_uvp < 0 || t->rows_tc < 0 || t->rows_uvp > c->P.imax || t->rows_tc > c->P.imax)
        return fail(CFD_ERR_ARG, "row-block heights must lie in [0, imax], got %d, %d", t->rows_uvp, t->rows_tc);
    if (!(t->waves_uvp >= 0 && t->waves_uvp <= 3))
        return fail(CFD_ERR_ARG, "waves_uvp must be 0, 1, 2 or 3, got %d", t->waves_uvp);
    if (!(t->waves_tc == 0 || t->waves_tc == 1 || t->waves_tc == 2 || t->waves_tc == 4))
        return fail(CFD_ERR_ARG, "waves_tc must be 0, 1, 2 or 4, got %d", t->waves_tc);
    return CFD_OK;
}

int cfd_ctx_set_tiles(cfd_ctx *c, const cfd_tiles *t)
{
    if (!c) return fail(CFD_ERR_STATE, "no context");
    if (!t) return fail(CFD_ERR_ARG, "tiles is NULL");
    CHK(tiles_valid(c, t));
    c->tiles = *t;
    c->tiles_set = 1;
    return CFD_OK;
}

int cfd_ctx_get_tiles(const cfd_ctx *c, cfd_tiles *t)
{
    if (!c) return fail(CFD_ERR_STATE, "no context");
    if (!t) return fail(CFD_ERR_ARG, "tiles is NULL");
    *t = c->tiles;
    return CFD_OK;
}

/* Mean ms per fused step over `steps` steps (after one untimed step) with the given tiles. */
static int time_tiles(cfd_ctx *c, const cfd_tiles *t, int32_t steps, double *ms_out)
{
    c->tiles = *t;
    CHK(step_fused(c, c->d_hist, NULL, 0));
    tau_swap(c);
    CHK(cfdk_zero_slots(c->d_hist, 6, NULL, c->s));
    HIPCHK(hipEventRecord(c->ev0, c->s));
    for (int32_t k = 0; k < steps; ++k) {
        CHK(step_fused(c, c->d_hist + 3 * (k & 1), c->d_hist + 3 * ((k + 1) & 1), 1));
        tau_swap(c);
    }
    CHK(pipe_join(c));
    HIPCHK(hipEventRecord(c->ev1, c->s));
    HIPCHK(hipEventSynchronize(c->ev1));
    float ms = 0.f;
    HIPCHK(hipEventElapsedTime(&ms, c->ev0, c->ev1));
    *ms_out = (double)ms / steps;
    return CFD_OK;
}

/* Candidate shapes. The u/v/p kernel holds ~300 registers per lane: one wave per SIMD
 * without spills, or two with the excess spilled to scratch (then more, shorter row blocks
 * fill the 2 x 256 wave slots); variant 3 runs the u and v chains in separate waves of a
 * 512-lane workgroup, each with half the register windows. The T/C kernel runs 3 waves per
 * SIMD as compiled, 4 with spills, or (variant 2) the T and C chains in separate waves of a
 * 512-lane workgroup. Shapes are timed with the pipeline as configured. */
/* small row blocks (4-12 rows) serve grids too small to fill the CUs, whose step time is
 * the row count of one block plus the stage lags */
static const cfd_tiles CAND_UVP[] = {
    {0, 0, 1, 0}, {4, 0, 1, 0}, {8, 0, 1, 0}, {12, 0, 1, 0}, {24, 0, 1, 0}, {32, 0, 1, 0}, {48, 0, 1, 0},
    {64, 0, 1, 0}, {96, 0, 1, 0}, {128, 0, 1, 0},
    {0, 0, 2, 0}, {16, 0, 2, 0}, {24, 0, 2, 0}, {32, 0, 2, 0}, {40, 0, 2, 0}, {48, 0, 2, 0}, {64, 0, 2, 0},
    {0, 0, 3, 0}, {4, 0, 3, 0}, {8, 0, 3, 0}, {12, 0, 3, 0}, {16, 0, 3, 0}, {24, 0, 3, 0}, {32, 0, 3, 0},
    {48, 0, 3, 0}, {64, 0, 3, 0}, {96, 0, 3, 0},
};
static const cfd_tiles CAND_TC[] = {
    {0, 0, 0, 1}, {0, 4, 0, 1}, {0, 8, 0, 1}, {0, 12, 0, 1}, {0, 16, 0, 1}, {0, 24, 0, 1}, {0, 32, 0, 1},
    {0, 64, 0, 1}, {0, 4, 0, 2},
    {0, 8, 0, 4}, {0, 12, 0, 4}, {0, 16, 0, 4}, {0, 24, 0, 4}, {0, 32, 0, 4},
    {0, 0, 0, 2}, {0, 8, 0, 2}, {0, 16, 0, 2}, {0, 32, 0, 2}, {0, 64, 0, 2},
};

int cfd_ctx_autotune(cfd_ctx *c, int32_t steps, cfd_tiles *out, double *out_ms)
{
    CHK(check_ready(c));
    if (steps < 1) steps = 1;
    if (!(c->kernel == 1 && c->have_fused && c->P.b_art == 0.0 && c->P.compat == CFD_COMPAT_FORTRAN))
        return fail(CFD_ERR_UNSUPPORTED, "autotune: this context does not run the fused kernels");
    CHK(materialize(c));
    /* snapshot: every state buffer's contents, the buffer roles and the step flags */
    static const char kinds[S_NSTATE] = {'u', 'v', 'u', 'v', 'u', 'v', 'p', 'p', 'p', 'p', 'p', 'p', 'p', 'p'};
    size_t total = 0, off[S_NSTATE];
    for (int k = 0; k < S_NSTATE; ++k) { off[k] = total; total += kind_count(c, kinds[k]); }
    double *snap = NULL;
    HIPCHK(hipMalloc((void **)&snap, total * sizeof(double)));
    double *roles[S_NSTATE];
    float *roles32[S_NSTATE];
    memcpy(roles, c->st, sizeof roles);
    memcpy(roles32, c->s32, sizeof roles32);
    const int f_bc = c->bc_clean, f_uv = c->uv_reads_pn, f_first = c->first_step, f_pend = c->pending_in_tau;
    const int64_t n_f = c->n_fused, n_e = c->n_exact;
    const double last_ms = c->last_ms_per_step;
    const cfd_tiles prev = c->tiles;
    int r = CFD_OK;
    for (int k = 0; k < S_NSTATE && r == CFD_OK; ++k)
        r = cfdk_copy(snap + off[k], c->st[k], (long)kind_count(c, kinds[k]), c->s);
    /* the fused kernels need boundary-consistent iterates: one exact step if not there yet */
    if (r == CFD_OK && !fused_ok(c)) {
        r = step_exact(c, c->d_hist, NULL);
        if (r == CFD_OK) tau_swap(c);
    }
    const int rows = c->P.imax - 2;
    cfd_tiles best_t = {0, 0, 1, 1};
    double best = INFINITY, t = 0.0;
    /* the two kernels are independent: tune the u/v/p shape with the T/C default, then the
     * T/C shape with the best u/v/p one */
    for (size_t q = 0; q < sizeof CAND_UVP / sizeof CAND_UVP[0] && r == CFD_OK; ++q) {
        if (CAND_UVP[q].rows_uvp > rows) continue;
        cfd_tiles cand = CAND_UVP[q];
        cand.rows_tc = 0; cand.waves_tc = 1;
        r = time_tiles(c, &cand, steps, &t);
        if (r == CFD_OK && t < best) { best = t; best_t.rows_uvp = cand.rows_uvp; best_t.waves_uvp = cand.waves_uvp; }
    }
    for (size_t q = 0; q < sizeof CAND_TC / sizeof CAND_TC[0] && r == CFD_OK; ++q) {
        if (CAND_TC[q].rows_tc > rows) continue;
        cfd_tiles cand = best_t;
        cand.rows_tc = CAND_TC[q].rows_tc; cand.waves_tc = CAND_TC[q].waves_tc;
        r = time_tiles(c, &cand, steps, &t);
        if (r == CFD_OK && t < best) { best = t; best_t = cand; }
    }
    /* restore the state exactly as it was (the FP64 buffers hold it; float copies are
     * rebuilt by the next fused step) */
    if (pipe_join(c) != CFD_OK && r == CFD_OK) r = CFD_ERR_HIP;
    memcpy(c->st, roles, sizeof roles);
    memcpy(c->s32, roles32, sizeof roles32);
    c->live32 = 0;
    for (int k = 0; k < S_NSTATE; ++k) {
        const int rr = cfdk_copy(c->st[k], snap + off[k], (long)kind_count(c, kinds[k]), c->s);
        if (r == CFD_OK) r = rr;
    }
    hipStreamSynchronize(c->s);
    hipFree(snap);
    c->bc_clean = f_bc; c->uv_reads_pn = f_uv; c->first_step = f_first; c->pending_in_tau = f_pend;
    c->n_fused = n_f; c->n_exact = n_e;
    c->last_ms_per_step = last_ms;
    if (r != CFD_OK) { c->tiles = prev; return r; }
    c->tiles = best_t;
    c->tiles_set = 1;
    tiles_to_cache(c);
    if (out) *out = best_t;
    if (out_ms) *out_ms = best;
    return CFD_OK;
}

int cfd_pseudo_steps_timed(cfd_ctx *c, int32_t nsteps, double res[3], double ms[4])
{
    CHK(check_ready(c));
    if (nsteps < 1) return fail(CFD_ERR_ARG, "nsteps < 1");
    if (!ms) return fail(CFD_ERR_ARG, "ms is NULL");
    if (nsteps > 100000) return fail(CFD_ERR_ARG, "nsteps > 100000");
    CHK(ensure_hist(c, 1));
    /* four events per step, all read after one synchronisation at the end: the steps run
     * back to back on the stream, as in an untimed run */
    hipEvent_t *ev = (hipEvent_t *)calloc((size_t)nsteps * 4, sizeof(hipEvent_t));
    if (!ev) return fail(CFD_ERR_ARG, "out of host memory");
    int r = CFD_OK;
    hipError_t e = hipSuccess;
    for (int32_t k = 0; k < 4 * nsteps && e == hipSuccess; ++k) e = hipEventCreate(&ev[k]);
    if (e != hipSuccess) r = fail(CFD_ERR_HIP, "timing events: %s", hipGetErrorString(e));
    int32_t nf = 0;
    for (int32_t k = 0; k < nsteps && r == CFD_OK; ++k) {
        if (fused_ok(c)) {
            r = step_fused_ev(c, c->d_hist, NULL, 0, ev + 4 * nf);
            ++nf;
        } else {
            r = one_step(c, c->d_hist, NULL, 0);
        }
        if (r == CFD_OK) tau_swap(c);
    }
    if (r == CFD_OK) e = hipMemcpyAsync(c->h_hist, c->d_hist, 3 * sizeof(uint64_t), hipMemcpyDeviceToHost, c->s);
    if (r == CFD_OK && e == hipSuccess) e = hipStreamSynchronize(c->s);
    double acc[4] = {0, 0, 0, 0};
    for (int32_t k = 0; k < nf && r == CFD_OK && e == hipSuccess; ++k) {
        float t[3];
        for (int q = 0; q < 3 && e == hipSuccess; ++q) e = hipEventElapsedTime(&t[q], ev[4 * k + q], ev[4 * k + q + 1]);
        if (e == hipSuccess) { acc[0] += t[0]; acc[1] += t[1]; acc[2] += t[2]; acc[3] += (double)t[0] + t[1] + t[2]; }
    }
    if (r == CFD_OK && e != hipSuccess) r = fail(CFD_ERR_HIP, "timed steps: %s", hipGetErrorString(e));
    if (r != CFD_OK) hipStreamSynchronize(c->s);     /* no event is destroyed while pending */
    for (int32_t k = 0; k < 4 * nsteps; ++k) if (ev[k]) hipEventDestroy(ev[k]);
    free(ev);
    if (r != CFD_OK) return r;
    for (int q = 0; q < 4; ++q) ms[q] = nf ? acc[q] / nf : 0.0;
    if (res) for (int q = 0; q < 3; ++q) res[q] = u64_to_res(c->h_hist[q]);
    if (!nf) return fail(CFD_ERR_UNSUPPORTED, "no step took the fused path");
    return CFD_OK;
}

int cfd_ctx_set_residual_l2(cfd_ctx *c, int32_t on)
{
    if (!c) return fail(CFD_ERR_STATE, "no context");
    CHK(pipe_join(c));
    HIPCHK(hipMemsetAsync(c->d_l2, 0, 3 * sizeof(double), c->s));
    HIPCHK(hipStreamSynchronize(c->s));
    c->l2_on = on ? 1 : 0;
    return CFD_OK;
}

int cfd_ctx_last_residual_l2(cfd_ctx *c, double l2[3])
{
    if (!c || !l2) return fail(CFD_ERR_ARG, "NULL argument");
    if (!c->l2_on) return fail(CFD_ERR_STATE, "the L2 norms are not folded: enable them with cfd_ctx_set_residual_l2");
    CHK(pipe_join(c));
    double h[3];
    HIPCHK(hipMemcpyAsync(h, c->d_l2, sizeof h, hipMemcpyDeviceToHost, c->s));
    HIPCHK(hipStreamSynchronize(c->s));
    for (int q = 0; q < 3; ++q) l2[q] = sqrt(h[q]);
    return CFD_OK;
}

int cfd_ctx_step_counts(const cfd_ctx *c, int64_t *fused, int64_t *exact)
{
    if (!c) return fail(CFD_ERR_STATE, "no context");
    if (fused) *fused = c->n_fused;
    if (exact) *exact = c->n_exact;
    return CFD_OK;
}

int cfd_ctx_last_timing(const cfd_ctx *c, double *ms)
{
    if (!c || !ms) return fail(CFD_ERR_ARG, "NULL argument");
    *ms = c->last_ms_per_step;
    return CFD_OK;
}

/* ====================================================================================== */
/* Legacy entry points (c_cuda/functions.h:7-21,36): a default context on the default     */
/* stream, configured by cfd_legacy_init / cfd_legacy_set_mesh, field arrays owned by the */
/* caller (device pointers in the C layout).                                              */
/* ====================================================================================== */
static cfd_ctx *g_legacy = NULL;

int cfd_legacy_init(const cfd_params *params)
{
    if (g_legacy) { ctx_free(g_legacy); free(g_legacy); g_legacy = NULL; }
    cfd_ctx *c = (cfd_ctx *)calloc(1, sizeof *c);
    if (!c) return fail(CFD_ERR_ARG, "out of host memory");
    int r = ctx_init(c, params, CFD_PREC_FP64, 0);
    if (r != CFD_OK) { ctx_free(c); free(c); return r; }
    g_legacy = c;
    return CFD_OK;
}

int cfd_legacy_set_params(const cfd_params *params)
{
    if (!g_legacy) return fail(CFD_ERR_STATE, "legacy context not initialised (cfd_legacy_init)");
    return cfd_ctx_set_params(g_legacy, params);
}

int cfd_legacy_set_mesh(const double *x, const double *y, const double *eps, const double *lp)
{
    if (!g_legacy) return fail(CFD_ERR_STATE, "legacy context not initialised (cfd_legacy_init)");
    CHK(cfd_ctx_set_mesh(g_legacy, x, y));
    if (eps) CHK(cfd_ctx_upload(g_legacy, CFD_F_EPS, eps, CFD_LAYOUT_C));
    if (lp) CHK(cfd_ctx_upload(g_legacy, CFD_F_LP, lp, CFD_LAYOUT_C));
    return CFD_OK;
}

cfd_ctx *cfd_legacy_context(void) { return g_legacy; }

void alocar_globais(void)
{
    if (!g_legacy) {
        cfd_params p;
        cfd_params_default(&p, 51, 84);
        if (cfd_legacy_init(&p) != CFD_OK) return;
    }
}

void desalocar_globais(void)
{
    if (g_legacy) { hipDeviceSynchronize(); ctx_free(g_legacy); free(g_legacy); g_legacy = NULL; }
}

#define LEGACY_CTX(ret)                                                                  \
    cfd_ctx *c = g_legacy;                                                               \
    if (!c) { fail(CFD_ERR_STATE, "legacy context not initialised (cfd_legacy_init)"); return ret; } \
    xenv e = {&c->g, &c->P, c->s};                                                       \
    (void)e;

static double legacy_read(cfd_ctx *c, uint64_t *slot)
{
    uint64_t b = 0;
    if (hipMemcpy(&b, slot, sizeof b, hipMemcpyDeviceToHost) != hipSuccess) {
        fail(CFD_ERR_HIP, "residual read-back failed");
        return NAN;
    }
    (void)c;
    return u64_to_res(b);
}

void RESU(double *dev_um, double *dev_vm, double *dev_p, double *dev_ru)
{
    LEGACY_CTX();
    cfdk_resu(&c->g, &c->P, 0, dev_um, dev_vm, dev_p, NULL, dev_ru, c->s);
}

void RESV(double *dev_um, double *dev_vm, double *dev_p, double *dev_t, double *dev_rv)
{
    LEGACY_CTX();
    cfdk_resv(&c->g, &c->P, 0, dev_um, dev_vm, dev_p, dev_t, dev_rv, c->s);
}

void RESZ(double *dev_um_n, double *dev_vm_n, double *dev_z, double *dev_rz)
{
    LEGACY_CTX();
    cfdk_ressc(&c->g, &c->P, 0, dev_um_n, dev_vm_n, dev_z, NULL, dev_rz, c->s);
}

void RESC(double *dev_um_n, double *dev_vm_n, double *dev_c, double *dev_rc)
{
    LEGACY_CTX();
    cfdk_ressc(&c->g, &c->P, 1, dev_um_n, dev_vm_n, dev_c, NULL, dev_rc, c->s);
}

void bcUV(double *dev_um, double *dev_vm)
{
    LEGACY_CTX();
    cfdk_bc_uv(&c->g, &c->P, dev_um, dev_vm, c->s);
}

void bcP(double *dev_pn) { LEGACY_CTX(); cfdk_bc_s(&c->g, &c->P, 0, dev_pn, c->s); }
void bcZ(double *dev_zt) { LEGACY_CTX(); cfdk_bc_s(&c->g, &c->P, 1, dev_zt, c->s); }
void bcC(double *dev_cn) { LEGACY_CTX(); cfdk_bc_s(&c->g, &c->P, 2, dev_cn, c->s); }

double solve_U(double *dev_um, double *dev_vm, double *dev_um_n, double *dev_um_tau, double *dev_vm_tau,
               double *dev_um_n_tau, double *dev_p, double *dev_ui, double *dev_ru, double *dev_res_u)
{
    LEGACY_CTX(NAN);
    (void)dev_vm; (void)dev_um_n;
    uint64_t *slot = c->d_hist;
    if (cfdk_zero_u64(slot, 1, c->s) != CFD_OK) return NAN;
    if (x_solve_u(&e, dev_um, dev_um_tau, dev_vm_tau, dev_um_n_tau, dev_p, NULL, dev_ui, dev_ru, dev_res_u, slot) != CFD_OK)
        return NAN;
    return legacy_read(c, slot);
}

double solve_V(double *dev_um, double *dev_vm, double *dev_vm_n, double *dev_um_tau, double *dev_vm_tau,
               double *dev_vm_n_tau, double *dev_p, double *dev_t, double *dev_vi, double *dev_rv, double *dev_res_v)
{
    LEGACY_CTX(NAN);
    (void)dev_um; (void)dev_vm_n;
    uint64_t *slot = c->d_hist + 1;
    if (cfdk_zero_u64(slot, 1, c->s) != CFD_OK) return NAN;
    if (x_solve_v(&e, dev_vm, dev_um_tau, dev_vm_tau, dev_vm_n_tau, dev_p, dev_t, dev_vi, dev_rv, dev_res_v, slot) != CFD_OK)
        return NAN;
    return legacy_read(c, slot);
}

double solve_P(double *dev_p, double *dev_um_n, double *dev_vm_n, double *dev_pn)
{
    LEGACY_CTX(NAN);
    uint64_t *slot = c->d_hist + 2;
    if (cfdk_zero_u64(slot, 1, c->s) != CFD_OK) return NAN;
    if (x_solve_p(&e, dev_p, dev_um_n, dev_vm_n, dev_pn, c->scr[X_RP], c->scr[X_PI], c->scr[X_RESP], slot) != CFD_OK)
        return NAN;
    return legacy_read(c, slot);
}

void solve_Z(double *dev_um_n, double *dev_vm_n, double *dev_z, double *dev_z_n_tau, double *dev_z_tau, double *dev_rz)
{
    LEGACY_CTX();
    x_solve_s(&e, 0, dev_um_n, dev_vm_n, dev_z, dev_z_n_tau, dev_z_tau, c->scr[X_ZI], dev_rz);
}

void solve_C(double *dev_um_n, double *dev_vm_n, double *dev_c, double *dev_c_n_tau, double *dev_c_tau, double *dev_rc)
{
    LEGACY_CTX();
    x_solve_s(&e, 1, dev_um_n, dev_vm_n, dev_c, dev_c_n_tau, dev_c_tau, c->scr[X_CI], dev_rc);
}

/* Strip-only residuals: the upwind rows and columns of RESU / RESV. */
void upwind_U_pair(double *dev_um, double *dev_vm, double *dev_p, double *dev_ru)
{
    LEGACY_CTX();
    cfdk_resu(&c->g, &c->P, 1, dev_um, dev_vm, dev_p, NULL, dev_ru, c->s);
}

void upwind_V_pair(double *dev_um, double *dev_vm, double *dev_p, double *dev_t, double *dev_rv)
{
    LEGACY_CTX();
    cfdk_resv(&c->g, &c->P, 1, dev_um, dev_vm, dev_p, dev_t, dev_rv, c->s);
}

/* max_reduce.cu:14-41: max |a| over the first imax*jmax linear entries. */
double max_reduce(double *dev_matriz_linearizada, int n_i, int n_j)
{
    LEGACY_CTX(NAN);
    uint64_t *slot = c->d_hist;
    if (cfdk_zero_u64(slot, 1, c->s) != CFD_OK) return NAN;
    if (cfdk_maxabs_linear(dev_matriz_linearizada, (long)n_i * n_j, slot, c->s) != CFD_OK) return NAN;
    return legacy_read(c, slot);
}
