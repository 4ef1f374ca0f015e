/*
 * cfd_oracle.c — CPU oracle of the pseudo-time iteration (TEST INFRASTRUCTURE ONLY).
 *
 * Used by tests/, __graft_entry__.smoke() and bench.py's cpu_baseline leg as the checker
 * or the timed CPU baseline. It is never linked into the product (libcfd_amd.so).
 *
 * Statement-by-statement restatement of OpenMp_Multicore/equations.f90 and boundary.f90
 * on the C layout (i-major, j contiguous, 1-based with padding index 0). Each expression
 * keeps the reference's association order; compile with -ffp-contract=off.
 *
 * Additions beyond the reference (each reduces to the reference when switched off):
 *   scheme UPWIND (the reference's upwind row operator on every row), BC set CHANNEL,
 *   Boussinesq buoyancy, mu(T) polynomial, steady mode, CUDA compat switches.
 */
#include "cfd_oracle.h"
#include <math.h>
#include <stdlib.h>
#include <string.h>

struct orc_ctx {
    cfd_params P;
    int imax, jmax;
    double *x, *y, *xm, *ym;
    double *aun, *aus, *aue, *auw, *avn, *avs, *ave, *avw;
    double *eps, *lp;
    /* scratch (module arrays RU/RV/RZ/RC/Pi/RP/res_* of comum.f90:121-146) */
    double *ru, *rv, *rz, *rc, *ui, *vi, *zi, *ci, *pi, *rp, *res_u, *res_v, *res_p;
    /* physical-level temperature of the k(T) diffusivity while solve_Z runs (else NULL) */
    const double *kT;
};

#define IU(i, j) ((size_t)(i) * (size_t)(o->jmax + 1) + (size_t)(j))   /* p-like and u */
#define IV(i, j) ((size_t)(i) * (size_t)(o->jmax + 2) + (size_t)(j))   /* v */

static size_t np_(const orc_ctx *o) { return (size_t)(o->imax + 1) * (size_t)(o->jmax + 1); }
static size_t nu_(const orc_ctx *o) { return (size_t)(o->imax + 2) * (size_t)(o->jmax + 1); }
static size_t nv_(const orc_ctx *o) { return (size_t)(o->imax + 1) * (size_t)(o->jmax + 2); }

static double *dup_(const double *a, size_t n)
{
    double *b = (double *)calloc(n, sizeof(double));
    if (a) memcpy(b, a, n * sizeof(double));
    return b;
}

static void alloc_common_(orc_ctx *o)
{
    o->ru = dup_(NULL, nu_(o)); o->ui = dup_(NULL, nu_(o));
    o->rv = dup_(NULL, nv_(o)); o->vi = dup_(NULL, nv_(o));
    o->rz = dup_(NULL, np_(o)); o->rc = dup_(NULL, np_(o));
    o->zi = dup_(NULL, np_(o)); o->ci = dup_(NULL, np_(o));
    o->pi = dup_(NULL, np_(o)); o->rp = dup_(NULL, np_(o));
    o->res_u = dup_(NULL, nu_(o)); o->res_v = dup_(NULL, nv_(o)); o->res_p = dup_(NULL, np_(o));
}

orc_ctx *orc_create_full(const cfd_params *prm, const double *x, const double *y,
                         const double *xm, const double *ym,
                         const double *aun, const double *aus, const double *aue, const double *auw,
                         const double *avn, const double *avs, const double *ave, const double *avw,
                         const double *eps, const double *lp)
{
    orc_ctx *o = (orc_ctx *)calloc(1, sizeof(orc_ctx));
    int im = prm->imax, jm = prm->jmax;
    o->P = *prm; o->imax = im; o->jmax = jm;
    o->x = dup_(x, im + 1);  o->y = dup_(y, jm + 1);
    o->xm = dup_(xm, im + 2); o->ym = dup_(ym, jm + 2);
    o->aun = dup_(aun, im + 2); o->aus = dup_(aus, im + 2);
    o->aue = dup_(aue, jm + 1); o->auw = dup_(auw, jm + 1);
    o->avn = dup_(avn, im + 1); o->avs = dup_(avs, im + 1);
    o->ave = dup_(ave, jm + 2); o->avw = dup_(avw, jm + 2);
    o->eps = dup_(eps, np_(o)); o->lp = dup_(lp, np_(o));
    alloc_common_(o);
    return o;
}

/* nonsymetric_mesh.f90:48-101 */
orc_ctx *orc_create(const cfd_params *prm, const double *x, const double *y,
                    const double *eps, const double *lp)
{
    int im = prm->imax, jm = prm->jmax;
    double *xm = dup_(NULL, im + 2), *ym = dup_(NULL, jm + 2);
    double *aun = dup_(NULL, im + 2), *aue = dup_(NULL, jm + 1);
    double *avn = dup_(NULL, im + 1), *ave = dup_(NULL, jm + 2);
    xm[1] = -(x[2] + x[1]) * 0.5;
    xm[im + 1] = (x[im] + x[im - 1]) * 0.5 + (x[im] - x[im - 1]);
    for (int i = 2; i <= im; ++i) xm[i] = (x[i] + x[i - 1]) * 0.5;
    ym[1] = y[1] - (y[2] - y[1]) * 0.5;
    ym[jm + 1] = (y[jm] + y[jm - 1]) * 0.5 + (y[jm] - y[jm - 1]);
    for (int j = 2; j <= jm; ++j) ym[j] = (y[j] + y[j - 1]) * 0.5;
    for (int i = 2; i <= im; ++i) aun[i] = x[i] - x[i - 1];
    for (int j = 1; j <= jm; ++j) aue[j] = ym[j + 1] - ym[j];
    for (int i = 1; i <= im; ++i) avn[i] = xm[i + 1] - xm[i];
    for (int j = 2; j <= jm; ++j) ave[j] = y[j] - y[j - 1];
    orc_ctx *o = orc_create_full(prm, x, y, xm, ym, aun, aun, aue, aue, avn, avn, ave, ave, eps, lp);
    free(xm); free(ym); free(aun); free(aue); free(avn); free(ave);
    return o;
}

void orc_destroy(orc_ctx *o)
{
    if (!o) return;
    double *arr[] = {o->x, o->y, o->xm, o->ym, o->aun, o->aus, o->aue, o->auw, o->avn, o->avs,
                     o->ave, o->avw, o->eps, o->lp, o->ru, o->rv, o->rz, o->rc, o->ui, o->vi,
                     o->zi, o->ci, o->pi, o->rp, o->res_u, o->res_v, o->res_p};
    for (size_t k = 0; k < sizeof(arr) / sizeof(arr[0]); ++k) free(arr[k]);
    free(o);
}

void orc_set_params(orc_ctx *o, const cfd_params *prm) { o->P = *prm; }

void orc_get_geometry(const orc_ctx *o, double *xm, double *ym, double *aun, double *aue,
                      double *avn, double *ave)
{
    memcpy(xm, o->xm, (o->imax + 2) * sizeof(double));
    memcpy(ym, o->ym, (o->jmax + 2) * sizeof(double));
    memcpy(aun, o->aun, (o->imax + 2) * sizeof(double));
    memcpy(aue, o->aue, (o->jmax + 1) * sizeof(double));
    memcpy(avn, o->avn, (o->imax + 1) * sizeof(double));
    memcpy(ave, o->ave, (o->jmax + 2) * sizeof(double));
}

/* ------------------------------------------------------------------------------------ */
/* Viscous factor eps/Re of the momentum D coefficients; with mu(T), eps*mu(T)/Re.       */
static inline double visc_(const orc_ctx *o, double eps, const double *T, size_t k)
{
    if (o->P.mu_model == CFD_MU_POLY4 && T) {
        const double t = T[k];
        const double *c = o->P.mu_coef;
        double mu = (((c[4] * t + c[3]) * t + c[2]) * t + c[1]) * t + c[0];
        return eps * mu / o->P.re;
    }
    return eps / o->P.re;
}

/* RU at one point. quick=1: equations.f90:14-104; quick=0: equations.f90:133-174. */
static void resu_point_(const orc_ctx *o, const double *um, const double *vm, const double *p,
                        const double *T, double *ru, int i, int j, int quick)
{
    const cfd_params *P = &o->P;
    const double e = o->eps[IU(i, j)];
    const double fn = 0.5 * (vm[IV(i, j + 1)] + vm[IV(i - 1, j + 1)]) * o->aun[i] / e;
    const double fs = 0.5 * (vm[IV(i, j)] + vm[IV(i - 1, j)]) * o->aus[i] / e;
    const double fe = 0.5 * (um[IU(i + 1, j)] + um[IU(i, j)]) * o->aue[j] / e;
    const double fw = 0.5 * (um[IU(i, j)] + um[IU(i - 1, j)]) * o->auw[j] / e;
    const double df = fe - fw + fn - fs;
    const double vis = visc_(o, e, T, IU(i, j));
    const double Dn = vis * o->aun[i] / (o->y[j + 1] - o->y[j]);
    const double Ds = vis * o->aus[i] / (o->y[j] - o->y[j - 1]);
    const double De = vis * o->aue[j] / (o->xm[i + 1] - o->xm[i]);
    const double Dw = vis * o->auw[j] / (o->xm[i] - o->xm[i - 1]);
    const double uW = um[IU(i - 1, j)], uE = um[IU(i + 1, j)];
    const double uS = um[IU(i, j - 1)], uN = um[IU(i, j + 1)];
    const double uP = um[IU(i, j)], vP = vm[IV(i, j)];
    double sum;
    if (quick) {
        const double afw = fw > 0.0 ? 1.0 : 0.0, afe = fe > 0.0 ? 1.0 : 0.0;
        const double afn = fn > 0.0 ? 1.0 : 0.0, afs = fs > 0.0 ? 1.0 : 0.0;
        const double aw = Dw + 0.75 * afw * fw + 0.125 * afe * fe + 0.375 * (1.0 - afw) * fw;
        const double ae = De - 0.375 * afe * fe - 0.75 * (1.0 - afe) * fe - 0.125 * (1.0 - afw) * fw;
        const double as = Ds + 0.75 * afs * fs + 0.125 * afn * fn + 0.375 * (1.0 - afs) * fs;
        const double an = Dn - 0.375 * afn * fn - 0.75 * (1.0 - afn) * fn - 0.125 * (1.0 - afs) * fs;
        const double aww = -(0.125 * afw * fw);
        const double aee = 0.125 * (1.0 - afe) * fe;
        const double ass = -(0.125 * afs * fs);
        const double ann = 0.125 * (1.0 - afn) * fn;
        const double ap = aw + ae + as + an + aww + aee + ass + ann + df;
        const double uWW = um[IU(i - 2, j)], uEE = um[IU(i + 2, j)];
        const double uSS = um[IU(i, j - 2)], uNN = um[IU(i, j + 2)];
        sum = -(ap * uP) + aww * uWW + aw * uW + aee * uEE + ae * uE + ass * uSS + as * uS
              + ann * uNN + an * uN;
    } else {
        const double aw = Dw + fmax(fw, 0.0);
        const double as = Ds + fmax(fs, 0.0);
        const double ae = De + fmax(0.0, -fe);
        const double an = Dn + fmax(0.0, -fn);
        const double ap = aw + ae + as + an + df;
        sum = -(ap * uP) + aw * uW + ae * uE + as * uS + an * uN;
    }
    const double dudxdx = o->aue[j] * (uE - uP) / (o->xm[i + 1] - o->xm[i])
                        - o->auw[j] * (uP - uW) / (o->xm[i] - o->xm[i - 1]);
    const double dxdvdy = o->aue[j] * (vm[IV(i, j + 1)] - vm[IV(i, j)]) / (o->ym[j + 1] - o->ym[j])
                        - o->auw[j] * (vm[IV(i - 1, j + 1)] - vm[IV(i - 1, j)]) / (o->ym[j + 1] - o->ym[j]);
    const double q_art = e * (p[IU(i, j)] - p[IU(i - 1, j)]) / (o->x[i] - o->x[i - 1])
                       - P->b_art * (dudxdx + dxdvdy);
    const double drag = uP / (P->re * P->darcy)
                      + P->cf / sqrt(e * P->darcy) * uP * sqrt(uP * uP + vP * vP);
    ru[IU(i, j)] = 1.0 / (o->x[i] - o->x[i - 1]) / (o->y[j] - o->y[j - 1]) * sum
                 - q_art - e * drag * o->lp[IU(i, j)] - P->g * e;
}

/* RV at one point. quick=1: equations.f90:317-408; quick=0: equations.f90:436-477. */
static void resv_point_(const orc_ctx *o, const double *um, const double *vm, const double *p,
                        const double *T, double *rv, int i, int j, int quick)
{
    const cfd_params *P = &o->P;
    const double e = o->eps[IU(i, j)];
    const double fn = 0.5 * (vm[IV(i, j)] + vm[IV(i, j + 1)]) * o->avn[i] / e;
    const double fs = 0.5 * (vm[IV(i, j)] + vm[IV(i, j - 1)]) * o->avs[i] / e;
    const double fe = 0.5 * (um[IU(i + 1, j)] + um[IU(i + 1, j - 1)]) * o->ave[j] / e;
    const double fw = 0.5 * (um[IU(i, j)] + um[IU(i, j - 1)]) * o->avw[j] / e;
    const double df = fe - fw + fn - fs;
    const double vis = visc_(o, e, T, IU(i, j));
    const double Dn = vis * o->avn[i] / (o->ym[j + 1] - o->ym[j]);
    const double Ds = vis * o->avs[i] / (o->ym[j] - o->ym[j - 1]);
    const double De = vis * o->ave[j] / (o->x[i + 1] - o->x[i]);
    const double Dw = vis * o->avw[j] / (o->x[i] - o->x[i - 1]);
    const double vW = vm[IV(i - 1, j)], vE = vm[IV(i + 1, j)];
    const double vS = vm[IV(i, j - 1)], vN = vm[IV(i, j + 1)];
    const double vP = vm[IV(i, j)], uP = um[IU(i, j)];
    double sum;
    if (quick) {
        const double afw = fw > 0.0 ? 1.0 : 0.0, afe = fe > 0.0 ? 1.0 : 0.0;
        const double afn = fn > 0.0 ? 1.0 : 0.0, afs = fs > 0.0 ? 1.0 : 0.0;
        const double aw = Dw + 0.75 * afw * fw + 0.125 * afe * fe + 0.375 * (1.0 - afw) * fw;
        const double ae = De - 0.375 * afe * fe - 0.75 * (1.0 - afe) * fe - 0.125 * (1.0 - afw) * fw;
        const double as = Ds + 0.75 * afs * fs + 0.125 * afn * fn + 0.375 * (1.0 - afs) * fs;
        const double an = Dn - 0.375 * afn * fn - 0.75 * (1.0 - afn) * fn - 0.125 * (1.0 - afs) * fs;
        const double aww = -(0.125 * afw * fw);
        const double aee = 0.125 * (1.0 - afe) * fe;
        const double ass = -(0.125 * afs * fs);
        const double ann = 0.125 * (1.0 - afn) * fn;
        const double ap = aw + ae + as + an + aww + aee + ass + ann + df;
        const double vWW = vm[IV(i - 2, j)], vEE = vm[IV(i + 2, j)];
        const double vSS = vm[IV(i, j - 2)], vNN = vm[IV(i, j + 2)];
        sum = -(ap * vP) + aww * vWW + aw * vW + aee * vEE + ae * vE + ass * vSS + as * vS
              + ann * vNN + an * vN;
    } else {
        const double aw = Dw + fmax(fw, 0.0);
        const double as = Ds + fmax(fs, 0.0);
        const double ae = De + fmax(0.0, -fe);
        const double an = Dn + fmax(0.0, -fn);
        const double ap = aw + ae + as + an + df;
        sum = -(ap * vP) + aw * vW + ae * vE + as * vS + an * vN;
    }
    const double dvdydy = o->avn[i] * (vN - vP) / (o->ym[j + 1] - o->ym[j])
                        - o->avs[i] * (vP - vS) / (o->ym[j] - o->ym[j - 1]);
    const double dydudx = o->avn[i] * (um[IU(i + 1, j)] - um[IU(i, j)]) / (o->xm[i + 1] - o->xm[i])
                        - o->avs[i] * (um[IU(i + 1, j - 1)] - um[IU(i, j - 1)]) / (o->xm[i + 1] - o->xm[i]);
    const double q_art = e * (p[IU(i, j)] - p[IU(i, j - 1)]) / (o->y[j] - o->y[j - 1])
                       - P->b_art * (dydudx + dvdydy);
    double buoy;
    if (P->buoyancy == CFD_BUOY_BOUSSINESQ)
        buoy = P->ri * ((T[IU(i, j)] + T[IU(i, j - 1)]) * 0.5 - P->t_ref);
    else
        buoy = P->invfr2 * (1.0 - 1.0 / ((T[IU(i, j)] + T[IU(i, j - 1)]) * 0.5));
    const double drag = vP / (P->re * P->darcy)
                      + P->cf / sqrt(e * P->darcy) * vP * sqrt(uP * uP + vP * vP);
    rv[IV(i, j)] = 1.0 / (o->x[i] - o->x[i - 1]) / (o->y[j] - o->y[j - 1]) * sum
                 - q_art + buoy - e * drag * o->lp[IU(i, j)];
}

void orc_upwind_ui(orc_ctx *o, const double *um, const double *vm, const double *p, const double *T, double *ru, int j)
{
    for (int i = 2; i <= o->imax; ++i) resu_point_(o, um, vm, p, T, ru, i, j, 0);
}
void orc_upwind_uj(orc_ctx *o, const double *um, const double *vm, const double *p, const double *T, double *ru, int i)
{
    for (int j = 2; j <= o->jmax - 1; ++j) resu_point_(o, um, vm, p, T, ru, i, j, 0);
}
void orc_upwind_vi(orc_ctx *o, const double *um, const double *vm, const double *p, const double *T, double *rv, int j)
{
    for (int i = 2; i <= o->imax - 1; ++i) resv_point_(o, um, vm, p, T, rv, i, j, 0);
}
void orc_upwind_vj(orc_ctx *o, const double *um, const double *vm, const double *p, const double *T, double *rv, int i)
{
    /* Fortran: j = 2..jmax (equations.f90:496); CUDA: j <= jmax-1 (upwind_V_pair.cu:22) */
    int jend = o->P.compat == CFD_COMPAT_CUDA ? o->jmax - 1 : o->jmax;
    for (int j = 2; j <= jend; ++j) resv_point_(o, um, vm, p, T, rv, i, j, 0);
}

/* equations.f90:2-116 */
void orc_resu(orc_ctx *o, const double *um, const double *vm, const double *p, const double *T, double *ru)
{
    const int im = o->imax, jm = o->jmax;
    if (o->P.scheme == CFD_SCHEME_UPWIND) {
        for (int i = 2; i <= im; ++i)
            for (int j = 2; j <= jm - 1; ++j) resu_point_(o, um, vm, p, T, ru, i, j, 0);
        return;
    }
    #pragma omp parallel for schedule(static)
    for (int i = 3; i <= im - 1; ++i)
        for (int j = 3; j <= jm - 2; ++j) resu_point_(o, um, vm, p, T, ru, i, j, 1);
    orc_upwind_ui(o, um, vm, p, T, ru, 2);
    orc_upwind_ui(o, um, vm, p, T, ru, jm - 1);
    orc_upwind_uj(o, um, vm, p, T, ru, 2);
    orc_upwind_uj(o, um, vm, p, T, ru, im);
}

/* equations.f90:305-420 */
void orc_resv(orc_ctx *o, const double *um, const double *vm, const double *p, const double *T, double *rv)
{
    const int im = o->imax, jm = o->jmax;
    if (o->P.scheme == CFD_SCHEME_UPWIND) {
        for (int i = 2; i <= im - 1; ++i)
            for (int j = 2; j <= jm; ++j) resv_point_(o, um, vm, p, T, rv, i, j, 0);
        return;
    }
    #pragma omp parallel for schedule(static)
    for (int i = 3; i <= im - 2; ++i)
        for (int j = 3; j <= jm - 1; ++j) resv_point_(o, um, vm, p, T, rv, i, j, 1);
    orc_upwind_vi(o, um, vm, p, T, rv, 2);
    /* Fortran strip row j=jmax (equations.f90:414); CUDA row j=jmax-1 (upwind_V_pair.cu:114-117) */
    orc_upwind_vi(o, um, vm, p, T, rv, o->P.compat == CFD_COMPAT_CUDA ? jm - 1 : jm);
    orc_upwind_vj(o, um, vm, p, T, rv, 2);
    orc_upwind_vj(o, um, vm, p, T, rv, im - 1);
}

/* equations.f90:735-762 */
void orc_resz(orc_ctx *o, const double *um, const double *vm, const double *z, double *rz)
{
    const int im = o->imax, jm = o->jmax;
    const double k0 = 1.0 / o->P.pe;
    const double *kT = o->P.k_model == CFD_K_POLY4 ? o->kT : NULL;
    const double *kc = o->P.k_coef;
    #pragma omp parallel for schedule(static)
    for (int i = 2; i <= im - 1; ++i)
        for (int j = 2; j <= jm - 1; ++j) {
            /* k(T)/Pe of the cell (the reference: 1/Pe, equations.f90:737) */
            double k = k0;
            if (kT) {
                const double t = kT[IU(i, j)];
                k = ((((kc[4] * t + kc[3]) * t + kc[2]) * t + kc[1]) * t + kc[0]) / o->P.pe;
            }
            const double dzudx = 0.5 * (z[IU(i + 1, j)] + z[IU(i, j)]) * um[IU(i + 1, j)] * o->aue[j]
                               - 0.5 * (z[IU(i - 1, j)] + z[IU(i, j)]) * um[IU(i, j)] * o->auw[j];
            const double dzvdy = 0.5 * (z[IU(i, j + 1)] + z[IU(i, j)]) * vm[IV(i, j + 1)] * o->avn[i]
                               - 0.5 * (z[IU(i, j - 1)] + z[IU(i, j)]) * vm[IV(i, j)] * o->avs[i];
            const double De = (o->ym[j + 1] - o->ym[j]) * k / (o->x[i + 1] - o->x[i]);
            const double Dw = (o->ym[j + 1] - o->ym[j]) * k / (o->x[i] - o->x[i - 1]);
            const double Dn = (o->xm[i + 1] - o->xm[i]) * k / (o->y[j + 1] - o->y[j]);
            const double Ds = (o->xm[i + 1] - o->xm[i]) * k / (o->y[j] - o->y[j - 1]);
            const double Dp = De + Dw + Dn + Ds;
            const double L = o->lp[IU(i, j)], e = o->eps[IU(i, j)];
            rz[IU(i, j)] = 1.0 / (o->xm[i + 1] - o->xm[i]) / (o->ym[j + 1] - o->ym[j])
                         * (-(Dp * z[IU(i, j)]) + De * z[IU(i + 1, j)] + Dw * z[IU(i - 1, j)]
                            + Dn * z[IU(i, j + 1)] + Ds * z[IU(i, j - 1)]
                            - (1.0 - L) * (dzudx + dzvdy))
                         / (L * (e - 1.0) + 1.0);
        }
}

/* equations.f90:836-864 */
void orc_resc(orc_ctx *o, const double *um, const double *vm, const double *c, double *rc)
{
    const int im = o->imax, jm = o->jmax;
    const double k = 1.0 / o->P.re / o->P.sc;
    #pragma omp parallel for schedule(static)
    for (int i = 2; i <= im - 1; ++i)
        for (int j = 2; j <= jm - 1; ++j) {
            const double dcudx = 0.5 * (c[IU(i + 1, j)] + c[IU(i, j)]) * um[IU(i + 1, j)] * o->aue[j]
                               - 0.5 * (c[IU(i - 1, j)] + c[IU(i, j)]) * um[IU(i, j)] * o->auw[j];
            const double dcvdy = 0.5 * (c[IU(i, j + 1)] + c[IU(i, j)]) * vm[IV(i, j + 1)] * o->avn[i]
                               - 0.5 * (c[IU(i, j - 1)] + c[IU(i, j)]) * vm[IV(i, j)] * o->avs[i];
            const double De = (o->ym[j + 1] - o->ym[j]) * k / (o->x[i + 1] - o->x[i]);
            const double Dw = (o->ym[j + 1] - o->ym[j]) * k / (o->x[i] - o->x[i - 1]);
            const double Dn = (o->xm[i + 1] - o->xm[i]) * k / (o->y[j + 1] - o->y[j]);
            const double Ds = (o->xm[i + 1] - o->xm[i]) * k / (o->y[j] - o->y[j - 1]);
            const double Dp = De + Dw + Dn + Ds;
            const double L = o->lp[IU(i, j)], e = o->eps[IU(i, j)];
            rc[IU(i, j)] = 1.0 / (o->xm[i + 1] - o->xm[i]) / (o->ym[j + 1] - o->ym[j])
                         * (-(Dp * c[IU(i, j)]) + De * c[IU(i + 1, j)] + Dw * c[IU(i - 1, j)]
                            + Dn * c[IU(i, j + 1)] + Ds * c[IU(i, j - 1)]
                            - (dcudx + dcvdy) / (L * (e - 1.0) + 1.0));
        }
}

/* ------------------------------------------------------------------------------------ */
/* Boundary conditions. Reference set: boundary.f90 (sequential order kept).           */
void orc_bcuv(orc_ctx *o, double *um, double *vm)
{
    const int im = o->imax, jm = o->jmax;
    if (o->P.bc == CFD_BC_CHANNEL) {
        for (int i = 1; i <= im; ++i) { vm[IV(i, 1)] = 0.0; vm[IV(i, 2)] = 0.0; }
        for (int i = 2; i <= im; ++i) um[IU(i, 1)] = 0.0;
        for (int i = 2; i <= im; ++i) { vm[IV(i, jm)] = 0.0; vm[IV(i, jm + 1)] = 0.0; }
        for (int i = 2; i <= im; ++i) um[IU(i, jm)] = 0.0;
        for (int j = 1; j <= jm; ++j) { um[IU(2, j)] = o->P.u_in; um[IU(1, j)] = o->P.u_in; }
        for (int j = 1; j <= jm + 1; ++j) vm[IV(1, j)] = 0.0;
        for (int j = 1; j <= jm; ++j) { um[IU(im, j)] = um[IU(im - 1, j)]; um[IU(im + 1, j)] = um[IU(im, j)]; }
        for (int j = 2; j <= jm; ++j) vm[IV(im, j)] = vm[IV(im - 1, j)];
        return;
    }
    /* boundary.f90:9-46 */
    for (int i = 1; i <= im; ++i) { vm[IV(i, 1)] = o->P.v_i; vm[IV(i, 2)] = vm[IV(i, 1)]; }
    for (int i = 2; i <= im; ++i) um[IU(i, 1)] = 0.0;
    for (int i = 2; i <= im; ++i) {
        vm[IV(i, jm)] = 2.0 * vm[IV(i, jm - 2)] - vm[IV(i, jm - 1)];
        vm[IV(i, jm + 1)] = 2.0 * vm[IV(i, jm - 1)] - vm[IV(i, jm)];
    }
    for (int i = 2; i <= im; ++i) um[IU(i, jm)] = um[IU(i, jm - 1)];
    for (int j = 1; j <= jm; ++j) { um[IU(2, j)] = 0.0; um[IU(1, j)] = 0.0; }
    for (int j = 1; j <= jm + 1; ++j) vm[IV(1, j)] = vm[IV(2, j)];
    for (int j = 1; j <= jm; ++j) { um[IU(im, j)] = 0.0; um[IU(im + 1, j)] = 0.0; }
    for (int j = 2; j <= jm; ++j) vm[IV(im, j)] = vm[IV(im - 1, j)];
}

/* boundary.f90:52-73 (reference); channel: zero-gradient walls/inlet, p=0 at the outlet. */
void orc_bcp(orc_ctx *o, double *p)
{
    const int im = o->imax, jm = o->jmax;
    if (o->P.bc == CFD_BC_CHANNEL) {
        for (int i = 1; i <= im; ++i) { p[IU(i, 1)] = p[IU(i, 2)]; p[IU(i, jm)] = p[IU(i, jm - 1)]; }
        for (int j = 1; j <= jm; ++j) { p[IU(1, j)] = p[IU(2, j)]; p[IU(im, j)] = 0.0; }
        return;
    }
    for (int i = 1; i <= im; ++i) {
        p[IU(i, 1)] = p[IU(i, 2)];
        p[IU(i, jm)] = p[IU(i, jm - 1)] + 1.0 * (p[IU(i, jm - 1)] - p[IU(i, jm - 2)]);
    }
    for (int j = 1; j <= jm; ++j) { p[IU(1, j)] = p[IU(2, j)]; p[IU(im, j)] = p[IU(im - 1, j)]; }
}

/* boundary.f90:76-116 (reference); channel: T = t_bottom / t_top on the walls. */
void orc_bcz(orc_ctx *o, double *z)
{
    const int im = o->imax, jm = o->jmax;
    if (o->P.bc == CFD_BC_CHANNEL) {
        for (int i = 1; i <= im; ++i) { z[IU(i, 1)] = o->P.t_bottom; z[IU(i, jm)] = o->P.t_top; }
        for (int j = 1; j <= jm; ++j) { z[IU(1, j)] = z[IU(2, j)]; z[IU(im, j)] = z[IU(im - 1, j)]; }
        return;
    }
    for (int i = 1; i <= im; ++i) {
        z[IU(i, 1)] = o->P.tinf;
        z[IU(i, jm)] = z[IU(i, jm - 1)] + 1.0 * (z[IU(i, jm - 1)] - z[IU(i, jm - 2)]);
    }
    for (int j = 1; j <= jm; ++j) { z[IU(1, j)] = z[IU(2, j)]; z[IU(im, j)] = z[IU(im - 1, j)]; }
}

/* boundary.f90:119-159 (reference); channel: slot value on the left wall, zero-gradient elsewhere. */
void orc_bcc(orc_ctx *o, double *c)
{
    const int im = o->imax, jm = o->jmax;
    if (o->P.bc == CFD_BC_CHANNEL) {
        for (int i = 1; i <= im; ++i) { c[IU(i, 1)] = c[IU(i, 2)]; c[IU(i, jm)] = c[IU(i, jm - 1)]; }
        for (int j = 1; j <= jm; ++j) {
            const double yj = o->y[j];
            c[IU(1, j)] = (yj >= o->P.slot_lo && yj <= o->P.slot_hi) ? o->P.slot_value : 0.0;
            c[IU(im, j)] = c[IU(im - 1, j)];
        }
        return;
    }
    for (int i = 1; i <= im; ++i) {
        c[IU(i, 1)] = 0.0;
        c[IU(i, jm)] = c[IU(i, jm - 1)] + 1.0 * (c[IU(i, jm - 1)] - c[IU(i, jm - 2)]);
    }
    for (int j = 1; j <= jm; ++j) { c[IU(1, j)] = c[IU(2, j)]; c[IU(im, j)] = c[IU(im - 1, j)]; }
}

double orc_maxabs(const double *a, int ld, int i0, int i1, int j0, int j1)
{
    double m = 0.0;
    int nan = 0;
    for (int i = i0; i <= i1; ++i)
        for (int j = j0; j <= j1; ++j) {
            double v = fabs(a[(size_t)i * ld + j]);
            if (v != v) nan = 1;
            if (v > m) m = v;
        }
    return nan ? NAN : m;
}

static double sumsq_(const double *a, int ld, int i0, int i1, int j0, int j1)
{
    double s = 0.0;
    for (int i = i0; i <= i1; ++i)
        for (int j = j0; j <= j1; ++j) {
            const double v = a[(size_t)i * ld + j];
            s += v * v;
        }
    return s;
}

/* Root of the sum of squares of the last solve_U / solve_V / solve_P final-stage residual
 * arrays over the update regions, the entries MAXVAL(ABS(res_*)) reads (equations.f90:298,
 * 600,662). The reference computes the maximum only; the L2 norm is an addition. */
void orc_last_res_l2(const orc_ctx *o, double l2[3])
{
    const int im = o->imax, jm = o->jmax;
    l2[0] = sqrt(sumsq_(o->res_u, jm + 1, 3, im - 1, 2, jm - 1));
    l2[1] = sqrt(sumsq_(o->res_v, jm + 2, 2, im - 1, 3, jm - 1));
    l2[2] = sqrt(sumsq_(o->res_p, jm + 1, 2, im - 1, 2, jm - 1));
}

/* CUDA max_reduce (max_reduce.cu:14-41): max|a| over the first n*m linear entries. */
static double maxabs_linear_(const double *a, long n)
{
    double m = fabs(a[0]);
    for (long k = 0; k < n; ++k) m = fmax(fabs(a[k]), m);
    return m;
}

/* ------------------------------------------------------------------------------------ */
/* equations.f90:244-301 */
double orc_solve_u(orc_ctx *o, const double *um, double *um_tau, double *vm_tau, double *um_n_tau,
                   const double *p, const double *T)
{
    const int im = o->imax, jm = o->jmax;
    const double dt = o->P.dt, dtau = o->P.dtau;
    double *ru = o->ru, *ui = o->ui, *res = o->res_u;
    orc_resu(o, um_tau, vm_tau, p, T, ru);
    for (int i = 3; i <= im - 1; ++i)
        for (int j = 2; j <= jm - 1; ++j) {
            size_t k = IU(i, j);
            res[k] = ((um[k] - um_tau[k]) + ru[k] * dt) * dtau;
            ui[k] = um_tau[k] + res[k];
        }
    orc_bcuv(o, ui, vm_tau);
    orc_resu(o, ui, vm_tau, p, T, ru);
    for (int i = 3; i <= im - 1; ++i)
        for (int j = 2; j <= jm - 1; ++j) {
            size_t k = IU(i, j);
            res[k] = ((um[k] - um_tau[k]) + ru[k] * dt) * dtau;
            ui[k] = 0.75 * um_tau[k] + 0.25 * (ui[k] + res[k]);
        }
    orc_bcuv(o, ui, vm_tau);
    orc_resu(o, ui, vm_tau, p, T, ru);
    for (int i = 3; i <= im - 1; ++i)
        for (int j = 2; j <= jm - 1; ++j) {
            size_t k = IU(i, j);
            res[k] = ((um[k] - um_tau[k]) + ru[k] * dt) * dtau;
            um_n_tau[k] = 1.0 / 3.0 * um_tau[k] + 2.0 / 3.0 * (ui[k] + res[k]);
        }
    orc_bcuv(o, um_n_tau, vm_tau);
    /* c_cuda/solve_U.cu:58: max_reduce(dev_res_u, imax-1, jmax-2) */
    if (o->P.compat == CFD_COMPAT_CUDA)
        return maxabs_linear_(res, (long)(im - 1) * (jm - 2));
    return orc_maxabs(res, jm + 1, 3, im - 1, 2, jm - 1);
}

/* equations.f90:546-603 */
double orc_solve_v(orc_ctx *o, const double *vm, double *um_tau, double *vm_tau, double *vm_n_tau,
                   const double *p, const double *T)
{
    const int im = o->imax, jm = o->jmax;
    const double dt = o->P.dt, dtau = o->P.dtau;
    double *rv = o->rv, *vi = o->vi, *res = o->res_v;
    orc_resv(o, um_tau, vm_tau, p, T, rv);
    for (int i = 2; i <= im - 1; ++i)
        for (int j = 3; j <= jm - 1; ++j) {
            size_t k = IV(i, j);
            res[k] = ((vm[k] - vm_tau[k]) + rv[k] * dt) * dtau;
            vi[k] = (vm_tau[k] + res[k]);
        }
    orc_bcuv(o, um_tau, vi);
    orc_resv(o, um_tau, vi, p, T, rv);
    for (int i = 2; i <= im - 1; ++i)
        for (int j = 3; j <= jm - 1; ++j) {
            size_t k = IV(i, j);
            res[k] = ((vm[k] - vm_tau[k]) + rv[k] * dt) * dtau;
            vi[k] = (0.75 * vm_tau[k] + 0.25 * (vi[k] + res[k]));
        }
    orc_bcuv(o, um_tau, vi);
    orc_resv(o, um_tau, vi, p, T, rv);
    for (int i = 2; i <= im - 1; ++i)
        for (int j = 3; j <= jm - 1; ++j) {
            size_t k = IV(i, j);
            res[k] = ((vm[k] - vm_tau[k]) + rv[k] * dt) * dtau;
            vm_n_tau[k] = 1.0 / 3.0 * vm_tau[k] + 2.0 / 3.0 * (vi[k] + res[k]);
        }
    orc_bcuv(o, um_tau, vm_n_tau);
    /* c_cuda/solve_V.cu:52: max_reduce(dev_res_v, imax-1, jmax) */
    if (o->P.compat == CFD_COMPAT_CUDA)
        return maxabs_linear_(res, (long)(im - 1) * jm);
    return orc_maxabs(res, jm + 2, 2, im - 1, 3, jm - 1);
}

/* equations.f90:607-665 */
double orc_solve_p(orc_ctx *o, const double *p, const double *um_n, const double *vm_n, double *pn)
{
    const int im = o->imax, jm = o->jmax;
    const double dtau = o->P.dtau, beta = o->P.beta;
    double *rp = o->rp, *pi = o->pi, *res = o->res_p;
    for (int i = 2; i <= im - 1; ++i)
        for (int j = 2; j <= jm - 1; ++j) {
            size_t k = IU(i, j);
            const double dudx = um_n[IU(i + 1, j)] * o->aue[j] - um_n[IU(i, j)] * o->auw[j];
            const double dvdy = vm_n[IV(i, j + 1)] * o->avn[i] - vm_n[IV(i, j)] * o->avs[i];
            rp[k] = -(dudx + dvdy);
            pi[k] = p[k] + dtau * rp[k] * beta;
        }
    orc_bcp(o, pi);
    /* c_cuda/solve_P.cu:18,149: stage 2 starts at j=3 */
    const int j2 = o->P.compat == CFD_COMPAT_CUDA ? 3 : 2;
    for (int i = 2; i <= im - 1; ++i)
        for (int j = j2; j <= jm - 1; ++j) {
            size_t k = IU(i, j);
            pi[k] = 0.75 * p[k] + 0.25 * (pi[k] + dtau * rp[k] * beta);
        }
    orc_bcp(o, pi);
    for (int i = 2; i <= im - 1; ++i)
        for (int j = 2; j <= jm - 1; ++j) {
            size_t k = IU(i, j);
            res[k] = dtau * rp[k] * beta;
            pn[k] = 1.0 / 3.0 * p[k] + 2.0 / 3.0 * (pi[k] + res[k]);
        }
    orc_bcp(o, pn);
    /* c_cuda/solve_P.cu:50: max_reduce(dev_res_p, imax, jmax) */
    if (o->P.compat == CFD_COMPAT_CUDA)
        return maxabs_linear_(res, (long)im * jm);
    return orc_maxabs(res, jm + 1, 2, im - 1, 2, jm - 1);
}

typedef void (*res_fn_)(orc_ctx *, const double *, const double *, const double *, double *);
typedef void (*bc_fn_)(orc_ctx *, double *);

/* equations.f90:669-719 (Z) and 769-819 (C) share this structure. */
static void solve_scalar_(orc_ctx *o, res_fn_ resf, bc_fn_ bcf, double *r, double *si,
                          const double *um_n, const double *vm_n, const double *z,
                          double *z_n_tau, const double *z_tau)
{
    const int im = o->imax, jm = o->jmax;
    const double dt = o->P.dt, dtau = o->P.dtau;
    resf(o, um_n, vm_n, z_tau, r);
    for (int i = 2; i <= im - 1; ++i)
        for (int j = 2; j <= jm - 1; ++j) {
            size_t k = IU(i, j);
            const double rs = ((z[k] - z_tau[k]) + r[k] * dt) * dtau;
            si[k] = z_tau[k] + rs;
        }
    bcf(o, si);
    resf(o, um_n, vm_n, si, r);
    for (int i = 2; i <= im - 1; ++i)
        for (int j = 2; j <= jm - 1; ++j) {
            size_t k = IU(i, j);
            const double rs = ((z[k] - z_tau[k]) + r[k] * dt) * dtau;
            si[k] = 0.75 * z_tau[k] + 0.25 * (si[k] + rs);
        }
    bcf(o, si);
    resf(o, um_n, vm_n, si, r);
    for (int i = 2; i <= im - 1; ++i)
        for (int j = 2; j <= jm - 1; ++j) {
            size_t k = IU(i, j);
            const double rs = ((z[k] - z_tau[k]) + r[k] * dt) * dtau;
            z_n_tau[k] = 1.0 / 3.0 * z_tau[k] + 2.0 / 3.0 * (si[k] + rs);
        }
    bcf(o, z_n_tau);
}

void orc_solve_z(orc_ctx *o, const double *um_n, const double *vm_n, const double *z,
                 double *z_n_tau, const double *z_tau)
{
    /* k(T) takes the physical-level temperature z, fixed through the three stages */
    o->kT = z;
    solve_scalar_(o, orc_resz, orc_bcz, o->rz, o->zi, um_n, vm_n, z, z_n_tau, z_tau);
    o->kT = NULL;
}

void orc_resz_kt(orc_ctx *o, const double *um, const double *vm, const double *z, const double *T, double *rz)
{
    o->kT = T;
    orc_resz(o, um, vm, z, rz);
    o->kT = NULL;
}

void orc_solve_c(orc_ctx *o, const double *um_n, const double *vm_n, const double *c,
                 double *c_n_tau, const double *c_tau)
{
    solve_scalar_(o, orc_resc, orc_bcc, o->rc, o->ci, um_n, vm_n, c, c_n_tau, c_tau);
}

/* ------------------------------------------------------------------------------------ */
static void copy_(double *dst, const double *src, size_t n) { memcpy(dst, src, n * sizeof(double)); }

/* main.f90:120-151. With steady params the physical level is the iterate itself. */
double orc_pseudo_step(orc_ctx *o, orc_state *s, double res[3], int do_swap)
{
    const int st = o->P.steady;
    const double *um = st ? s->um_tau : s->um;
    const double *vm = st ? s->vm_tau : s->vm;
    const double *T = st ? s->T_tau : s->T;
    const double *C = st ? s->C_tau : s->C;
    double r[3];
    r[0] = orc_solve_u(o, um, s->um_tau, s->vm_tau, s->um_n_tau, s->pn, T);
    r[1] = orc_solve_v(o, vm, s->um_tau, s->vm_tau, s->vm_n_tau, s->pn, T);
    r[2] = orc_solve_p(o, s->p, s->um_n_tau, s->vm_n_tau, s->pn);
    orc_solve_z(o, s->um_n_tau, s->vm_n_tau, T, s->T_n_tau, s->T_tau);
    orc_solve_c(o, s->um_n_tau, s->vm_n_tau, C, s->C_n_tau, s->C_tau);
    double err = fmax(fmax(r[0], r[1]), r[2]);
    if (r[0] != r[0] || r[1] != r[1] || r[2] != r[2]) err = NAN;
    if (res) { res[0] = r[0]; res[1] = r[1]; res[2] = r[2]; }
    if (do_swap) {
        copy_(s->um_tau, s->um_n_tau, nu_(o));
        copy_(s->vm_tau, s->vm_n_tau, nv_(o));
        copy_(s->p, s->pn, np_(o));
        copy_(s->T_tau, s->T_n_tau, np_(o));
        copy_(s->C_tau, s->C_n_tau, np_(o));
    }
    return err;
}

void orc_pseudo_steps(orc_ctx *o, orc_state *s, int n, double *history)
{
    for (int k = 0; k < n; ++k) {
        double r[3];
        orc_pseudo_step(o, s, r, 1);
        if (history) { history[3 * k] = r[0]; history[3 * k + 1] = r[1]; history[3 * k + 2] = r[2]; }
    }
}

/* main.f90:116-157: one pass of the physical-time loop body. */
long orc_physical_step(orc_ctx *o, orc_state *s, double eps, long itc_max, int first, double res[3])
{
    long itc = first ? 1 : 0, taken = 0;
    double r[3] = {0, 0, 0};
    while (itc < itc_max) {
        double err = orc_pseudo_step(o, s, r, 0);
        ++taken;
        if (itc != 1 && err < eps) break;   /* NaN compares false: never "converged" */
        copy_(s->um_tau, s->um_n_tau, nu_(o));
        copy_(s->vm_tau, s->vm_n_tau, nv_(o));
        copy_(s->p, s->pn, np_(o));
        copy_(s->T_tau, s->T_n_tau, np_(o));
        copy_(s->C_tau, s->C_n_tau, np_(o));
        ++itc;
    }
    copy_(s->um, s->um_n_tau, nu_(o));
    copy_(s->vm, s->vm_n_tau, nv_(o));
    copy_(s->T, s->T_n_tau, np_(o));
    copy_(s->C, s->C_n_tau, np_(o));
    if (res) { res[0] = r[0]; res[1] = r[1]; res[2] = r[2]; }
    return taken;
}

/* nonsymetric_mesh.f90:7-29 (x direction), domain centred on 0 with width L. */
void orc_mesh_ref_x(int n, double L, double P, double Q, double *X)
{
    const double dxc = L / (n - 1);
    for (int i = 1; i <= n; ++i) X[i] = ((double)i - 1.0) * dxc - L / 2.0;
    const int c = n / 2 + 1;
    for (int i = c; i <= n; ++i) {
        const double eta = (X[i] - X[n]) / (X[c] - X[n]);
        const double s = P * eta + (1.0 - P) * (1.0 - (tanh(Q * (1.0 - eta))) / tanh(Q));
        X[i] = X[n] - s * (X[n] - X[c]);
    }
    for (int i = c; i >= 1; --i) {
        const double eta = (X[i] + X[n]) / (X[c] + X[n]);
        const double s = P * eta + (1.0 - P) * (1.0 - (tanh(Q * (1.0 - eta))) / tanh(Q));
        X[i] = -X[n] - s * (-X[n] - X[c]);
    }
}
