/* cfd_batch.c — batched solves: many independent cases in one persistent launch per kernel.
 *
 * Every case has its own grid size, parameters, mesh and fields; all of them live in one
 * device arena. A batched pseudo-step is five launches (residual slots, the persistent
 * u/v/p kernel walking the tile list of every case, the persistent T/C kernel, the boundary
 * cells, the per-case exit test of main.f90:139-150). The exit test runs on the device, so
 * a run captures `steps_per_graph` pseudo-steps in a hipGraph and replays it, reading one
 * integer (the number of cases still running) per replay. A case stops updating when it
 * converges (itc != 1 and max residual < tol, without the tau update, as the reference),
 * produces a non-finite residual, diverges, or reaches itc_max.
 *
 * Fields are stored in the batch precision (FP64, or FP32 with FP64 pressure). The tau and
 * n roles of each case are tracked by a parity bit on the device (one argument set per
 * parity), so no pointer swap needs the host.
 */
#include <math.h>
#include <stdarg.h>
#include <stdio.h>
#include <stdlib.h>
#include <string.h>

#include "cfd_internal.h"

int cfd__fail(int code, const char *fmt, ...);   /* cfd_host.c: thread-local last error */

#define BCHK(x) do { int r_ = (x); if (r_ != CFD_OK) return r_; } while (0)
#define BHIP(x) do { hipError_t e_ = (x); if (e_ != hipSuccess) \
    return cfd__fail(CFD_ERR_HIP, "%s: %s", #x, hipGetErrorString(e_)); } while (0)

enum { B_UM, B_VM, B_UMT, B_VMT, B_UMN, B_VMN, B_P, B_PN, B_T, B_TT, B_TN, B_C, B_CT, B_CN, B_EPS, B_LP, B_NBUF };

typedef struct bcase {
    cfd_params P;
    size_t np, nu, nv;
    void *buf[B_NBUF];          /* device pointers into the arena */
    double *y, *pack;
    int have_mesh;
} bcase;

struct cfd_batch {
    int32_t n, prec, steady, mupoly;
    size_t esz;                 /* bytes per field element (8 or 4) */
    bcase *cs;
    void *arena;                /* state + porosity */
    size_t state_bytes;         /* bytes of the state part (snapshotted by the autotuner) */
    void *geo_arena;            /* y and pack of every case */
    hipStream_t s;
    /* device control */
    int32_t *ctl, *status, *nactive;
    int64_t *itc;
    double *best, *last;
    uint64_t *res;
    void *d_k1, *d_k2, *d_mat, *d_t1, *d_t2;
    int32_t ntiles1, ntiles2, max_mat, cus;
    int32_t ntiles_all1, ntiles_all2;  /* list lengths with every case (the buffers' size) */
    int32_t lists_partial;             /* 1: the tile lists hold the running cases only */
    int32_t span;                      /* columns per workgroup of the batch kernels */
    cfd_tiles tiles;
    int args_dirty;
    int eager;                  /* 1: launch the steps directly instead of replaying a graph */
    hipEvent_t ev0, ev1;
    double last_ms;
    int64_t last_steps, last_cells;
};

static size_t align256(size_t b) { return (b + 255) & ~(size_t)255; }

static size_t field_count(const bcase *c, int k)
{
    if (k == B_UM || k == B_UMT || k == B_UMN) return c->nu;
    if (k == B_VM || k == B_VMT || k == B_VMN) return c->nv;
    return c->np;
}

static size_t field_esz(const cfd_batch *b, int k) { return (k == B_P || k == B_PN) ? 8 : b->esz; }

int cfd_batch_create(int32_t n, const cfd_params *params, int32_t precision, cfd_batch **out)
{
    if (!out) return cfd__fail(CFD_ERR_ARG, "out is NULL");
    *out = NULL;
    if (n < 1 || !params) return cfd__fail(CFD_ERR_ARG, "need at least one case");
    if (precision != CFD_PREC_FP64 && precision != CFD_PREC_FP32) return cfd__fail(CFD_ERR_ARG, "unknown precision");
    for (int32_t c = 0; c < n; ++c) {
        const cfd_params *P = &params[c];
        if (P->imax < 8 || P->jmax < 8) return cfd__fail(CFD_ERR_ARG, "case %d: grid %dx%d too small", c, P->imax, P->jmax);
        if (P->b_art != 0.0 || P->compat != CFD_COMPAT_FORTRAN)
            return cfd__fail(CFD_ERR_UNSUPPORTED, "case %d: batches run the fused path (b_art = 0, Fortran semantics)", c);
        if (P->k_model != CFD_K_CONST)
            return cfd__fail(CFD_ERR_UNSUPPORTED, "case %d: batches use the constant energy diffusivity 1/Pe", c);
        if (P->steady != params[0].steady || P->mu_model != params[0].mu_model)
            return cfd__fail(CFD_ERR_ARG, "case %d: steady and mu_model must be the same for every case", c);
        if (!(P->re > 0 && P->darcy > 0 && P->pe > 0 && P->sc > 0 && P->dtau > 0 && P->dt > 0))
            return cfd__fail(CFD_ERR_ARG, "case %d: Re, Da, Pe, Sc, dtau, dt must be positive", c);
    }
    cfd_batch *b = (cfd_batch *)calloc(1, sizeof *b);
    if (!b) return cfd__fail(CFD_ERR_ARG, "out of host memory");
    b->n = n; b->prec = precision; b->steady = params[0].steady; b->mupoly = params[0].mu_model == CFD_MU_POLY4;
    b->esz = precision == CFD_PREC_FP32 ? 4 : 8;
    b->cs = (bcase *)calloc((size_t)n, sizeof(bcase));
    int r = CFD_OK;
    if (!b->cs) { free(b); return cfd__fail(CFD_ERR_ARG, "out of host memory"); }
    size_t tot = 0, gtot = 0;
    for (int32_t c = 0; c < n; ++c) {
        bcase *k = &b->cs[c];
        k->P = params[c];
        const int im = k->P.imax, jm = k->P.jmax;
        k->np = (size_t)(im + 1) * (jm + 1);
        k->nu = (size_t)(im + 2) * (jm + 1);
        k->nv = (size_t)(im + 1) * (jm + 2);
        for (int f = 0; f < B_EPS; ++f) tot += align256(field_count(k, f) * field_esz(b, f));
        gtot += align256((size_t)(jm + 1) * 8) + align256(((size_t)4 * (im + 2) + (size_t)4 * (jm + 2)) * 8);
    }
    b->state_bytes = tot;
    for (int32_t c = 0; c < n; ++c) tot += 2 * align256(b->cs[c].np * b->esz);
    hipError_t e = hipStreamCreateWithFlags(&b->s, hipStreamNonBlocking);
    if (e == hipSuccess) e = hipMalloc(&b->arena, tot);
    if (e == hipSuccess) e = hipMemset(b->arena, 0, tot);
    if (e == hipSuccess) e = hipMalloc(&b->geo_arena, gtot);
    if (e == hipSuccess) e = hipMemset(b->geo_arena, 0, gtot);
    if (e == hipSuccess) e = hipMalloc((void **)&b->ctl, (size_t)n * sizeof(int32_t));
    if (e == hipSuccess) e = hipMalloc((void **)&b->status, (size_t)n * sizeof(int32_t));
    if (e == hipSuccess) e = hipMalloc((void **)&b->nactive, sizeof(int32_t));
    if (e == hipSuccess) e = hipMalloc((void **)&b->itc, (size_t)n * sizeof(int64_t));
    if (e == hipSuccess) e = hipMalloc((void **)&b->best, (size_t)n * sizeof(double));
    if (e == hipSuccess) e = hipMalloc((void **)&b->last, (size_t)n * 3 * sizeof(double));
    if (e == hipSuccess) e = hipMalloc((void **)&b->res, (size_t)n * 3 * sizeof(uint64_t));
    if (e == hipSuccess) e = hipMemset(b->status, 0, (size_t)n * sizeof(int32_t));
    if (e == hipSuccess) e = hipMemset(b->itc, 0, (size_t)n * sizeof(int64_t));
    if (e == hipSuccess) e = hipMemset(b->last, 0, (size_t)n * 3 * sizeof(double));
    if (e == hipSuccess) e = hipEventCreate(&b->ev0);
    if (e == hipSuccess) e = hipEventCreate(&b->ev1);
    if (e == hipSuccess) {
        int dev = 0;
        e = hipGetDevice(&dev);
        if (e == hipSuccess) e = hipDeviceGetAttribute(&b->cus, hipDeviceAttributeMultiprocessorCount, dev);
    }
    if (e != hipSuccess) {
        r = cfd__fail(CFD_ERR_HIP, "batch allocation: %s", hipGetErrorString(e));
        cfd_batch_destroy(b);
        return r;
    }
    char *p = (char *)b->arena, *g = (char *)b->geo_arena;
    for (int32_t c = 0; c < n; ++c) {
        bcase *k = &b->cs[c];
        for (int f = 0; f < B_EPS; ++f) { k->buf[f] = p; p += align256(field_count(k, f) * field_esz(b, f)); }
    }
    for (int32_t c = 0; c < n; ++c) {
        bcase *k = &b->cs[c];
        k->buf[B_EPS] = p; p += align256(k->np * b->esz);
        k->buf[B_LP] = p; p += align256(k->np * b->esz);
        k->y = (double *)g; g += align256((size_t)(k->P.jmax + 1) * 8);
        k->pack = (double *)g; g += align256(((size_t)4 * (k->P.imax + 2) + (size_t)4 * (k->P.jmax + 2)) * 8);
    }
    /* eps = 1 until uploaded; every case starts active with parity 0 */
    int32_t *h = (int32_t *)calloc((size_t)n, sizeof(int32_t));
    for (int32_t c = 0; c < n && h; ++c) h[c] = 1;
    e = h ? hipMemcpy(b->ctl, h, (size_t)n * sizeof(int32_t), hipMemcpyHostToDevice) : hipErrorOutOfMemory;
    free(h);
    for (int32_t c = 0; c < n && e == hipSuccess; ++c) {
        bcase *k = &b->cs[c];
        void *tmp = malloc(k->np * b->esz);
        if (!tmp) { e = hipErrorOutOfMemory; break; }
        for (size_t q = 0; q < k->np; ++q) {
            if (b->prec == CFD_PREC_FP32) ((float *)tmp)[q] = 1.0f; else ((double *)tmp)[q] = 1.0;
        }
        e = hipMemcpy(k->buf[B_EPS], tmp, k->np * b->esz, hipMemcpyHostToDevice);
        free(tmp);
    }
    if (e != hipSuccess) {
        r = cfd__fail(CFD_ERR_HIP, "batch initialisation: %s", hipGetErrorString(e));
        cfd_batch_destroy(b);
        return r;
    }
    b->args_dirty = 1;
    *out = b;
    return CFD_OK;
}

int cfd_batch_destroy(cfd_batch *b)
{
    if (!b) return CFD_OK;
    if (b->s) hipStreamSynchronize(b->s);
    void *d[] = {b->arena, b->geo_arena, b->ctl, b->status, b->nactive, b->itc, b->best, b->last, b->res,
                 b->d_k1, b->d_k2, b->d_mat, b->d_t1, b->d_t2};
    for (size_t k = 0; k < sizeof d / sizeof d[0]; ++k) if (d[k]) hipFree(d[k]);
    if (b->ev0) hipEventDestroy(b->ev0);
    if (b->ev1) hipEventDestroy(b->ev1);
    if (b->s) hipStreamDestroy(b->s);
    free(b->cs);
    free(b);
    return CFD_OK;
}

static int case_ok(const cfd_batch *b, int32_t c)
{
    if (!b) return cfd__fail(CFD_ERR_STATE, "no batch");
    if (c < 0 || c >= b->n) return cfd__fail(CFD_ERR_ARG, "case %d out of range [0, %d)", c, b->n);
    return CFD_OK;
}

int cfd_batch_set_mesh(cfd_batch *b, int32_t c, const double *x, const double *y)
{
    BCHK(case_ok(b, c));
    if (!x || !y) return cfd__fail(CFD_ERR_ARG, "x or y is NULL");
    bcase *k = &b->cs[c];
    const int im = k->P.imax, jm = k->P.jmax;
    for (int i = 2; i <= im; ++i) if (!(x[i] > x[i - 1])) return cfd__fail(CFD_ERR_ARG, "case %d: x must increase", c);
    for (int j = 2; j <= jm; ++j) if (!(y[j] > y[j - 1])) return cfd__fail(CFD_ERR_ARG, "case %d: y must increase", c);
    const size_t R = (size_t)im + 2, C = (size_t)jm + 2;
    double *h = (double *)calloc(12 * (R + 1) + 12 * (C + 1) + 4 * R + 4 * C, sizeof(double));
    if (!h) return cfd__fail(CFD_ERR_ARG, "out of host memory");
    double *xm = h, *aun = xm + R + 1, *aus = aun + R + 1, *avn = aus + R + 1, *avs = avn + R + 1;
    double *ym = avs + R + 1, *aue = ym + C + 1, *auw = aue + C + 1, *ave = auw + C + 1, *avw = ave + C + 1;
    double *pack = avw + C + 1;
    int r = cfd_geometry_derive(im, jm, x, y, xm, ym, aun, aus, aue, auw, avn, avs, ave, avw);
    if (r == CFD_OK) {
        /* the packed 1-D factors of the fused kernels, same arithmetic as the context's */
        double *pr = pack, *pc = pack + 4 * R;
        for (size_t t = 0; t < R; ++t) {
            const double a = (t >= 2 && t <= (size_t)im) ? 1.0 / (x[t] - x[t - 1]) : 0.0;
            const double bb = (t >= 2 && t <= (size_t)im + 1) ? 1.0 / (xm[t] - xm[t - 1]) : 0.0;
            pr[4 * t] = (t >= 2 && t <= (size_t)im) ? aun[t] : 0.0;
            pr[4 * t + 1] = (t >= 1 && t <= (size_t)im) ? avn[t] : 0.0;
            pr[4 * t + 2] = a;
            pr[4 * t + 3] = bb;
        }
        for (size_t t = 0; t < C; ++t) {
            const double a = (t >= 2 && t <= (size_t)jm) ? 1.0 / (y[t] - y[t - 1]) : 0.0;
            const double bb = (t >= 2 && t <= (size_t)jm + 1) ? 1.0 / (ym[t] - ym[t - 1]) : 0.0;
            pc[t] = (t >= 1 && t <= (size_t)jm) ? aue[t] : 0.0;
            pc[C + t] = (t >= 2 && t <= (size_t)jm) ? ave[t] : 0.0;
            pc[2 * C + t] = a;
            pc[3 * C + t] = bb;
        }
        hipError_t e = hipMemcpy(k->pack, pack, (4 * R + 4 * C) * sizeof(double), hipMemcpyHostToDevice);
        if (e == hipSuccess) e = hipMemcpy(k->y, y, (size_t)(jm + 1) * sizeof(double), hipMemcpyHostToDevice);
        if (e != hipSuccess) r = cfd__fail(CFD_ERR_HIP, "mesh upload: %s", hipGetErrorString(e));
    }
    free(h);
    if (r == CFD_OK) k->have_mesh = 1;
    return r;
}

/* Parity and status of one case (host copies of the device control words). */
static int case_control(cfd_batch *b, int32_t c, int *parity, int *status)
{
    int32_t cs = 0, st = 0;
    BHIP(hipStreamSynchronize(b->s));
    BHIP(hipMemcpy(&cs, b->ctl + c, sizeof cs, hipMemcpyDeviceToHost));
    BHIP(hipMemcpy(&st, b->status + c, sizeof st, hipMemcpyDeviceToHost));
    *parity = (cs >> 1) & 1;
    *status = st;
    return CFD_OK;
}

/* Buffer holding field f of case c now: the tau iterate of the current parity, or the n
 * buffer when the case stopped on convergence (the reference exits without the update). */
static int field_buffer(cfd_batch *b, int32_t c, int32_t field, int *bufidx)
{
    int par = 0, st = 0;
    BCHK(case_control(b, c, &par, &st));
    const int flip = par ^ (st == 1);
    int k;
    switch (field) {
    case CFD_F_UM: k = b->steady ? (flip ? B_UMN : B_UMT) : B_UM; break;
    case CFD_F_VM: k = b->steady ? (flip ? B_VMN : B_VMT) : B_VM; break;
    case CFD_F_T: k = b->steady ? (flip ? B_TN : B_TT) : B_T; break;
    case CFD_F_C: k = b->steady ? (flip ? B_CN : B_CT) : B_C; break;
    case CFD_F_UM_TAU: k = flip ? B_UMN : B_UMT; break;
    case CFD_F_VM_TAU: k = flip ? B_VMN : B_VMT; break;
    case CFD_F_T_TAU: k = flip ? B_TN : B_TT; break;
    case CFD_F_C_TAU: k = flip ? B_CN : B_CT; break;
    case CFD_F_P: case CFD_F_PN: k = flip ? B_PN : B_P; break;
    case CFD_F_EPS: k = B_EPS; break;
    case CFD_F_LP: k = B_LP; break;
    default: return cfd__fail(CFD_ERR_ARG, "unknown field %d", field);
    }
    *bufidx = k;
    return CFD_OK;
}

int cfd_batch_upload(cfd_batch *b, int32_t c, int32_t field, const double *host)
{
    BCHK(case_ok(b, c));
    if (!host) return cfd__fail(CFD_ERR_ARG, "host pointer is NULL");
    if (field == CFD_F_PN) return cfd__fail(CFD_ERR_ARG, "CFD_F_PN is download-only");
    int k = 0;
    BCHK(field_buffer(b, c, field, &k));
    bcase *q = &b->cs[c];
    const size_t n = field_count(q, k), esz = field_esz(b, k);
    hipError_t e;
    if (esz == 8) {
        e = hipMemcpy(q->buf[k], host, n * 8, hipMemcpyHostToDevice);
    } else {
        float *t = (float *)malloc(n * sizeof(float));
        if (!t) return cfd__fail(CFD_ERR_ARG, "out of host memory");
        for (size_t i = 0; i < n; ++i) t[i] = (float)host[i];
        e = hipMemcpy(q->buf[k], t, n * sizeof(float), hipMemcpyHostToDevice);
        free(t);
    }
    if (e != hipSuccess) return cfd__fail(CFD_ERR_HIP, "upload: %s", hipGetErrorString(e));
    return CFD_OK;
}

int cfd_batch_download(cfd_batch *b, int32_t c, int32_t field, double *host)
{
    BCHK(case_ok(b, c));
    if (!host) return cfd__fail(CFD_ERR_ARG, "host pointer is NULL");
    int k = 0;
    BCHK(field_buffer(b, c, field, &k));
    bcase *q = &b->cs[c];
    const size_t n = field_count(q, k), esz = field_esz(b, k);
    hipError_t e;
    if (esz == 8) {
        e = hipMemcpy(host, q->buf[k], n * 8, hipMemcpyDeviceToHost);
    } else {
        float *t = (float *)malloc(n * sizeof(float));
        if (!t) return cfd__fail(CFD_ERR_ARG, "out of host memory");
        e = hipMemcpy(t, q->buf[k], n * sizeof(float), hipMemcpyDeviceToHost);
        for (size_t i = 0; i < n && e == hipSuccess; ++i) host[i] = (double)t[i];
        free(t);
    }
    if (e != hipSuccess) return cfd__fail(CFD_ERR_HIP, "download: %s", hipGetErrorString(e));
    return CFD_OK;
}

int cfd_batch_set_tiles(cfd_batch *b, const cfd_tiles *t)
{
    if (!b || !t) return cfd__fail(CFD_ERR_ARG, "NULL argument");
    if (t->rows_uvp < 0 || t->rows_tc < 0) return cfd__fail(CFD_ERR_ARG, "row-block heights must be >= 0");
    if (!(t->waves_uvp >= 0 && t->waves_uvp <= 3)) return cfd__fail(CFD_ERR_ARG, "waves_uvp must be 0, 1, 2 or 3");
    if (!(t->waves_tc == 0 || t->waves_tc == 1 || t->waves_tc == 2 || t->waves_tc == 4))
        return cfd__fail(CFD_ERR_ARG, "waves_tc must be 0, 1, 2 or 4");
    b->tiles = *t;
    b->args_dirty = 1;
    return CFD_OK;
}

int cfd_batch_get_tiles(const cfd_batch *b, cfd_tiles *t)
{
    if (!b || !t) return cfd__fail(CFD_ERR_ARG, "NULL argument");
    *t = b->tiles;
    return CFD_OK;
}

/* Device argument tables and tile lists for the current shapes. */
static int build_tables(cfd_batch *b)
{
    if (!b->args_dirty) return CFD_OK;
    for (int32_t c = 0; c < b->n; ++c)
        if (!b->cs[c].have_mesh) return cfd__fail(CFD_ERR_STATE, "case %d has no mesh", c);
    size_t s1, s2, sm;
    cfdk_batch_arg_sizes(b->prec, &s1, &s2, &sm);
    cfd_batch_case *cases = (cfd_batch_case *)calloc((size_t)b->n, sizeof(cfd_batch_case));
    int32_t *hb1 = (int32_t *)calloc((size_t)b->n, sizeof(int32_t)), *hb2 = (int32_t *)calloc((size_t)b->n, sizeof(int32_t));
    void *h1 = malloc(2 * s1 * b->n), *h2 = malloc(2 * s2 * b->n), *hm = malloc(2 * sm * b->n);
    int r = CFD_OK;
    if (!cases || !hb1 || !hb2 || !h1 || !h2 || !hm) r = cfd__fail(CFD_ERR_ARG, "out of host memory");
    int64_t nt1 = 0, nt2 = 0;
    int maxm = 0;
    if (r == CFD_OK) {
        int32_t *jm = (int32_t *)malloc((size_t)b->n * sizeof(int32_t));
        if (!jm) r = cfd__fail(CFD_ERR_ARG, "out of host memory");
        for (int32_t c = 0; c < b->n && r == CFD_OK; ++c) jm[c] = b->cs[c].P.jmax;
        if (r == CFD_OK) b->span = cfdk_batch_span(b->n, jm);
        free(jm);
    }
    for (int32_t c = 0; c < b->n && r == CFD_OK; ++c) {
        bcase *k = &b->cs[c];
        cfd_batch_case *bc = &cases[c];
        bc->P = k->P; bc->y = k->y; bc->pack = k->pack;
        hb1[c] = b->tiles.rows_uvp; hb2[c] = b->tiles.rows_tc;
        for (int q = 0; q < 2; ++q) {
            cfd_fused_args *a = &bc->a[q];
            const int t_u = q ? B_UMN : B_UMT, n_u = q ? B_UMT : B_UMN;
            const int t_v = q ? B_VMN : B_VMT, n_v = q ? B_VMT : B_VMN;
            const int t_p = q ? B_PN : B_P, n_p = q ? B_P : B_PN;
            const int t_T = q ? B_TN : B_TT, n_T = q ? B_TT : B_TN;
            const int t_C = q ? B_CN : B_CT, n_C = q ? B_CT : B_CN;
            a->um_tau = k->buf[t_u]; a->vm_tau = k->buf[t_v]; a->T_tau = k->buf[t_T]; a->C_tau = k->buf[t_C];
            a->um = b->steady ? k->buf[t_u] : k->buf[B_UM];
            a->vm = b->steady ? k->buf[t_v] : k->buf[B_VM];
            a->T = b->steady ? k->buf[t_T] : k->buf[B_T];
            a->C = b->steady ? k->buf[t_C] : k->buf[B_C];
            a->p = (const double *)k->buf[t_p];
            a->um_n = k->buf[n_u]; a->vm_n = k->buf[n_v]; a->T_n = k->buf[n_T]; a->C_n = k->buf[n_C];
            a->pn = (double *)k->buf[n_p];
            a->epsr = k->buf[B_EPS]; a->lpr = k->buf[B_LP];
            a->res_max = b->res + 3 * c;
            a->res_l2 = NULL;      /* batches report the maxima only */
            a->prec = b->prec;
        }
        nt1 += cfdk_batch_case_tiles(1, b->span, k->P.imax, k->P.jmax, hb1[c], c, NULL);
        nt2 += cfdk_batch_case_tiles(2, b->span, k->P.imax, k->P.jmax, hb2[c], c, NULL);
        const int im = k->P.imax, jm = k->P.jmax;
        const int m = 2 * (im + 1) + 4 * (jm - 2) + 4 * im + 2 * (jm - 3) + 3 * (2 * im + 2 * (jm - 2));
        if (m > maxm) maxm = m;
    }
    int32_t *t1 = NULL, *t2 = NULL;
    if (r == CFD_OK) {
        t1 = (int32_t *)malloc((size_t)nt1 * 4 * sizeof(int32_t));
        t2 = (int32_t *)malloc((size_t)nt2 * 4 * sizeof(int32_t));
        if (!t1 || !t2) r = cfd__fail(CFD_ERR_ARG, "out of host memory");
    }
    if (r == CFD_OK) {
        /* case-major tile lists, largest cases first so the persistent grid ends evenly */
        int32_t *order = (int32_t *)malloc((size_t)b->n * sizeof(int32_t));
        if (!order) r = cfd__fail(CFD_ERR_ARG, "out of host memory");
        for (int32_t c = 0; c < b->n && r == CFD_OK; ++c) order[c] = c;
        for (int32_t i = 1; i < b->n && r == CFD_OK; ++i) {      /* insertion sort by cells, descending */
            const int32_t v = order[i];
            const int64_t cv = (int64_t)b->cs[v].P.imax * b->cs[v].P.jmax;
            int32_t j = i - 1;
            while (j >= 0 && (int64_t)b->cs[order[j]].P.imax * b->cs[order[j]].P.jmax < cv) { order[j + 1] = order[j]; --j; }
            order[j + 1] = v;
        }
        int64_t o1 = 0, o2 = 0;
        for (int32_t q = 0; q < b->n && r == CFD_OK; ++q) {
            const int32_t c = order[q];
            o1 += cfdk_batch_case_tiles(1, b->span, b->cs[c].P.imax, b->cs[c].P.jmax, hb1[c], c, t1 + 4 * o1);
            o2 += cfdk_batch_case_tiles(2, b->span, b->cs[c].P.imax, b->cs[c].P.jmax, hb2[c], c, t2 + 4 * o2);
        }
        free(order);
        if (r == CFD_OK) r = cfdk_batch_build_args(b->prec, b->n, cases, hb1, hb2, h1, h2, hm, b->span);
    }
    if (r == CFD_OK) {
        void *old[] = {b->d_k1, b->d_k2, b->d_mat, b->d_t1, b->d_t2};
        for (size_t q = 0; q < 5; ++q) if (old[q]) hipFree(old[q]);
        b->d_k1 = b->d_k2 = b->d_mat = b->d_t1 = b->d_t2 = NULL;
        hipError_t e = hipMalloc(&b->d_k1, 2 * s1 * b->n);
        if (e == hipSuccess) e = hipMalloc(&b->d_k2, 2 * s2 * b->n);
        if (e == hipSuccess) e = hipMalloc(&b->d_mat, 2 * sm * b->n);
        if (e == hipSuccess) e = hipMalloc(&b->d_t1, (size_t)nt1 * 16);
        if (e == hipSuccess) e = hipMalloc(&b->d_t2, (size_t)nt2 * 16);
        if (e == hipSuccess) e = hipMemcpy(b->d_k1, h1, 2 * s1 * b->n, hipMemcpyHostToDevice);
        if (e == hipSuccess) e = hipMemcpy(b->d_k2, h2, 2 * s2 * b->n, hipMemcpyHostToDevice);
        if (e == hipSuccess) e = hipMemcpy(b->d_mat, hm, 2 * sm * b->n, hipMemcpyHostToDevice);
        if (e == hipSuccess) e = hipMemcpy(b->d_t1, t1, (size_t)nt1 * 16, hipMemcpyHostToDevice);
        if (e == hipSuccess) e = hipMemcpy(b->d_t2, t2, (size_t)nt2 * 16, hipMemcpyHostToDevice);
        if (e != hipSuccess) r = cfd__fail(CFD_ERR_HIP, "batch tables: %s", hipGetErrorString(e));
        b->ntiles1 = (int32_t)nt1; b->ntiles2 = (int32_t)nt2; b->max_mat = maxm;
        b->ntiles_all1 = b->ntiles1; b->ntiles_all2 = b->ntiles2;
        b->lists_partial = 0;
    }
    free(cases); free(hb1); free(hb2); free(h1); free(h2); free(hm); free(t1); free(t2);
    if (r == CFD_OK) b->args_dirty = 0;
    return r;
}

/* Rewrites the tile lists for the cases whose control word has the running bit (every case
 * when ctl is NULL or none runs), in the order build_tables uses, into the buffers it sized
 * for every case. The batch kernels hand out tiles by a fixed stride over the list, so the
 * tiles of stopped cases would still take a block each and, once the list is longer than
 * the grid, put running tiles into a second round. Which block computes a tile does not
 * change its arithmetic. */
static int compact_lists(cfd_batch *b, const int32_t *ctl)
{
    if (b->args_dirty) return cfd__fail(CFD_ERR_STATE, "tile lists not built");
    int any = 0;
    for (int32_t c = 0; ctl && c < b->n && !any; ++c) any = ctl[c] & 1;
    if (!any) ctl = NULL;
    if (!ctl && !b->lists_partial) return CFD_OK;
    int32_t *order = (int32_t *)malloc((size_t)b->n * sizeof(int32_t));
    if (!order) return cfd__fail(CFD_ERR_ARG, "out of host memory");
    int32_t m = 0;
    int64_t nt1 = 0, nt2 = 0;
    for (int32_t c = 0; c < b->n; ++c) {
        if (ctl && !(ctl[c] & 1)) continue;
        order[m++] = c;
        nt1 += cfdk_batch_case_tiles(1, b->span, b->cs[c].P.imax, b->cs[c].P.jmax, b->tiles.rows_uvp, c, NULL);
        nt2 += cfdk_batch_case_tiles(2, b->span, b->cs[c].P.imax, b->cs[c].P.jmax, b->tiles.rows_tc, c, NULL);
    }
    if (nt1 > b->ntiles_all1 || nt2 > b->ntiles_all2) { free(order); return cfd__fail(CFD_ERR_STATE, "tile lists overflow"); }
    for (int32_t i = 1; i < m; ++i) {      /* insertion sort by cells, descending (stable) */
        const int32_t v = order[i];
        const int64_t cv = (int64_t)b->cs[v].P.imax * b->cs[v].P.jmax;
        int32_t j = i - 1;
        while (j >= 0 && (int64_t)b->cs[order[j]].P.imax * b->cs[order[j]].P.jmax < cv) { order[j + 1] = order[j]; --j; }
        order[j + 1] = v;
    }
    int32_t *t1 = (int32_t *)malloc((size_t)nt1 * 4 * sizeof(int32_t)), *t2 = (int32_t *)malloc((size_t)nt2 * 4 * sizeof(int32_t));
    int r = (t1 && t2) ? CFD_OK : cfd__fail(CFD_ERR_ARG, "out of host memory");
    int64_t o1 = 0, o2 = 0;
    for (int32_t q = 0; q < m && r == CFD_OK; ++q) {
        const int32_t c = order[q];
        o1 += cfdk_batch_case_tiles(1, b->span, b->cs[c].P.imax, b->cs[c].P.jmax, b->tiles.rows_uvp, c, t1 + 4 * o1);
        o2 += cfdk_batch_case_tiles(2, b->span, b->cs[c].P.imax, b->cs[c].P.jmax, b->tiles.rows_tc, c, t2 + 4 * o2);
    }
    hipError_t e = hipSuccess;
    if (r == CFD_OK) e = hipStreamSynchronize(b->s);
    if (r == CFD_OK && e == hipSuccess) e = hipMemcpy(b->d_t1, t1, (size_t)nt1 * 16, hipMemcpyHostToDevice);
    if (r == CFD_OK && e == hipSuccess) e = hipMemcpy(b->d_t2, t2, (size_t)nt2 * 16, hipMemcpyHostToDevice);
    if (r == CFD_OK && e != hipSuccess) r = cfd__fail(CFD_ERR_HIP, "tile lists: %s", hipGetErrorString(e));
    if (r == CFD_OK) {
        b->ntiles1 = (int32_t)nt1; b->ntiles2 = (int32_t)nt2;
        b->lists_partial = ctl != NULL;
    }
    free(order); free(t1); free(t2);
    return r;
}

static int make_launch(cfd_batch *b, cfd_batch_launch *L, double tol, int64_t itc_max, double diverge)
{
    BCHK(build_tables(b));
    memset(L, 0, sizeof *L);
    L->prec = b->prec; L->steady = b->steady; L->mupoly = b->mupoly;
    L->occ1 = b->tiles.waves_uvp >= 2 ? b->tiles.waves_uvp : 1;
    L->occ2 = (b->tiles.waves_tc == 2 || b->tiles.waves_tc == 4) ? b->tiles.waves_tc : 1;
    L->ncases = b->n; L->max_mat = b->max_mat; L->span = b->span;
    L->k1 = b->d_k1; L->k2 = b->d_k2; L->mat = b->d_mat; L->tiles1 = b->d_t1; L->tiles2 = b->d_t2;
    L->ntiles1 = b->ntiles1; L->ntiles2 = b->ntiles2;
    int o1 = 1, o2 = 1;
    BCHK(cfdk_batch_occupancy(b->prec, b->steady, b->mupoly, L->occ1, L->occ2, L->span, &o1, &o2));
    const int64_t g1 = (int64_t)b->cus * o1, g2 = (int64_t)b->cus * o2;
    L->grid1 = (int32_t)(g1 < b->ntiles1 ? g1 : b->ntiles1);
    L->grid2 = (int32_t)(g2 < b->ntiles2 ? g2 : b->ntiles2);
    L->ctl = b->ctl; L->res = b->res; L->status = b->status; L->itc = b->itc; L->best = b->best; L->last = b->last;
    L->nactive = b->nactive;
    L->tol = tol; L->itc_max = itc_max; L->diverge = diverge;
    return CFD_OK;
}

int cfd_batch_apply_bc(cfd_batch *b)
{
    if (!b) return cfd__fail(CFD_ERR_STATE, "no batch");
    for (int32_t c = 0; c < b->n; ++c)
        if (b->cs[c].P.bc != CFD_BC_CHANNEL)
            return cfd__fail(CFD_ERR_UNSUPPORTED, "case %d: cfd_batch_apply_bc needs the channel BC set", c);
    cfd_batch_launch L;
    BCHK(make_launch(b, &L, 0.0, 1, 0.0));
    BCHK(cfdk_batch_apply_bc(&L, b->s));
    BHIP(hipStreamSynchronize(b->s));
    return CFD_OK;
}

/* Start of a run: every case active, counters cleared; a case that stopped on convergence
 * first takes the tau update it skipped. */
static int reset_control(cfd_batch *b)
{
    const int32_t n = b->n;
    int32_t *ctl = (int32_t *)malloc((size_t)n * sizeof(int32_t)), *st = (int32_t *)malloc((size_t)n * sizeof(int32_t));
    double *best = (double *)malloc((size_t)n * sizeof(double));
    int r = CFD_OK;
    if (!ctl || !st || !best) r = cfd__fail(CFD_ERR_ARG, "out of host memory");
    hipError_t e = hipSuccess;
    if (r == CFD_OK) e = hipStreamSynchronize(b->s);
    if (r == CFD_OK && e == hipSuccess) e = hipMemcpy(ctl, b->ctl, (size_t)n * sizeof(int32_t), hipMemcpyDeviceToHost);
    if (r == CFD_OK && e == hipSuccess) e = hipMemcpy(st, b->status, (size_t)n * sizeof(int32_t), hipMemcpyDeviceToHost);
    if (r == CFD_OK && e == hipSuccess) {
        for (int32_t c = 0; c < n; ++c) {
            int par = (ctl[c] >> 1) & 1;
            if (st[c] == 1) par ^= 1;
            ctl[c] = 1 | (par << 1);
            st[c] = 0;
            best[c] = INFINITY;
        }
        e = hipMemcpy(b->ctl, ctl, (size_t)n * sizeof(int32_t), hipMemcpyHostToDevice);
        if (e == hipSuccess) e = hipMemcpy(b->status, st, (size_t)n * sizeof(int32_t), hipMemcpyHostToDevice);
        if (e == hipSuccess) e = hipMemcpy(b->best, best, (size_t)n * sizeof(double), hipMemcpyHostToDevice);
        if (e == hipSuccess) e = hipMemset(b->itc, 0, (size_t)n * sizeof(int64_t));
    }
    if (r == CFD_OK && e != hipSuccess) r = cfd__fail(CFD_ERR_HIP, "batch reset: %s", hipGetErrorString(e));
    free(ctl); free(st); free(best);
    return r;
}

static int batch_run(cfd_batch *b, int resume, double tol, int64_t itc_max, int32_t steps_per_graph,
                     double diverge_factor, int64_t *iters, double *res, int32_t *status)
{
    if (!b) return cfd__fail(CFD_ERR_STATE, "no batch");
    if (itc_max < 1) return cfd__fail(CFD_ERR_ARG, "itc_max must be >= 1");
    if (steps_per_graph < 1) steps_per_graph = 1;
    BCHK(build_tables(b));
    if (!resume) {
        BCHK(reset_control(b));
        BCHK(compact_lists(b, NULL));
    } else {
        /* cases stopped only by the previous itc_max go on; the others stay stopped */
        const int32_t n = b->n;
        int32_t *ctl = (int32_t *)malloc((size_t)n * sizeof(int32_t)), *st = (int32_t *)malloc((size_t)n * sizeof(int32_t));
        hipError_t e = (ctl && st) ? hipStreamSynchronize(b->s) : hipErrorOutOfMemory;
        if (e == hipSuccess) e = hipMemcpy(ctl, b->ctl, (size_t)n * sizeof(int32_t), hipMemcpyDeviceToHost);
        if (e == hipSuccess) e = hipMemcpy(st, b->status, (size_t)n * sizeof(int32_t), hipMemcpyDeviceToHost);
        if (e == hipSuccess) {
            for (int32_t c = 0; c < n; ++c)
                if (st[c] == 4 || st[c] == 0) { ctl[c] |= 1; st[c] = 0; }
            e = hipMemcpy(b->ctl, ctl, (size_t)n * sizeof(int32_t), hipMemcpyHostToDevice);
            if (e == hipSuccess) e = hipMemcpy(b->status, st, (size_t)n * sizeof(int32_t), hipMemcpyHostToDevice);
        }
        /* the cases that stay stopped leave the tile lists */
        const int rc = e == hipSuccess ? compact_lists(b, ctl) : CFD_OK;
        free(ctl); free(st);
        if (e != hipSuccess) return cfd__fail(CFD_ERR_HIP, "batch resume: %s", hipGetErrorString(e));
        BCHK(rc);
    }
    cfd_batch_launch L;
    BCHK(make_launch(b, &L, tol, itc_max, diverge_factor));
    /* the slots of the first step; each step's exit test clears them for the next */
    BCHK(cfdk_batch_clear(&L, b->s));
    /* capture steps_per_graph batched steps + the active count (eager mode launches the
     * same sequence directly each round) */
    hipGraph_t graph = NULL;
    hipGraphExec_t exec = NULL;
    hipError_t e = hipSuccess;
    if (!b->eager) {
        BHIP(hipStreamBeginCapture(b->s, hipStreamCaptureModeThreadLocal));
        int r = CFD_OK;
        for (int32_t k = 0; k < steps_per_graph && r == CFD_OK; ++k) r = cfdk_batch_step(&L, b->s);
        if (r == CFD_OK) r = cfdk_batch_count(&L, b->s);
        e = hipStreamEndCapture(b->s, &graph);
        if (r != CFD_OK) { if (graph) hipGraphDestroy(graph); return cfd__fail(r, "batch capture failed"); }
        if (e != hipSuccess) return cfd__fail(CFD_ERR_HIP, "batch capture: %s", hipGetErrorString(e));
        e = hipGraphInstantiate(&exec, graph, NULL, NULL, 0);
        if (e != hipSuccess) { hipGraphDestroy(graph); return cfd__fail(CFD_ERR_HIP, "graph instantiate: %s", hipGetErrorString(e)); }
    }
    /* the active count is read back into pinned memory */
    int32_t *active = NULL;
    if (hipHostMalloc((void **)&active, sizeof *active, hipHostMallocDefault) != hipSuccess) {
        if (exec) hipGraphExecDestroy(exec);
        if (graph) hipGraphDestroy(graph);
        return cfd__fail(CFD_ERR_HIP, "batch run: pinned allocation failed");
    }
    *active = b->n;
    int64_t replays = 0;
    int r = CFD_OK;
    const int64_t max_replays = (itc_max + steps_per_graph - 1) / steps_per_graph;
    e = hipEventRecord(b->ev0, b->s);
    while (e == hipSuccess && r == CFD_OK && *active > 0 && replays < max_replays) {
        if (!b->eager) {
            e = hipGraphLaunch(exec, b->s);
        } else {
            for (int32_t k = 0; k < steps_per_graph && r == CFD_OK; ++k) r = cfdk_batch_step(&L, b->s);
            if (r == CFD_OK) r = cfdk_batch_count(&L, b->s);
        }
        if (e == hipSuccess && r == CFD_OK)
            e = hipMemcpyAsync(active, b->nactive, sizeof *active, hipMemcpyDeviceToHost, b->s);
        if (e == hipSuccess && r == CFD_OK) e = hipStreamSynchronize(b->s);
        ++replays;
    }
    if (e == hipSuccess && r == CFD_OK) e = hipEventRecord(b->ev1, b->s);
    if (e == hipSuccess && r == CFD_OK) e = hipEventSynchronize(b->ev1);
    float ms = 0.f;
    if (e == hipSuccess && r == CFD_OK) e = hipEventElapsedTime(&ms, b->ev0, b->ev1);
    if (exec) hipGraphExecDestroy(exec);
    if (graph) hipGraphDestroy(graph);
    hipHostFree(active);
    if (r != CFD_OK) return cfd__fail(r, "batch step launch failed");
    if (e != hipSuccess) return cfd__fail(CFD_ERR_HIP, "batch run: %s", hipGetErrorString(e));
    /* results */
    const int32_t n = b->n;
    int64_t *it = (int64_t *)malloc((size_t)n * sizeof(int64_t));
    int32_t *st = (int32_t *)malloc((size_t)n * sizeof(int32_t));
    double *last = (double *)malloc((size_t)n * 3 * sizeof(double));
    if (!it || !st || !last) { free(it); free(st); free(last); return cfd__fail(CFD_ERR_ARG, "out of host memory"); }
    e = hipMemcpy(it, b->itc, (size_t)n * sizeof(int64_t), hipMemcpyDeviceToHost);
    if (e == hipSuccess) e = hipMemcpy(st, b->status, (size_t)n * sizeof(int32_t), hipMemcpyDeviceToHost);
    if (e == hipSuccess) e = hipMemcpy(last, b->last, (size_t)n * 3 * sizeof(double), hipMemcpyDeviceToHost);
    int64_t cells = 0;
    int failed = 0, unconv = 0;
    for (int32_t c = 0; c < n && e == hipSuccess; ++c) {
        cells += it[c] * (int64_t)b->cs[c].P.imax * b->cs[c].P.jmax;
        if (iters) iters[c] = it[c];
        if (status) status[c] = st[c];
        if (res) memcpy(res + 3 * c, last + 3 * c, 3 * sizeof(double));
        if (st[c] == 2 || st[c] == 3) ++failed;
        else if (st[c] != 1) ++unconv;
    }
    free(it); free(st); free(last);
    if (e != hipSuccess) return cfd__fail(CFD_ERR_HIP, "batch results: %s", hipGetErrorString(e));
    b->last_ms = ms;
    b->last_steps = replays * steps_per_graph;
    b->last_cells = cells;
    if (failed) return cfd__fail(CFD_ERR_NONFINITE, "%d case(s) produced a non-finite or diverging residual", failed);
    if (tol > 0 && unconv) return cfd__fail(CFD_ERR_NOCONV, "%d case(s) did not reach %.3e in %lld steps", unconv, tol, (long long)itc_max);
    return CFD_OK;
}

int cfd_batch_run(cfd_batch *b, double tol, int64_t itc_max, int32_t steps_per_graph, double diverge_factor,
                  int64_t *iters, double *res, int32_t *status)
{
    return batch_run(b, 0, tol, itc_max, steps_per_graph, diverge_factor, iters, res, status);
}

int cfd_batch_resume(cfd_batch *b, double tol, int64_t itc_max, int32_t steps_per_graph, double diverge_factor,
                     int64_t *iters, double *res, int32_t *status)
{
    return batch_run(b, 1, tol, itc_max, steps_per_graph, diverge_factor, iters, res, status);
}

int cfd_batch_set_graph(cfd_batch *b, int32_t use_graph)
{
    if (!b) return cfd__fail(CFD_ERR_STATE, "no batch");
    b->eager = use_graph ? 0 : 1;
    return CFD_OK;
}

int cfd_batch_stop_case(cfd_batch *b, int32_t c, int32_t status)
{
    if (!b) return cfd__fail(CFD_ERR_STATE, "no batch");
    if (c < 0 || c >= b->n) return cfd__fail(CFD_ERR_ARG, "case %d out of range", c);
    if (status < 5) return cfd__fail(CFD_ERR_ARG, "status %d: codes below 5 are the run's own", status);
    int32_t ctl = 0;
    hipError_t e = hipStreamSynchronize(b->s);
    if (e == hipSuccess) e = hipMemcpy(&ctl, b->ctl + c, sizeof ctl, hipMemcpyDeviceToHost);
    ctl &= ~1;
    if (e == hipSuccess) e = hipMemcpy(b->ctl + c, &ctl, sizeof ctl, hipMemcpyHostToDevice);
    if (e == hipSuccess) e = hipMemcpy(b->status + c, &status, sizeof status, hipMemcpyHostToDevice);
    if (e != hipSuccess) return cfd__fail(CFD_ERR_HIP, "batch stop: %s", hipGetErrorString(e));
    return CFD_OK;
}

int cfd_batch_last_timing(const cfd_batch *b, double *ms, int64_t *steps, int64_t *cell_updates)
{
    if (!b) return cfd__fail(CFD_ERR_STATE, "no batch");
    if (ms) *ms = b->last_ms;
    if (steps) *steps = b->last_steps;
    if (cell_updates) *cell_updates = b->last_cells;
    return CFD_OK;
}

/* Autotuner: candidate shapes timed on the current state, which is restored afterwards. */
int cfd_batch_autotune(cfd_batch *b, int32_t steps, cfd_tiles *out, double *out_ms)
{
    if (!b) return cfd__fail(CFD_ERR_STATE, "no batch");
    if (steps < 1) steps = 1;
    static const cfd_tiles cand1[] = {{16, 0, 1, 0}, {24, 0, 1, 0}, {32, 0, 1, 0}, {48, 0, 1, 0}, {64, 0, 1, 0},
                                      {16, 0, 2, 0}, {24, 0, 2, 0}, {32, 0, 2, 0}, {48, 0, 2, 0}, {64, 0, 2, 0},
                                      {16, 0, 3, 0}, {32, 0, 3, 0}, {48, 0, 3, 0}, {64, 0, 3, 0}};
    static const cfd_tiles cand2[] = {{0, 8, 0, 1}, {0, 16, 0, 1}, {0, 24, 0, 1}, {0, 32, 0, 1}, {0, 48, 0, 1},
                                      {0, 8, 0, 4}, {0, 16, 0, 4}, {0, 24, 0, 4}, {0, 32, 0, 4},
                                      {0, 8, 0, 2}, {0, 16, 0, 2}, {0, 32, 0, 2}, {0, 48, 0, 2}};
    void *snap = NULL;
    BHIP(hipStreamSynchronize(b->s));
    BHIP(hipMalloc(&snap, b->state_bytes));
    const size_t nctl = (size_t)b->n * sizeof(int32_t);
    int32_t *ctl = (int32_t *)malloc(nctl), *st = (int32_t *)malloc(nctl);
    hipError_t e = (ctl && st) ? hipMemcpy(snap, b->arena, b->state_bytes, hipMemcpyDeviceToDevice) : hipErrorOutOfMemory;
    if (e == hipSuccess) e = hipMemcpy(ctl, b->ctl, nctl, hipMemcpyDeviceToHost);
    if (e == hipSuccess) e = hipMemcpy(st, b->status, nctl, hipMemcpyDeviceToHost);
    int r = e == hipSuccess ? CFD_OK : cfd__fail(CFD_ERR_HIP, "autotune snapshot: %s", hipGetErrorString(e));
    const cfd_tiles prev = b->tiles;
    cfd_tiles best_t = {32, 16, 1, 1};
    double best = INFINITY;
    for (int pass = 0; pass < 2 && r == CFD_OK; ++pass) {
        const cfd_tiles *cand = pass ? cand2 : cand1;
        const size_t nc = pass ? sizeof cand2 / sizeof cand2[0] : sizeof cand1 / sizeof cand1[0];
        for (size_t q = 0; q < nc && r == CFD_OK; ++q) {
            cfd_tiles t = best_t;
            if (pass == 0) { t.rows_uvp = cand[q].rows_uvp; t.waves_uvp = cand[q].waves_uvp; }
            else { t.rows_tc = cand[q].rows_tc; t.waves_tc = cand[q].waves_tc; }
            b->tiles = t;
            b->args_dirty = 1;
            /* tol 0: nothing converges; one untimed graph replay, then a timed one */
            r = cfd_batch_run(b, 0.0, 2 * steps, steps, 0.0, NULL, NULL, NULL);
            if (r == CFD_ERR_NONFINITE) r = CFD_OK;   /* a failing case does not change the timing */
            if (r == CFD_OK) {
                const double ms = b->last_ms / (double)b->last_steps;
                if (ms < best) { best = ms; best_t = t; }
            }
        }
    }
    /* restore the state and the control words */
    hipError_t e2 = hipMemcpy(b->arena, snap, b->state_bytes, hipMemcpyDeviceToDevice);
    if (e2 == hipSuccess && ctl) e2 = hipMemcpy(b->ctl, ctl, nctl, hipMemcpyHostToDevice);
    if (e2 == hipSuccess && st) e2 = hipMemcpy(b->status, st, nctl, hipMemcpyHostToDevice);
    hipFree(snap);
    free(ctl); free(st);
    if (r == CFD_OK && e2 != hipSuccess) r = cfd__fail(CFD_ERR_HIP, "autotune restore: %s", hipGetErrorString(e2));
    if (r != CFD_OK) { b->tiles = prev; b->args_dirty = 1; return r; }
    b->tiles = best_t;
    b->args_dirty = 1;
    if (out) *out = best_t;
    if (out_ms) *out_ms = best;
    return CFD_OK;
}

/* Copies every case's current iterates (and eps, L) from src into dst, converting between
 * precisions on the device; the cases must have the same grid sizes. dst continues from
 * them like from an upload (its next run starts a fresh exit-test count). */
int cfd_batch_copy_state(cfd_batch *dst, cfd_batch *src)
{
    if (!dst || !src) return cfd__fail(CFD_ERR_STATE, "no batch");
    if (dst->n != src->n || dst->steady != src->steady) return cfd__fail(CFD_ERR_ARG, "batches differ in cases or mode");
    for (int32_t c = 0; c < dst->n; ++c)
        if (dst->cs[c].P.imax != src->cs[c].P.imax || dst->cs[c].P.jmax != src->cs[c].P.jmax)
            return cfd__fail(CFD_ERR_ARG, "case %d: grid sizes differ", c);
    static const int32_t fields[] = {CFD_F_UM_TAU, CFD_F_VM_TAU, CFD_F_P, CFD_F_T_TAU, CFD_F_C_TAU, CFD_F_UM, CFD_F_VM,
                                     CFD_F_T, CFD_F_C, CFD_F_EPS, CFD_F_LP};
    const size_t nf = dst->steady ? 5 : 9;
    BHIP(hipStreamSynchronize(src->s));
    for (int32_t c = 0; c < dst->n; ++c) {
        for (size_t q = 0; q < nf + 2; ++q) {
            const int32_t f = q < nf ? fields[q] : fields[9 + (q - nf)];
            int ks = 0, kd = 0;
            BCHK(field_buffer(src, c, f, &ks));
            BCHK(field_buffer(dst, c, f, &kd));
            const bcase *a = &src->cs[c];
            bcase *b = &dst->cs[c];
            const size_t n = field_count(a, ks), es = field_esz(src, ks), ed = field_esz(dst, kd);
            if (es == ed) {
                BHIP(hipMemcpyAsync(b->buf[kd], a->buf[ks], n * es, hipMemcpyDeviceToDevice, dst->s));
            } else if (es == 4) {
                BCHK(cfdk_f2d((double *)b->buf[kd], (const float *)a->buf[ks], (long)n, dst->s));
            } else {
                BCHK(cfdk_d2f((float *)b->buf[kd], (const double *)a->buf[ks], (long)n, dst->s));
            }
        }
    }
    BHIP(hipStreamSynchronize(dst->s));
    return CFD_OK;
}
