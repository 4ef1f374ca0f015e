/* cfd_fortran.c — Fortran-callable hot-path subroutines (include/cfd_fortran.h).
 *
 * The reference's equations.f90 / boundary.f90 subroutines with their own names and
 * argument lists, on host column-major arrays. Each call stages its arrays into device
 * buffers in the C layout (k_from_fortran / k_to_fortran transposes on the device), runs the
 * per-operator path behind include/cfd_legacy.h on the default context and stream, and
 * copies back what the reference writes. */
#include <math.h>
#include <stdlib.h>
#include <string.h>
#include <hip/hip_runtime_api.h>

#include "cfd_internal.h"
#include "../../include/cfd_legacy.h"
#include "../../include/cfd_fortran.h"

/* Transfer buffers: device arrays in the C layout per argument slot, a raw column-major
 * staging buffer, and the scratch arrays the C entry points take (ui/ru/res_u, vi/rv/res_v,
 * Zi/RZ). */
enum { FU_A, FU_B, FU_C, FU_D, FU_S1, FU_S2, FU_S3, FU_N };
enum { FV_A, FV_B, FV_C, FV_D, FV_S1, FV_S2, FV_S3, FV_N };
enum { FP_A, FP_B, FP_C, FP_D, FP_S1, FP_N };
static struct {
    int32_t imax, jmax;
    double *u[FU_N], *v[FV_N], *p[FP_N];
    double *raw;
} g_f;

static void free_buffers(void)
{
    for (int k = 0; k < FU_N; ++k) if (g_f.u[k]) hipFree(g_f.u[k]);
    for (int k = 0; k < FV_N; ++k) if (g_f.v[k]) hipFree(g_f.v[k]);
    for (int k = 0; k < FP_N; ++k) if (g_f.p[k]) hipFree(g_f.p[k]);
    if (g_f.raw) hipFree(g_f.raw);
    memset(&g_f, 0, sizeof g_f);
}

static size_t cnt(char kind)
{
    const size_t im = (size_t)g_f.imax, jm = (size_t)g_f.jmax;
    return kind == 'u' ? (im + 2) * (jm + 1) : kind == 'v' ? (im + 1) * (jm + 2) : (im + 1) * (jm + 1);
}

/* Fortran extents (n x m) and the C row stride of each kind. */
static void dims(char kind, int *n, int *m, int *ld)
{
    const int im = g_f.imax, jm = g_f.jmax;
    if (kind == 'u') { *n = im + 1; *m = jm; *ld = jm + 1; }
    else if (kind == 'v') { *n = im; *m = jm + 1; *ld = jm + 2; }
    else { *n = im; *m = jm; *ld = jm + 1; }
}

static int alloc_kind(double **p, char kind)
{
    if (hipMalloc((void **)p, cnt(kind) * sizeof(double)) != hipSuccess) return CFD_ERR_HIP;
    return hipMemset(*p, 0, cnt(kind) * sizeof(double)) == hipSuccess ? CFD_OK : CFD_ERR_HIP;
}

static int ready(void)
{
    if (!g_f.raw || !cfd_legacy_context())
        return cfd__fail(CFD_ERR_STATE, "Fortran entry points: call cfd_fortran_setup first");
    return CFD_OK;
}

/* host column-major -> device C layout (NULL stream: ordered with the legacy calls) */
static int up(double *dst, const double *host, char kind)
{
    int n, m, ld;
    dims(kind, &n, &m, &ld);
    if (hipMemcpy(g_f.raw, host, (size_t)n * m * sizeof(double), hipMemcpyHostToDevice) != hipSuccess)
        return cfd__fail(CFD_ERR_HIP, "Fortran entry points: upload failed");
    if (cfdk_from_fortran(dst, g_f.raw, n, m, ld, 0) != CFD_OK)
        return cfd__fail(CFD_ERR_HIP, "Fortran entry points: layout conversion failed");
    return CFD_OK;
}

static int down(double *host, const double *src, char kind)
{
    int n, m, ld;
    dims(kind, &n, &m, &ld);
    if (cfdk_to_fortran(g_f.raw, src, n, m, ld, 0) != CFD_OK)
        return cfd__fail(CFD_ERR_HIP, "Fortran entry points: layout conversion failed");
    if (hipMemcpy(host, g_f.raw, (size_t)n * m * sizeof(double), hipMemcpyDeviceToHost) != hipSuccess)
        return cfd__fail(CFD_ERR_HIP, "Fortran entry points: download failed");
    return CFD_OK;
}

int cfd_fortran_setup(const cfd_params *p, const double *x, const double *y, const double *epsilon1,
                      const double *liga_poros)
{
    if (!p || !x || !y) return cfd__fail(CFD_ERR_ARG, "cfd_fortran_setup: NULL argument");
    free_buffers();
    int r = cfd_legacy_init(p);
    if (r != CFD_OK) return r;
    const int32_t im = p->imax, jm = p->jmax;
    /* node coordinates with the C layout's unused element 0 */
    double *x0 = (double *)calloc((size_t)im + 1, sizeof(double)), *y0 = (double *)calloc((size_t)jm + 1, sizeof(double));
    if (!x0 || !y0) { free(x0); free(y0); return cfd__fail(CFD_ERR_ARG, "out of host memory"); }
    memcpy(x0 + 1, x, (size_t)im * sizeof(double));
    memcpy(y0 + 1, y, (size_t)jm * sizeof(double));
    cfd_ctx *c = cfd_legacy_context();
    r = cfd_ctx_set_mesh(c, x0, y0);
    free(x0); free(y0);
    if (r == CFD_OK && epsilon1) r = cfd_ctx_upload(c, CFD_F_EPS, epsilon1, CFD_LAYOUT_FORTRAN);
    if (r == CFD_OK && liga_poros) r = cfd_ctx_upload(c, CFD_F_LP, liga_poros, CFD_LAYOUT_FORTRAN);
    if (r != CFD_OK) return r;
    g_f.imax = im; g_f.jmax = jm;
    for (int k = 0; k < FU_N && r == CFD_OK; ++k) r = alloc_kind(&g_f.u[k], 'u');
    for (int k = 0; k < FV_N && r == CFD_OK; ++k) r = alloc_kind(&g_f.v[k], 'v');
    for (int k = 0; k < FP_N && r == CFD_OK; ++k) r = alloc_kind(&g_f.p[k], 'p');
    if (r == CFD_OK) r = alloc_kind(&g_f.raw, cnt('u') > cnt('v') ? 'u' : 'v');
    if (r != CFD_OK) {
        free_buffers();
        return cfd__fail(CFD_ERR_HIP, "cfd_fortran_setup: device allocation failed");
    }
    return CFD_OK;
}

void cfd_fortran_finalize(void)
{
    hipDeviceSynchronize();
    free_buffers();
    desalocar_globais();
}

/* A legacy call reports failure through the last-error text (its residual is NaN); the
 * void ones are checked with the stream's error state. */
static int launched_ok(void)
{
    hipError_t e = hipGetLastError();
    return e == hipSuccess ? CFD_OK : cfd__fail(CFD_ERR_HIP, "Fortran entry points: %s", hipGetErrorString(e));
}

void solve_u_(const double *um, const double *vm, const double *um_n, double *um_tau, double *vm_tau,
              double *um_n_tau, const double *p, double *residual_u)
{
    (void)vm; (void)um_n;   /* unused by the reference too */
    *residual_u = NAN;
    if (ready() != CFD_OK) return;
    double **U = g_f.u, **V = g_f.v, **P = g_f.p;
    if (up(U[FU_A], um, 'u') || up(U[FU_B], um_tau, 'u') || up(V[FV_A], vm_tau, 'v') ||
        up(U[FU_C], um_n_tau, 'u') || up(P[FP_A], p, 'p'))
        return;
    const double r = solve_U(U[FU_A], NULL, NULL, U[FU_B], V[FV_A], U[FU_C], P[FP_A], U[FU_S1], U[FU_S2], U[FU_S3]);
    if (isnan(r)) return;
    if (down(um_n_tau, U[FU_C], 'u') || down(vm_tau, V[FV_A], 'v')) return;
    *residual_u = r;
}

void solve_v_(const double *um, const double *vm, const double *vm_n, double *um_tau, double *vm_tau,
              double *vm_n_tau, const double *p, const double *T, double *residual_v)
{
    (void)um; (void)vm_n;
    *residual_v = NAN;
    if (ready() != CFD_OK) return;
    double **U = g_f.u, **V = g_f.v, **P = g_f.p;
    if (up(V[FV_A], vm, 'v') || up(U[FU_A], um_tau, 'u') || up(V[FV_B], vm_tau, 'v') ||
        up(V[FV_C], vm_n_tau, 'v') || up(P[FP_A], p, 'p') || up(P[FP_B], T, 'p'))
        return;
    const double r = solve_V(NULL, V[FV_A], NULL, U[FU_A], V[FV_B], V[FV_C], P[FP_A], P[FP_B], V[FV_S1], V[FV_S2],
                             V[FV_S3]);
    if (isnan(r)) return;
    if (down(vm_n_tau, V[FV_C], 'v') || down(um_tau, U[FU_A], 'u')) return;
    *residual_v = r;
}

void solve_p_(const double *p, const double *um_n, const double *vm_n, double *pn, double *residual_p)
{
    *residual_p = NAN;
    if (ready() != CFD_OK) return;
    double **U = g_f.u, **V = g_f.v, **P = g_f.p;
    if (up(P[FP_A], p, 'p') || up(U[FU_A], um_n, 'u') || up(V[FV_A], vm_n, 'v') || up(P[FP_B], pn, 'p')) return;
    const double r = solve_P(P[FP_A], U[FU_A], V[FV_A], P[FP_B]);
    if (isnan(r)) return;
    if (down(pn, P[FP_B], 'p')) return;
    *residual_p = r;
}

static void solve_scalar(int species, const double *um_n, const double *vm_n, const double *z, double *z_n_tau,
                         const double *z_tau)
{
    if (ready() != CFD_OK) return;
    double **U = g_f.u, **V = g_f.v, **P = g_f.p;
    if (up(U[FU_A], um_n, 'u') || up(V[FV_A], vm_n, 'v') || up(P[FP_A], z, 'p') || up(P[FP_B], z_n_tau, 'p') ||
        up(P[FP_C], z_tau, 'p'))
        return;
    if (species) solve_C(U[FU_A], V[FV_A], P[FP_A], P[FP_B], P[FP_C], P[FP_S1]);
    else solve_Z(U[FU_A], V[FV_A], P[FP_A], P[FP_B], P[FP_C], P[FP_S1]);
    if (launched_ok() != CFD_OK) return;
    down(z_n_tau, P[FP_B], 'p');
}

void solve_z_(const double *um_n, const double *vm_n, const double *Z, double *Z_n_tau, const double *Z_tau)
{
    solve_scalar(0, um_n, vm_n, Z, Z_n_tau, Z_tau);
}

void solve_c_(const double *um_n, const double *vm_n, const double *C, double *C_n_tau, const double *C_tau)
{
    solve_scalar(1, um_n, vm_n, C, C_n_tau, C_tau);
}

void resu_(const double *um, const double *vm, const double *p, double *RU)
{
    if (ready() != CFD_OK) return;
    double **U = g_f.u, **V = g_f.v, **P = g_f.p;
    if (up(U[FU_A], um, 'u') || up(V[FV_A], vm, 'v') || up(P[FP_A], p, 'p') || up(U[FU_B], RU, 'u')) return;
    RESU(U[FU_A], V[FV_A], P[FP_A], U[FU_B]);
    if (launched_ok() == CFD_OK) down(RU, U[FU_B], 'u');
}

void resv_(const double *um, const double *vm, const double *p, const double *T, double *RV)
{
    if (ready() != CFD_OK) return;
    double **U = g_f.u, **V = g_f.v, **P = g_f.p;
    if (up(U[FU_A], um, 'u') || up(V[FV_A], vm, 'v') || up(P[FP_A], p, 'p') || up(P[FP_B], T, 'p') ||
        up(V[FV_B], RV, 'v'))
        return;
    RESV(U[FU_A], V[FV_A], P[FP_A], P[FP_B], V[FV_B]);
    if (launched_ok() == CFD_OK) down(RV, V[FV_B], 'v');
}

static void res_scalar(int species, const double *um_n, const double *vm_n, const double *z, double *rz)
{
    if (ready() != CFD_OK) return;
    double **U = g_f.u, **V = g_f.v, **P = g_f.p;
    if (up(U[FU_A], um_n, 'u') || up(V[FV_A], vm_n, 'v') || up(P[FP_A], z, 'p') || up(P[FP_B], rz, 'p')) return;
    if (species) RESC(U[FU_A], V[FV_A], P[FP_A], P[FP_B]);
    else RESZ(U[FU_A], V[FV_A], P[FP_A], P[FP_B]);
    if (launched_ok() == CFD_OK) down(rz, P[FP_B], 'p');
}

void resz_(const double *um_n, const double *vm_n, const double *Z, double *RZ) { res_scalar(0, um_n, vm_n, Z, RZ); }
void resc_(const double *um_n, const double *vm_n, const double *C, double *RC) { res_scalar(1, um_n, vm_n, C, RC); }

void bcuv_(double *um, double *vm)
{
    if (ready() != CFD_OK) return;
    if (up(g_f.u[FU_A], um, 'u') || up(g_f.v[FV_A], vm, 'v')) return;
    bcUV(g_f.u[FU_A], g_f.v[FV_A]);
    if (launched_ok() != CFD_OK) return;
    if (down(um, g_f.u[FU_A], 'u')) return;
    down(vm, g_f.v[FV_A], 'v');
}

static void bc_scalar(int kind, double *a)
{
    if (ready() != CFD_OK) return;
    if (up(g_f.p[FP_A], a, 'p')) return;
    if (kind == 0) bcP(g_f.p[FP_A]);
    else if (kind == 1) bcZ(g_f.p[FP_A]);
    else bcC(g_f.p[FP_A]);
    if (launched_ok() == CFD_OK) down(a, g_f.p[FP_A], 'p');
}

void bcp_(double *pn) { bc_scalar(0, pn); }
void bcz_(double *Zt) { bc_scalar(1, Zt); }
void bcc_(double *Cn) { bc_scalar(2, Cn); }
