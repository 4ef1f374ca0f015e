// cfd_fused.hip — fused pseudo-step kernels and their support launchers.
#include "cfd_device.h"
#include "cfd_internal.h"

using namespace cfd;

namespace {
inline int cdiv(long a, long b) { return (int)((a + b - 1) / b); }
inline int launched() { return hipGetLastError() == hipSuccess ? CFD_OK : CFD_ERR_HIP; }

inline Geo mk(const cfd_devgeo *g)
{
    Geo G;
    G.x = g->x; G.y = g->y; G.xm = g->xm; G.ym = g->ym;
    G.aun = g->aun; G.aus = g->aus; G.aue = g->aue; G.auw = g->auw;
    G.avn = g->avn; G.avs = g->avs; G.ave = g->ave; G.avw = g->avw;
    G.eps = g->eps; G.lp = g->lp; G.imax = g->imax; G.jmax = g->jmax;
    return G;
}

__global__ void k_factors(Geo g, double *rdx, double *rdy, double *rdxm, double *rdym)
{
    const int t = blockIdx.x * blockDim.x + threadIdx.x;
    if (t >= 2 && t <= g.imax) rdx[t] = 1.0 / (g.x[t] - g.x[t - 1]);
    if (t >= 2 && t <= g.jmax) rdy[t] = 1.0 / (g.y[t] - g.y[t - 1]);
    if (t >= 2 && t <= g.imax + 1) rdxm[t] = 1.0 / (g.xm[t] - g.xm[t - 1]);
    if (t >= 2 && t <= g.jmax + 1) rdym[t] = 1.0 / (g.ym[t] - g.ym[t - 1]);
}

// Boundary cells of the five outputs of a fused step, from their interior values. The
// reference v corner (imax, jmax+1) reads the pre-BC column imax, which in the reference
// holds the previous step's output: that is the step's input iterate vm_prev.
__global__ void k_materialize(Geo g, cfd_params P, double *u, double *v, double *p, double *T,
                              double *C, const double *vprev)
{
    const int im = g.imax, jm = g.jmax;
    const int t = blockIdx.x * blockDim.x + threadIdx.x;
    const bool chan = P.bc == CFD_BC_CHANNEL;
    // u: rows j=1,jm (i=1..im+1) and columns i=1,2,im,im+1 (j=2..jm-1)
    const int nu_r = im + 1, nu_c = jm - 2;
    const int n_u = 2 * nu_r + 4 * nu_c;
    // v: rows j=1,2,jm,jm+1 (i=1..im) and columns i=1,im (j=3..jm-1)
    const int n_v = 4 * im + 2 * (jm - 3);
    // scalars: rows j=1,jm (i=1..im) and columns i=1,im (j=2..jm-1), three arrays
    const int n_s = 2 * im + 2 * (jm - 2);
    if (t < n_u) {
        int i, j;
        if (t < 2 * nu_r) { i = 1 + t % nu_r; j = t < nu_r ? 1 : jm; }
        else { const int q = t - 2 * nu_r, c = q / nu_c; j = 2 + q % nu_c; i = c == 0 ? 1 : (c == 1 ? 2 : (c == 2 ? im : im + 1)); }
        u[iu(g, i, j)] = chan ? bc_u_chan(g, P, u, i, j) : bc_u_ref(g, u, i, j);
        return;
    }
    int q = t - n_u;
    if (q < n_v) {
        int i, j;
        if (q < 4 * im) { const int r = q / im; i = 1 + q % im; j = r == 0 ? 1 : (r == 1 ? 2 : (r == 2 ? jm : jm + 1)); }
        else { const int c = q - 4 * im, nc = jm - 3; i = c < nc ? 1 : im; j = 3 + c % nc; }
        double val;
        if (!chan && i == im && j == jm + 1) {
            const double a = vprev[iv(g, im, jm - 2)], b = vprev[iv(g, im, jm - 1)];
            val = 2.0 * b - (2.0 * a - b);
        } else {
            val = chan ? bc_v_chan(g, v, i, j) : bc_v_ref(g, P, v, i, j);
        }
        v[iv(g, i, j)] = val;
        return;
    }
    q -= n_v;
    if (q < 3 * n_s) {
        const int kind = q / n_s, r = q % n_s;
        double *a = kind == 0 ? p : (kind == 1 ? T : C);
        int i, j;
        if (r < 2 * im) { i = 1 + r % im; j = r < im ? 1 : jm; }
        else { const int c = r - 2 * im, nc = jm - 2; i = c < nc ? 1 : im; j = 2 + c % nc; }
        a[iu(g, i, j)] = chan ? bc_s_chan(g, P, kind, a, i, j) : bc_s_ref(g, P, kind, a, i, j);
    }
}
}  // namespace

extern "C" {

int cfdk_fused_available(void) { return 0; }

int cfdk_factors(const cfd_devgeo *g, cfd_devfactors *f, hipStream_t s)
{
    const int n = (g->imax > g->jmax ? g->imax : g->jmax) + 2;
    hipLaunchKernelGGL(k_factors, dim3(cdiv(n, 256)), dim3(256), 0, s, mk(g), f->rdx, f->rdy, f->rdxm, f->rdym);
    return launched();
}

int cfdk_materialize_bc(const cfd_devgeo *g, const cfd_params *P, double *um_n, double *vm_n,
                        double *pn, double *T_n, double *C_n, const double *vm_prev, hipStream_t s)
{
    const int im = g->imax, jm = g->jmax;
    const int n = 2 * (im + 1) + 4 * (jm - 2) + 4 * im + 2 * (jm - 3) + 3 * (2 * im + 2 * (jm - 2));
    hipLaunchKernelGGL(k_materialize, dim3(cdiv(n, 256)), dim3(256), 0, s, mk(g), *P, um_n, vm_n, pn, T_n, C_n, vm_prev);
    return launched();
}

int cfdk_fused_step(const cfd_devgeo *g, const cfd_devfactors *f, const cfd_params *P,
                    const cfd_fused_args *a, int kernel, hipStream_t s)
{
    (void)g; (void)f; (void)P; (void)a; (void)kernel; (void)s;
    return CFD_ERR_UNSUPPORTED;
}

}  // extern "C"
