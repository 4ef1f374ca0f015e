// cfd_device.h — device-side building blocks shared by the HIP kernels of libcfd_amd.
//
// Layout: the reference C layout (c_cuda/main.cu:59-89). p-like and u arrays use
// i*(jmax+1)+j, v arrays i*(jmax+2)+j; 1-based, index 0 is padding; j is contiguous, so
// the fast thread index always walks j and every global access is coalesced.
#pragma once
#include <hip/hip_runtime.h>
#include "../../include/cfd.h"

namespace cfd {

// Geometry and porosity, device pointers (sizes as c_cuda/comum.cu:133-164).
struct Geo {
    const double *x, *y, *xm, *ym;
    const double *aun, *aus, *aue, *auw, *avn, *avs, *ave, *avw;
    const double *eps, *lp;
    int imax, jmax;
};

__device__ __forceinline__ size_t iu(const Geo &g, int i, int j) { return (size_t)i * (size_t)(g.jmax + 1) + (size_t)j; }
__device__ __forceinline__ size_t iv(const Geo &g, int i, int j) { return (size_t)i * (size_t)(g.jmax + 2) + (size_t)j; }

// ---------------------------------------------------------------------------------------
// Literal point operators: same operations in the same order as the reference
// (OpenMp_Multicore/equations.f90); compiled with -ffp-contract=off they reproduce the
// CPU oracle bit for bit. Used by the drop-in entry points (RESU, solve_U, ...).
// ---------------------------------------------------------------------------------------
__device__ __forceinline__ double visc_lit(const cfd_params &P, double e, const double *T, size_t k)
{
    if (P.mu_model == CFD_MU_POLY4 && T) {
        const double t = T[k];
        const double mu = (((P.mu_coef[4] * t + P.mu_coef[3]) * t + P.mu_coef[2]) * t + P.mu_coef[1]) * t + P.mu_coef[0];
        return e * mu / P.re;
    }
    return e / P.re;
}

// equations.f90:14-104 (quick) and 133-174 (upwind)
__device__ __forceinline__ double resu_lit(const Geo &g, const cfd_params &P, const double *um, const double *vm,
                                           const double *p, const double *T, int i, int j, bool quick)
{
    const double e = g.eps[iu(g, i, j)];
    const double fn = 0.5 * (vm[iv(g, i, j + 1)] + vm[iv(g, i - 1, j + 1)]) * g.aun[i] / e;
    const double fs = 0.5 * (vm[iv(g, i, j)] + vm[iv(g, i - 1, j)]) * g.aus[i] / e;
    const double fe = 0.5 * (um[iu(g, i + 1, j)] + um[iu(g, i, j)]) * g.aue[j] / e;
    const double fw = 0.5 * (um[iu(g, i, j)] + um[iu(g, i - 1, j)]) * g.auw[j] / e;
    const double df = fe - fw + fn - fs;
    const double vis = visc_lit(P, e, T, iu(g, i, j));
    const double Dn = vis * g.aun[i] / (g.y[j + 1] - g.y[j]);
    const double Ds = vis * g.aus[i] / (g.y[j] - g.y[j - 1]);
    const double De = vis * g.aue[j] / (g.xm[i + 1] - g.xm[i]);
    const double Dw = vis * g.auw[j] / (g.xm[i] - g.xm[i - 1]);
    const double uW = um[iu(g, i - 1, j)], uE = um[iu(g, i + 1, j)];
    const double uS = um[iu(g, i, j - 1)], uN = um[iu(g, i, j + 1)];
    const double uP = um[iu(g, i, j)], vP = vm[iv(g, i, j)];
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
        sum = -(ap * uP) + aww * um[iu(g, i - 2, j)] + aw * uW + aee * um[iu(g, i + 2, j)] + ae * uE
              + ass * um[iu(g, i, j - 2)] + as * uS + ann * um[iu(g, i, j + 2)] + an * uN;
    } else {
        const double aw = Dw + fmax(fw, 0.0);
        const double as = Ds + fmax(fs, 0.0);
        const double ae = De + fmax(0.0, -fe);
        const double an = Dn + fmax(0.0, -fn);
        const double ap = aw + ae + as + an + df;
        sum = -(ap * uP) + aw * uW + ae * uE + as * uS + an * uN;
    }
    const double dudxdx = g.aue[j] * (uE - uP) / (g.xm[i + 1] - g.xm[i])
                        - g.auw[j] * (uP - uW) / (g.xm[i] - g.xm[i - 1]);
    const double dxdvdy = g.aue[j] * (vm[iv(g, i, j + 1)] - vm[iv(g, i, j)]) / (g.ym[j + 1] - g.ym[j])
                        - g.auw[j] * (vm[iv(g, i - 1, j + 1)] - vm[iv(g, i - 1, j)]) / (g.ym[j + 1] - g.ym[j]);
    const double q_art = e * (p[iu(g, i, j)] - p[iu(g, i - 1, j)]) / (g.x[i] - g.x[i - 1])
                       - P.b_art * (dudxdx + dxdvdy);
    const double drag = uP / (P.re * P.darcy) + P.cf / sqrt(e * P.darcy) * uP * sqrt(uP * uP + vP * vP);
    return 1.0 / (g.x[i] - g.x[i - 1]) / (g.y[j] - g.y[j - 1]) * sum
         - q_art - e * drag * g.lp[iu(g, i, j)] - P.g * e;
}

// equations.f90:317-408 (quick) and 436-477 (upwind)
__device__ __forceinline__ double resv_lit(const Geo &g, const cfd_params &P, const double *um, const double *vm,
                                           const double *p, const double *T, int i, int j, bool quick)
{
    const double e = g.eps[iu(g, i, j)];
    const double fn = 0.5 * (vm[iv(g, i, j)] + vm[iv(g, i, j + 1)]) * g.avn[i] / e;
    const double fs = 0.5 * (vm[iv(g, i, j)] + vm[iv(g, i, j - 1)]) * g.avs[i] / e;
    const double fe = 0.5 * (um[iu(g, i + 1, j)] + um[iu(g, i + 1, j - 1)]) * g.ave[j] / e;
    const double fw = 0.5 * (um[iu(g, i, j)] + um[iu(g, i, j - 1)]) * g.avw[j] / e;
    const double df = fe - fw + fn - fs;
    const double vis = visc_lit(P, e, T, iu(g, i, j));
    const double Dn = vis * g.avn[i] / (g.ym[j + 1] - g.ym[j]);
    const double Ds = vis * g.avs[i] / (g.ym[j] - g.ym[j - 1]);
    const double De = vis * g.ave[j] / (g.x[i + 1] - g.x[i]);
    const double Dw = vis * g.avw[j] / (g.x[i] - g.x[i - 1]);
    const double vW = vm[iv(g, i - 1, j)], vE = vm[iv(g, i + 1, j)];
    const double vS = vm[iv(g, i, j - 1)], vN = vm[iv(g, i, j + 1)];
    const double vP = vm[iv(g, i, j)], uP = um[iu(g, i, j)];
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
        sum = -(ap * vP) + aww * vm[iv(g, i - 2, j)] + aw * vW + aee * vm[iv(g, i + 2, j)] + ae * vE
              + ass * vm[iv(g, i, j - 2)] + as * vS + ann * vm[iv(g, i, j + 2)] + an * vN;
    } else {
        const double aw = Dw + fmax(fw, 0.0);
        const double as = Ds + fmax(fs, 0.0);
        const double ae = De + fmax(0.0, -fe);
        const double an = Dn + fmax(0.0, -fn);
        const double ap = aw + ae + as + an + df;
        sum = -(ap * vP) + aw * vW + ae * vE + as * vS + an * vN;
    }
    const double dvdydy = g.avn[i] * (vN - vP) / (g.ym[j + 1] - g.ym[j])
                        - g.avs[i] * (vP - vS) / (g.ym[j] - g.ym[j - 1]);
    const double dydudx = g.avn[i] * (um[iu(g, i + 1, j)] - um[iu(g, i, j)]) / (g.xm[i + 1] - g.xm[i])
                        - g.avs[i] * (um[iu(g, i + 1, j - 1)] - um[iu(g, i, j - 1)]) / (g.xm[i + 1] - g.xm[i]);
    const double q_art = e * (p[iu(g, i, j)] - p[iu(g, i, j - 1)]) / (g.y[j] - g.y[j - 1])
                       - P.b_art * (dydudx + dvdydy);
    double buoy;
    if (P.buoyancy == CFD_BUOY_BOUSSINESQ)
        buoy = P.ri * ((T[iu(g, i, j)] + T[iu(g, i, j - 1)]) * 0.5 - P.t_ref);
    else
        buoy = P.invfr2 * (1.0 - 1.0 / ((T[iu(g, i, j)] + T[iu(g, i, j - 1)]) * 0.5));
    const double drag = vP / (P.re * P.darcy) + P.cf / sqrt(e * P.darcy) * vP * sqrt(uP * uP + vP * vP);
    return 1.0 / (g.x[i] - g.x[i - 1]) / (g.y[j] - g.y[j - 1]) * sum
         - q_art + buoy - e * drag * g.lp[iu(g, i, j)];
}

// k(T)/Pe of the energy equation's diffusion (k_model CFD_K_POLY4): the conductivity over its
// reference value as a polynomial in the temperature (no counterpart in the reference's
// RESZ, whose diffusivity is the constant 1/Pe, equations.f90:737)
__device__ __forceinline__ double cond_lit(const cfd_params &P, double t)
{
    const double *c = P.k_coef;
    return ((((c[4] * t + c[3]) * t + c[2]) * t + c[1]) * t + c[0]) / P.pe;
}

// equations.f90:735-762 (k = 1/Pe, conv_in_porous = false) and 836-864 (k = 1/Re/Sc, true)
__device__ __forceinline__ double ressc_lit(const Geo &g, double k, bool species, const double *um, const double *vm,
                                            const double *z, int i, int j)
{
    const double dzudx = 0.5 * (z[iu(g, i + 1, j)] + z[iu(g, i, j)]) * um[iu(g, i + 1, j)] * g.aue[j]
                       - 0.5 * (z[iu(g, i - 1, j)] + z[iu(g, i, j)]) * um[iu(g, i, j)] * g.auw[j];
    const double dzvdy = 0.5 * (z[iu(g, i, j + 1)] + z[iu(g, i, j)]) * vm[iv(g, i, j + 1)] * g.avn[i]
                       - 0.5 * (z[iu(g, i, j - 1)] + z[iu(g, i, j)]) * vm[iv(g, i, j)] * g.avs[i];
    const double De = (g.ym[j + 1] - g.ym[j]) * k / (g.x[i + 1] - g.x[i]);
    const double Dw = (g.ym[j + 1] - g.ym[j]) * k / (g.x[i] - g.x[i - 1]);
    const double Dn = (g.xm[i + 1] - g.xm[i]) * k / (g.y[j + 1] - g.y[j]);
    const double Ds = (g.xm[i + 1] - g.xm[i]) * k / (g.y[j] - g.y[j - 1]);
    const double Dp = De + Dw + Dn + Ds;
    const double L = g.lp[iu(g, i, j)], e = g.eps[iu(g, i, j)];
    const double diff = -(Dp * z[iu(g, i, j)]) + De * z[iu(g, i + 1, j)] + Dw * z[iu(g, i - 1, j)]
                      + Dn * z[iu(g, i, j + 1)] + Ds * z[iu(g, i, j - 1)];
    const double vol = 1.0 / (g.xm[i + 1] - g.xm[i]) / (g.ym[j + 1] - g.ym[j]);
    if (species)
        return vol * (diff - (dzudx + dzvdy) / (L * (e - 1.0) + 1.0));
    return vol * (diff - (1.0 - L) * (dzudx + dzvdy)) / (L * (e - 1.0) + 1.0);
}

// ---------------------------------------------------------------------------------------
// Boundary conditions as closed-form final values.
//
// The reference applies its BC loops in sequence (boundary.f90); the value a boundary cell
// holds after the whole sequence is written here as a function of pre-BC interior values
// (reads never touch a cell the same BC writes, except the v corner (imax, jmax+1) of the
// reference set, handled by the caller). These give the same result as the sequential
// loops and let every boundary cell be produced by an independent thread.
// ---------------------------------------------------------------------------------------

// u, reference set (boundary.f90:14-16, 24-26, 29-32, 39-42)
template <typename F>
__device__ __forceinline__ double bc_u_ref(const Geo &g, const F *u, int i, int j)
{
    const int im = g.imax, jm = g.jmax;
    if (i <= 2 || i >= im) return 0.0;
    if (j == 1) return 0.0;
    return u[iu(g, i, jm - 1)];                 // j == jmax
}
// u, channel set: left inflow u_in, right zero-gradient, no-slip walls
template <typename F>
__device__ __forceinline__ double bc_u_chan(const Geo &g, const cfd_params &P, const F *u, int i, int j)
{
    const int im = g.imax, jm = g.jmax;
    if (i <= 2) return P.u_in;
    if (i >= im) return (j == 1 || j == jm) ? 0.0 : u[iu(g, im - 1, j)];
    return 0.0;                                  // j == 1 or j == jmax, 3 <= i <= imax-1
}
// v, reference set (boundary.f90:9-12, 19-22, 34-36, 44-46); (imax, jmax+1) excluded.
template <typename F>
__device__ __forceinline__ double bc_v_top_ref(const Geo &g, const F *v, int i)
{
    return 2.0 * (double)v[iv(g, i, g.jmax - 2)] - (double)v[iv(g, i, g.jmax - 1)];
}
template <typename F>
__device__ __forceinline__ double bc_v_ref(const Geo &g, const cfd_params &P, const F *v, int i, int j)
{
    const int im = g.imax, jm = g.jmax;
    if (j <= 2) return P.v_i;
    if (j == jm) {
        const int ii = i == 1 ? 2 : (i == im ? im - 1 : i);
        return bc_v_top_ref(g, v, ii);
    }
    if (j == jm + 1) {                           // i <= imax-1 here
        const int ii = i == 1 ? 2 : i;
        return 2.0 * (double)v[iv(g, ii, jm - 1)] - bc_v_top_ref(g, v, ii);
    }
    return i == 1 ? v[iv(g, 2, j)] : v[iv(g, im - 1, j)];
}
template <typename F>
__device__ __forceinline__ double bc_v_chan(const Geo &g, const F *v, int i, int j)
{
    const int im = g.imax, jm = g.jmax;
    if (j <= 2 || j >= jm || i == 1) return 0.0;
    return v[iv(g, im - 1, j)];                  // i == imax, 3 <= j <= jmax-1
}
// p (kind 0), T (kind 1), C (kind 2); reference set boundary.f90:52-159, channel set as the oracle.
template <typename F>
__device__ __forceinline__ double bc_s_ref(const Geo &g, const cfd_params &P, int kind, const F *a, int i, int j)
{
    const int im = g.imax, jm = g.jmax;
    const int ii = i == 1 ? 2 : (i == im ? im - 1 : i);
    if (j == 1) return kind == 0 ? a[iu(g, ii, 2)] : (kind == 1 ? P.tinf : 0.0);
    if (j == jm) return (double)a[iu(g, ii, jm - 1)] + 1.0 * ((double)a[iu(g, ii, jm - 1)] - (double)a[iu(g, ii, jm - 2)]);
    return a[iu(g, ii, j)];
}
__device__ __forceinline__ double slot_value(const cfd_params &P, double yj)
{
    return (yj >= P.slot_lo && yj <= P.slot_hi) ? P.slot_value : 0.0;
}
template <typename F>
__device__ __forceinline__ double bc_s_chan(const Geo &g, const cfd_params &P, int kind, const F *a, int i, int j)
{
    const int im = g.imax, jm = g.jmax;
    if (kind == 0) {
        if (i == im) return 0.0;
        const int ii = i == 1 ? 2 : i;
        const int jj = j == 1 ? 2 : (j == jm ? jm - 1 : j);
        return a[iu(g, ii, jj)];
    }
    if (kind == 1) {
        if (j == 1) return P.t_bottom;
        if (j == jm) return P.t_top;
        return a[iu(g, i == 1 ? 2 : im - 1, j)];
    }
    if (i == 1) return slot_value(P, g.y[j]);
    const int ii = i == im ? im - 1 : i;
    const int jj = j == 1 ? 2 : (j == jm ? jm - 1 : j);
    return a[iu(g, ii, jj)];
}

}  // namespace cfd
