// cfd_fused.hip — the fused pseudo-step: two row-streaming kernels and their support.
//
// One pseudo-step (main.f90:269-279: solve_U, solve_V, solve_P, solve_Z, solve_C) runs as
//   K1  k_uvp    the three SSP-RK3 stages of u and v and the pressure update,
//   K2  k_scal   the three stages of T and of C,
//   k_materialize  the boundary cells of the five outputs (the kernels themselves read
//                  boundary values of in-flight stages through the BC formulas).
//
// Decomposition. Lanes walk j (the contiguous axis): every global access of a wavefront is
// one coalesced row segment. A workgroup owns a column strip and a block of rows and marches
// down its rows; stage k of a chain is computed a fixed number of rows behind stage k-1
// (QUICK reaches two rows), so all three RK stages and the pressure update of a row happen
// in one pass over the inputs. i-direction neighbours come from per-lane register windows
// (Win<N>), j-direction neighbours of the centre row from small LDS rings. Each input row is
// read from HBM once per pass (halo columns/rows of neighbouring blocks hit L2/MALL).
//
// Numerics: flux-form QUICK (face value 1/2 (U+D) - 1/8 curvature at the upstream cell,
// which equals the reference's coefficient assembly of equations.f90:51-72), reciprocals of
// the 1-D spacings precomputed, FMA contraction allowed: results differ from the literal
// reference arithmetic by rounding only (tested at 1e-10 relative).
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

// Sliding window of the last N rows of one lane's column; [0] is the newest row.
template <int N>
struct Win {
    double v[N];
    __device__ __forceinline__ void fill(double x)
    {
#pragma unroll
        for (int k = 0; k < N; ++k) v[k] = x;
    }
    __device__ __forceinline__ void push(double x)
    {
#pragma unroll
        for (int k = 0; k < N - 1; ++k) v[k] = v[k + 1];
        v[N - 1] = x;
    }
    __device__ __forceinline__ double operator[](int back) const { return v[N - 1 - back]; }
    __device__ __forceinline__ void set(int back, double x) { v[N - 1 - back] = x; }
};

__device__ __forceinline__ int s2(int r) { return (r + 24) & 1; }
__device__ __forceinline__ int s3(int r) { return (r + 24) % 3; }
__device__ __forceinline__ int s4(int r) { return (r + 24) & 3; }

__device__ __forceinline__ unsigned long long dbits(double v)
{
    return (unsigned long long)__double_as_longlong(fabs(v));
}

__device__ __forceinline__ void fold_max(unsigned long long *out, unsigned long long m)
{
    for (int o = 32; o > 0; o >>= 1) {
        const unsigned long long other = __shfl_xor(m, o, 64);
        m = other > m ? other : m;
    }
    if ((threadIdx.x & 63) == 0 && m) atomicMax(out, m);
}

// QUICK face value in flux form: f > 0 takes the side-a upstream curvature.
__device__ __forceinline__ double quick_face(double f, double s, double ca, double cb)
{
    return 0.5 * s - 0.125 * (f > 0.0 ? ca : cb);
}

// ---------------------------------------------------------------------------------------
// K1: u, v (three RK stages each) and p.
// ---------------------------------------------------------------------------------------
constexpr int NT1 = 256;
constexpr int HL1 = 6, HR1 = 7;                 // column halo of the U/V/P chain
constexpr int W1 = NT1 - HL1 - HR1;             // output columns per workgroup
constexpr int PAD = 3;                          // LDS padding for j-3 .. j+2 reads
constexpr int LROW1 = NT1 + 2 * PAD;

struct K1Args {
    const double *um, *vm, *ut, *vt, *p, *T;    // inputs
    double *un, *vn, *pn;                       // outputs (interior cells)
    unsigned long long *res;                    // max |r3| bit patterns: u, v, p
    const double *eps, *lp;
    const double *aun, *aue, *avn, *ave;        // face areas
    const double *rdx, *rdy, *rdxm, *rdym;      // reciprocal spacings
    int imax, jmax, nbi, nbj, hb;
    cfd_params P;
};

// u of the centre row at column j+dj, the reference BC formula on boundary columns
// (boundary.f90:14-16,24-26 / channel: no-slip walls). Centre rows are interior rows.
__device__ __forceinline__ double lds_u(const double *row, int x, int j, int dj, int jm, bool chan)
{
    const int jj = j + dj;
    if (jj == 1) return 0.0;
    if (jj == jm) return chan ? 0.0 : row[PAD + x + dj - 1];
    return row[PAD + x + dj];
}

// v of the centre row at column j+dj (boundary.f90:9-12,19-22 / channel walls).
__device__ __forceinline__ double lds_v(const double *row, int x, int j, int dj, int jm, bool chan, double vi)
{
    const int jj = j + dj;
    if (jj <= 2) return chan ? 0.0 : vi;
    if (jj == jm) return chan ? 0.0 : 2.0 * row[PAD + x + dj - 2] - row[PAD + x + dj - 1];
    if (jj == jm + 1) {
        if (chan) return 0.0;
        const double a = row[PAD + x + dj - 3], b = row[PAD + x + dj - 2];
        return 2.0 * b - (2.0 * a - b);
    }
    return row[PAD + x + dj];
}

template <bool STEADY>
__global__ __launch_bounds__(NT1, 1) void k_uvp(K1Args a)
{
    const int G = a.nbi * a.nbj, per = (G + 7) / 8;
    const int b = blockIdx.x, L = (b % 8) * per + b / 8;   // blocks of one XCD are neighbours
    if (L >= G) return;
    const int bi = L / a.nbj, bj = L % a.nbj;
    const int im = a.imax, jm = a.jmax;
    const int ilo = 2 + bi * a.hb, ihi = min(ilo + a.hb, im);        // output rows [ilo, ihi)
    const int jlo = 2 + bj * W1, jhi = min(jlo + W1, jm);            // output cols [jlo, jhi)
    const int x = threadIdx.x;
    const int j = jlo - HL1 + x;
    const cfd_params &P = a.P;
    const bool chan = P.bc == CFD_BC_CHANNEL;
    const bool upwind_all = P.scheme == CFD_SCHEME_UPWIND;
    const bool mu_poly = P.mu_model == CFD_MU_POLY4;
    const bool bouss = P.buoyancy == CFD_BUOY_BOUSSINESQ;
    const double dt = P.dt, dtau = P.dtau, vi = P.v_i;
    const double inv_re = 1.0 / P.re, inv_reda = 1.0 / (P.re * P.darcy);

    __shared__ double sUT[4][LROW1], sVT[4][LROW1], sP[2][LROW1], sT[2][LROW1];
    __shared__ double sU1[3][LROW1], sU2[3][LROW1], sV1[3][LROW1], sV2[3][LROW1], sV3[2][LROW1];
    for (int k = x; k < LROW1; k += NT1) {
        for (int q = 0; q < 4; ++q) { sUT[q][k] = 0.0; sVT[q][k] = 0.0; }
        for (int q = 0; q < 3; ++q) { sU1[q][k] = 0.0; sU2[q][k] = 0.0; sV1[q][k] = 0.0; sV2[q][k] = 0.0; }
        for (int q = 0; q < 2; ++q) { sP[q][k] = 0.0; sT[q][k] = 1.0; sV3[q][k] = 0.0; }
    }

    // per-column geometry (clamped indices; values outside the ranges are never used)
    const bool jvalid = j >= 1 && j <= jm + 1;
    const int jc = min(max(j, 2), jm);
    const double aue_j = a.aue[min(max(j, 1), jm)];
    const double ave_j = a.ave[jc];
    const double rdy_j = a.rdy[jc], rdy_n = a.rdy[min(jc + 1, jm)];
    const double rdym_j = a.rdym[jc], rdym_n = a.rdym[min(jc + 1, jm + 1)];
    const bool ucol = j >= 2 && j <= jm - 1;           // u / p interior column
    const bool vcol = j >= 3 && j <= jm - 1;           // v interior column
    const bool u_upw_col = upwind_all || j == 2 || j == jm - 1;
    const bool out_col = j >= jlo && j < jhi;

    Win<7> ut, Su;        // u_tau rows t-6..t ; Su = u_tau(r,j) + u_tau(r,j-1)
    Win<8> vt;            // v_tau rows t-7..t
    Win<6> Svn;           // v_tau(r,j+1) + v_tau(r-1,j+1), rows t-6..t-1
    Win<6> pw;            // p rows t-7..t-2
    Win<5> ie, ew, nuw, k1w, k2w, dpj, bw, du, dv;   // per-cell data, rows t-6..t-2
    Win<5> w1, w2, x1, x2;                            // u1, u2, v1, v2 own windows
    Win<2> w3, x3;                                    // u3, v3
    ut.fill(0.0); Su.fill(0.0); vt.fill(0.0); Svn.fill(0.0); pw.fill(0.0);
    ie.fill(1.0); ew.fill(1.0); nuw.fill(0.0); k1w.fill(0.0); k2w.fill(0.0); dpj.fill(0.0); bw.fill(0.0);
    du.fill(0.0); dv.fill(0.0);
    w1.fill(0.0); w2.fill(0.0); x1.fill(0.0); x2.fill(0.0); w3.fill(0.0); x3.fill(0.0);
    unsigned long long mu_ = 0, mv_ = 0, mp_ = 0;
    __syncthreads();

    const size_t su = (size_t)(jm + 1), sv = (size_t)(jm + 2);
    for (int t = ilo - 6; t <= ihi + 6; ++t) {
        // ---------------- phase A: rows t (iterates) and t-2 (pointwise inputs)
        double ut_new = 0.0, vt_new = 0.0;
        if (jvalid && t >= 1) {
            if (t <= im + 1 && j <= jm) ut_new = a.ut[(size_t)t * su + j];
            if (t <= im) vt_new = a.vt[(size_t)t * sv + j];
        }
        const int rp = t - 2;
        double p_rp = 0.0, T_rp = 1.0, e = 1.0, Lp = 0.0, um_rp = 0.0, vm_rp = 0.0;
        if (jvalid && rp >= 1 && rp <= im) {
            if (j <= jm) {
                const size_t k = (size_t)rp * su + j;
                p_rp = a.p[k]; T_rp = a.T[k]; e = a.eps[k]; Lp = a.lp[k];
                if (!STEADY) um_rp = a.um[k];
            }
            if (!STEADY) vm_rp = a.vm[(size_t)rp * sv + j];
        }
        ut.push(ut_new);
        vt.push(vt_new);
        sUT[s4(t)][PAD + x] = ut_new;
        sVT[s4(t)][PAD + x] = vt_new;
        sP[s2(rp)][PAD + x] = p_rp;
        sT[s2(rp)][PAD + x] = T_rp;
        __syncthreads();

        // ---------------- phase B: per-cell data of row t-2, neighbour sums
        Su.push(ut_new + sUT[s4(t)][PAD + x - 1]);
        Svn.push(sVT[s4(t - 1)][PAD + x + 1] + sVT[s4(t - 2)][PAD + x + 1]);
        pw.push(p_rp);
        {
            const double iev = 1.0 / e;
            double nu = e * inv_re;
            if (mu_poly) {
                const double tt = T_rp;
                nu *= (((P.mu_coef[4] * tt + P.mu_coef[3]) * tt + P.mu_coef[2]) * tt + P.mu_coef[1]) * tt + P.mu_coef[0];
            }
            double k1 = 0.0, k2 = 0.0;
            if (Lp != 0.0) {
                k1 = e * Lp * inv_reda;
                k2 = e * Lp * P.cf / sqrt(e * P.darcy);
            }
            ie.push(iev); ew.push(e); nuw.push(nu); k1w.push(k1); k2w.push(k2);
            dpj.push(p_rp - sP[s2(rp)][PAD + x - 1]);
            const double Tm = sT[s2(rp)][PAD + x - 1];
            bw.push(bouss ? P.ri * ((T_rp + Tm) * 0.5 - P.t_ref) : P.invfr2 * (1.0 - 1.0 / ((T_rp + Tm) * 0.5)));
            if (!STEADY) { du.push(um_rp - ut[2]); dv.push(vm_rp - vt[2]); }
        }

        // ---------------- U stage k at row r = t - 2k (u update: i = 3..imax-1, j = 2..jmax-1)
#define U_STAGE(K, WP, RING_C, WOUT, RING_OUT)                                                   \
        {                                                                                         \
            const int r = t - 2 * K;                                                              \
            const int bk = 2 * K - 2; /* back index of row r in the rows t-6..t-2 windows */      \
            double val = 0.0, rs = 0.0;                                                           \
            bool interior = false;                                                                \
            if (r >= 3 && r <= im - 1) {                                                          \
                if (ucol) {                                                                       \
                    interior = true;                                                              \
                    const double uP = WP[2], uW = WP[3], uWW = WP[4], uE = WP[1], uEE = WP[0];    \
                    const double *c = RING_C;                                                     \
                    const double uS = lds_u(c, x, j, -1, jm, chan), uN = lds_u(c, x, j, 1, jm, chan); \
                    const double iev = ie[bk], nu = nuw[bk], ev = ew[bk];                         \
                    const double aun_r = a.aun[r], rdx_r = a.rdx[r];                              \
                    const double cns = 0.5 * aun_r * iev, cew = 0.5 * aue_j * iev;                \
                    const double sE = uE + uP, sW = uP + uW, sS = uS + uP, sN = uP + uN;          \
                    const double fn = Svn[2 * K - 1] * cns;                                       \
                    const double fs = (vt[2 * K] + vt[2 * K + 1]) * cns;                          \
                    const double fe = sE * cew, fw = sW * cew;                                    \
                    const double nA = nu * aun_r, nB = nu * aue_j;                                \
                    const double diff = nB * a.rdxm[r] * (uW - uP) + nB * a.rdxm[r + 1] * (uE - uP) \
                                      + nA * rdy_j * (uS - uP) + nA * rdy_n * (uN - uP);          \
                    double pw_, pe_, ps_, pn_;                                                    \
                    if (u_upw_col) {                                                              \
                        pw_ = fw > 0.0 ? uW : uP; pe_ = fe > 0.0 ? uP : uE;                       \
                        ps_ = fs > 0.0 ? uS : uP; pn_ = fn > 0.0 ? uP : uN;                       \
                    } else {                                                                      \
                        const double uSS = lds_u(c, x, j, -2, jm, chan), uNN = lds_u(c, x, j, 2, jm, chan); \
                        const double cW = uWW - 2.0 * uW + uP, cPi = uW - 2.0 * uP + uE, cE = uP - 2.0 * uE + uEE; \
                        const double cS = uSS - 2.0 * uS + uP, cPj = uS - 2.0 * uP + uN, cN = uP - 2.0 * uN + uNN; \
                        pw_ = quick_face(fw, sW, cW, cPi); pe_ = quick_face(fe, sE, cPi, cE);     \
                        ps_ = quick_face(fs, sS, cS, cPj); pn_ = quick_face(fn, sN, cPj, cN);     \
                    }                                                                             \
                    const double conv = fw * pw_ - fe * pe_ + fs * ps_ - fn * pn_;                \
                    double R = rdx_r * rdy_j * (diff + conv) - ev * (pw[bk] - pw[bk + 1]) * rdx_r - P.g * ev; \
                    const double k2 = k2w[bk];                                                    \
                    if (k2 != 0.0) {                                                              \
                        const double vP = vt[2 * K];                                              \
                        R -= uP * (k1w[bk] + k2 * sqrt(uP * uP + vP * vP));                       \
                    }                                                                             \
                    rs = ((STEADY ? 0.0 : du[bk]) + R * dt) * dtau;                               \
                    const double tau = ut[2 * K];                                                 \
                    if (K == 1) val = tau + rs;                                                   \
                    else if (K == 2) val = 0.75 * tau + 0.25 * (uP + rs);                         \
                    else val = 1.0 / 3.0 * tau + 2.0 / 3.0 * (uP + rs);                          \
                }                                                                                 \
            } else if (r <= 2) {                                                                  \
                val = chan ? P.u_in : 0.0;                                                        \
            } else {                                                                              \
                val = (chan && ucol) ? WOUT[0] : 0.0;                                             \
            }                                                                                     \
            WOUT.push(val);                                                                       \
            RING_OUT                                                                              \
            if (K == 3 && interior && out_col && r >= ilo && r < ihi) {                           \
                a.un[(size_t)r * su + j] = val;                                                   \
                const unsigned long long bb = dbits(rs); mu_ = bb > mu_ ? bb : mu_;               \
            }                                                                                     \
        }

        // ---------------- V stage k at row r = t - 2k (v update: i = 2..imax-1, j = 3..jmax-1)
#define V_STAGE(K, WP, RING_C, WOUT, RING_OUT)                                                   \
        {                                                                                         \
            const int r = t - 2 * K;                                                              \
            const int bk = 2 * K - 2;                                                             \
            double val = 0.0, rs = 0.0;                                                           \
            bool interior = false;                                                                \
            if (j == 2) {                                                                         \
                val = chan ? 0.0 : vi;                                                            \
            } else if (r >= 2 && r <= im - 1) {                                                   \
                if (vcol) {                                                                       \
                    interior = true;                                                              \
                    const double vP = WP[2], vW = WP[3], vWW = WP[4], vE = WP[1], vEE = WP[0];    \
                    const double *c = RING_C;                                                     \
                    const double vS = lds_v(c, x, j, -1, jm, chan, vi), vN = lds_v(c, x, j, 1, jm, chan, vi); \
                    const double iev = ie[bk], nu = nuw[bk], ev = ew[bk];                         \
                    const double avn_r = a.avn[r], rdx_r = a.rdx[r];                              \
                    const double cns = 0.5 * avn_r * iev, cew = 0.5 * ave_j * iev;                \
                    const double sN = vP + vN, sS = vP + vS, sE = vE + vP, sW = vP + vW;          \
                    const double fn = sN * cns, fs = sS * cns;                                    \
                    const double fe = Su[2 * K - 1] * cew, fw = Su[2 * K] * cew;                  \
                    const double nA = nu * avn_r, nB = nu * ave_j;                                \
                    const double diff = nB * rdx_r * (vW - vP) + nB * a.rdx[r + 1] * (vE - vP)    \
                                      + nA * rdym_j * (vS - vP) + nA * rdym_n * (vN - vP);        \
                    double pw_, pe_, ps_, pn_;                                                    \
                    if (upwind_all || r == 2 || r == im - 1) {                                    \
                        pw_ = fw > 0.0 ? vW : vP; pe_ = fe > 0.0 ? vP : vE;                       \
                        ps_ = fs > 0.0 ? vS : vP; pn_ = fn > 0.0 ? vP : vN;                       \
                    } else {                                                                      \
                        const double vSS = lds_v(c, x, j, -2, jm, chan, vi), vNN = lds_v(c, x, j, 2, jm, chan, vi); \
                        const double cW = vWW - 2.0 * vW + vP, cPi = vW - 2.0 * vP + vE, cE = vP - 2.0 * vE + vEE; \
                        const double cS = vSS - 2.0 * vS + vP, cPj = vS - 2.0 * vP + vN, cN = vP - 2.0 * vN + vNN; \
                        pw_ = quick_face(fw, sW, cW, cPi); pe_ = quick_face(fe, sE, cPi, cE);     \
                        ps_ = quick_face(fs, sS, cS, cPj); pn_ = quick_face(fn, sN, cPj, cN);     \
                    }                                                                             \
                    const double conv = fw * pw_ - fe * pe_ + fs * ps_ - fn * pn_;                \
                    double R = rdx_r * rdy_j * (diff + conv) - ev * dpj[bk] * rdy_j + bw[bk];     \
                    const double k2 = k2w[bk];                                                    \
                    if (k2 != 0.0) {                                                              \
                        const double uP = ut[2 * K];                                              \
                        R -= vP * (k1w[bk] + k2 * sqrt(uP * uP + vP * vP));                       \
                    }                                                                             \
                    rs = ((STEADY ? 0.0 : dv[bk]) + R * dt) * dtau;                               \
                    const double tau = vt[2 * K];                                                 \
                    if (K == 1) val = tau + rs;                                                   \
                    else if (K == 2) val = 0.75 * tau + 0.25 * (vP + rs);                         \
                    else val = 1.0 / 3.0 * tau + 2.0 / 3.0 * (vP + rs);                           \
                    if (r == 2 && !chan) WOUT.set(0, val);   /* row 1 = row 2 (boundary.f90:34-36) */ \
                }                                                                                 \
            } else if (r == im) {                                                                 \
                val = vcol ? WOUT[0] : 0.0;                                                       \
            }                                                                                     \
            WOUT.push(val);                                                                       \
            RING_OUT                                                                              \
            if (K == 3 && interior && out_col && r >= ilo && r < ihi) {                           \
                a.vn[(size_t)r * sv + j] = val;                                                   \
                const unsigned long long bb = dbits(rs); mv_ = bb > mv_ ? bb : mv_;               \
            }                                                                                     \
        }

        U_STAGE(1, ut, sUT[s4(t - 2)], w1, { sU1[s3(t - 2)][PAD + x] = val; })
        V_STAGE(1, vt, sVT[s4(t - 2)], x1, { sV1[s3(t - 2)][PAD + x] = val; })
        U_STAGE(2, w1, sU1[s3(t - 4)], w2, { sU2[s3(t - 4)][PAD + x] = val; })
        V_STAGE(2, x1, sV1[s3(t - 4)], x2, { sV2[s3(t - 4)][PAD + x] = val; })
        U_STAGE(3, w2, sU2[s3(t - 6)], w3, {})
        V_STAGE(3, x2, sV2[s3(t - 6)], x3, { sV3[s2(t - 6)][PAD + x] = val; })
#undef U_STAGE
#undef V_STAGE

        // ---------------- P at row r = t - 7 (equations.f90:620-658)
        {
            const int r = t - 7;
            if (r >= 2 && r <= im - 1 && ucol && out_col && r >= ilo && r < ihi) {
                const double u3p = w3[1], u3e = w3[0];
                const double v3p = x3[1];
                const double v3n = lds_v(sV3[s2(r)], x, j, 1, jm, chan, vi);
                const double avn_r = a.avn[r];
                const double rpv = -((u3e * aue_j - u3p * aue_j) + (v3n * avn_r - v3p * avn_r));
                const double dd = dtau * rpv * P.beta;
                const double pr = pw[5];
                double pi = pr + dd;
                pi = 0.75 * pr + 0.25 * (pi + dd);
                a.pn[(size_t)r * su + j] = 1.0 / 3.0 * pr + 2.0 / 3.0 * (pi + dd);
                const unsigned long long bb = dbits(dd); mp_ = bb > mp_ ? bb : mp_;
            }
        }
    }
    fold_max(a.res + 0, mu_);
    fold_max(a.res + 1, mv_);
    fold_max(a.res + 2, mp_);
}

// ---------------------------------------------------------------------------------------
// K2: T (energy, 1/Pe) and C (species, 1/(Re Sc)), three RK stages each.
// ---------------------------------------------------------------------------------------
constexpr int NT2 = 256;
constexpr int HL2 = 3, HR2 = 3;
constexpr int W2 = NT2 - HL2 - HR2;
constexpr int LROW2 = NT2 + 2 * PAD;

struct K2Args {
    const double *un, *vn, *T, *Tt, *C, *Ct;
    double *Tn, *Cn;
    const double *eps, *lp, *y;
    const double *aue, *avn, *rdx, *rdy, *rdxm, *rdym;
    int imax, jmax, nbi, nbj, hb;
    cfd_params P;
};

// T of the centre row at column j+dj (boundary.f90:85-88 / channel wall temperatures).
__device__ __forceinline__ double lds_t(const double *row, int x, int j, int dj, int jm, bool chan, const cfd_params &P)
{
    const int jj = j + dj;
    if (jj == 1) return chan ? P.t_bottom : P.tinf;
    if (jj == jm) {
        if (chan) return P.t_top;
        const double a1 = row[PAD + x + dj - 1], a2 = row[PAD + x + dj - 2];
        return a1 + 1.0 * (a1 - a2);
    }
    return row[PAD + x + dj];
}

// C of the centre row (boundary.f90:128-131 / channel: zero-gradient walls).
__device__ __forceinline__ double lds_c(const double *row, int x, int j, int dj, int jm, bool chan)
{
    const int jj = j + dj;
    if (jj == 1) return chan ? row[PAD + x + dj + 1] : 0.0;
    if (jj == jm) {
        if (chan) return row[PAD + x + dj - 1];
        const double a1 = row[PAD + x + dj - 1], a2 = row[PAD + x + dj - 2];
        return a1 + 1.0 * (a1 - a2);
    }
    return row[PAD + x + dj];
}

template <bool STEADY>
__global__ __launch_bounds__(NT2, 1) void k_scal(K2Args a)
{
    const int G = a.nbi * a.nbj, per = (G + 7) / 8;
    const int b = blockIdx.x, L = (b % 8) * per + b / 8;
    if (L >= G) return;
    const int bi = L / a.nbj, bj = L % a.nbj;
    const int im = a.imax, jm = a.jmax;
    const int ilo = 2 + bi * a.hb, ihi = min(ilo + a.hb, im);
    const int jlo = 2 + bj * W2, jhi = min(jlo + W2, jm);
    const int x = threadIdx.x;
    const int j = jlo - HL2 + x;
    const cfd_params &P = a.P;
    const bool chan = P.bc == CFD_BC_CHANNEL;
    const double dt = P.dt, dtau = P.dtau;
    const double kt = 1.0 / P.pe, kc = 1.0 / P.re / P.sc;

    __shared__ double sTT[3][LROW2], sCT[3][LROW2], sV[2][LROW2];
    __shared__ double sZ1[2][LROW2], sZ2[2][LROW2], sC1[2][LROW2], sC2[2][LROW2];
    for (int k = x; k < LROW2; k += NT2) {
        for (int q = 0; q < 3; ++q) { sTT[q][k] = 0.0; sCT[q][k] = 0.0; }
        for (int q = 0; q < 2; ++q) { sV[q][k] = 0.0; sZ1[q][k] = 0.0; sZ2[q][k] = 0.0; sC1[q][k] = 0.0; sC2[q][k] = 0.0; }
    }
    const bool jvalid = j >= 1 && j <= jm;
    const int jc = min(max(j, 2), jm - 1);
    const double aue_j = a.aue[min(max(j, 1), jm)];
    const double rdy_j = a.rdy[jc], rdy_n = a.rdy[jc + 1];
    const double rdym_n = a.rdym[jc + 1];
    const bool scol = j >= 2 && j <= jm - 1;
    const bool out_col = j >= jlo && j < jhi;
    const double cslot = chan ? slot_value(P, a.y[min(max(j, 1), jm)]) : 0.0;

    Win<4> tt, ct, uw, vw, vnw, id, oml, dT, dC;
    Win<3> z1, z2, c1, c2;
    tt.fill(0.0); ct.fill(0.0); uw.fill(0.0); vw.fill(0.0); vnw.fill(0.0); id.fill(1.0); oml.fill(1.0);
    dT.fill(0.0); dC.fill(0.0); z1.fill(0.0); z2.fill(0.0); c1.fill(0.0); c2.fill(0.0);
    __syncthreads();

    const size_t su = (size_t)(jm + 1), sv = (size_t)(jm + 2);
    for (int t = ilo - 3; t <= ihi + 2; ++t) {
        // ---------------- phase A: row t
        double tnew = 0.0, cnew = 0.0, unew = 0.0, vnew = 0.0, Tp = 0.0, Cp = 0.0, e = 1.0, Lp = 0.0;
        if (jvalid && t >= 1 && t <= im + 1) {
            if (t <= im) {
                const size_t k = (size_t)t * su + j;
                tnew = a.Tt[k]; cnew = a.Ct[k]; e = a.eps[k]; Lp = a.lp[k];
                if (!STEADY) { Tp = a.T[k]; Cp = a.C[k]; }
                // v3: interior rows/columns as written by K1, boundary column 2 from its BC
                if (j == 2) vnew = chan ? 0.0 : P.v_i;
                else if (t >= 2 && t <= im - 1 && j <= jm - 1) vnew = a.vn[(size_t)t * sv + j];
            }
            // u3: interior rows as written by K1, boundary rows from the BC formulas
            if (t <= 2) unew = chan ? P.u_in : 0.0;
            else if (t >= im) unew = chan ? uw[0] : 0.0;
            else unew = a.un[(size_t)t * su + j];
        }
        tt.push(tnew); ct.push(cnew); uw.push(unew); vw.push(vnew);
        sTT[s3(t)][PAD + x] = tnew;
        sCT[s3(t)][PAD + x] = cnew;
        sV[s2(t)][PAD + x] = vnew;
        __syncthreads();

        // ---------------- phase B: per-cell data of row t
        {
            const int jn = j + 1;
            double vn_ = sV[s2(t)][PAD + x + 1];
            if (jn == jm) vn_ = chan ? 0.0 : 2.0 * sV[s2(t)][PAD + x - 1] - sV[s2(t)][PAD + x];
            vnw.push(vn_);
            id.push(Lp != 0.0 ? 1.0 / (Lp * (e - 1.0) + 1.0) : 1.0);
            oml.push(1.0 - Lp);
            if (!STEADY) { dT.push(Tp - tnew); dC.push(Cp - cnew); }
        }


        // stage k of a scalar at row r = t - k; newest own row of WP is t-k+1
#define SCAL_STAGE(K, WP, CROW, ISC, WOUT, RINGW, OUTP, KAPPA, DW, TAUW)                         \
        {                                                                                         \
            const int r = t - K;                                                                  \
            const int bq = K;            /* back index of row r in windows whose newest is t */   \
            double val = 0.0;                                                                     \
            bool interior = false;                                                                \
            if (r >= 2 && r <= im - 1) {                                                          \
                if (scol) {                                                                       \
                    interior = true;                                                              \
                    const double zE = WP[0], zP = WP[1], zW = WP[2];                              \
                    const double *c = CROW;                                                       \
                    const double zS = ISC ? lds_c(c, x, j, -1, jm, chan) : lds_t(c, x, j, -1, jm, chan, P); \
                    const double zN = ISC ? lds_c(c, x, j, 1, jm, chan) : lds_t(c, x, j, 1, jm, chan, P);   \
                    const double avn_r = a.avn[r];                                                \
                    const double kA = KAPPA * aue_j, kB = KAPPA * avn_r;                          \
                    const double diff = kA * a.rdx[r + 1] * (zE - zP) + kA * a.rdx[r] * (zW - zP) \
                                      + kB * rdy_n * (zN - zP) + kB * rdy_j * (zS - zP);          \
                    const double uE = uw[bq - 1], uP = uw[bq], vP = vw[bq], vN = vnw[bq];         \
                    const double conv = 0.5 * ((zE + zP) * uE - (zW + zP) * uP) * aue_j           \
                                      + 0.5 * ((zN + zP) * vN - (zS + zP) * vP) * avn_r;          \
                    const double vols = a.rdxm[r + 1] * rdym_n;                                   \
                    const double R = ISC ? vols * (diff - conv * id[bq])                          \
                                         : vols * (diff - oml[bq] * conv) * id[bq];               \
                    const double rs = ((STEADY ? 0.0 : DW[bq]) + R * dt) * dtau;                  \
                    const double tau = TAUW[bq];                                                  \
                    if (K == 1) val = tau + rs;                                                   \
                    else if (K == 2) val = 0.75 * tau + 0.25 * (zP + rs);                         \
                    else val = 1.0 / 3.0 * tau + 2.0 / 3.0 * (zP + rs);                          \
                    if (r == 2) WOUT.set(0, (ISC && chan) ? cslot : val);  /* row 1 */            \
                }                                                                                 \
            } else if (r == im) {                                                                 \
                val = scol ? WOUT[0] : 0.0;                                                       \
            } else if (r == 1) {                                                                  \
                val = (ISC && chan) ? cslot : 0.0;                                                \
            }                                                                                     \
            WOUT.push(val);                                                                       \
            RINGW                                                                                 \
            if (OUTP != nullptr && interior && out_col && r >= ilo && r < ihi)                    \
                OUTP[(size_t)r * su + j] = val;                                                   \
        }

        double *const noout = nullptr;
        SCAL_STAGE(1, tt, sTT[s3(t - 1)], false, z1, { sZ1[s2(t - 1)][PAD + x] = val; }, noout, kt, dT, tt)
        SCAL_STAGE(1, ct, sCT[s3(t - 1)], true, c1, { sC1[s2(t - 1)][PAD + x] = val; }, noout, kc, dC, ct)
        SCAL_STAGE(2, z1, sZ1[s2(t - 2)], false, z2, { sZ2[s2(t - 2)][PAD + x] = val; }, noout, kt, dT, tt)
        SCAL_STAGE(2, c1, sC1[s2(t - 2)], true, c2, { sC2[s2(t - 2)][PAD + x] = val; }, noout, kc, dC, ct)
        {
            Win<3> z3dummy, c3dummy;
            z3dummy.fill(0.0); c3dummy.fill(0.0);
            SCAL_STAGE(3, z2, sZ2[s2(t - 3)], false, z3dummy, {}, a.Tn, kt, dT, tt)
            SCAL_STAGE(3, c2, sC2[s2(t - 3)], true, c3dummy, {}, a.Cn, kc, dC, ct)
        }
#undef SCAL_STAGE
    }
}
}  // namespace

extern "C" {

int cfdk_fused_available(void) { return 1; }

int cfdk_factors(const cfd_devgeo *g, cfd_devfactors *f, hipStream_t s);

int cfdk_fused_step(const cfd_devgeo *g, const cfd_devfactors *f, const cfd_params *P,
                    const cfd_fused_args *a, int kernel, hipStream_t s)
{
    (void)kernel;
    const int im = g->imax, jm = g->jmax;
    if (P->b_art != 0.0 || P->compat != CFD_COMPAT_FORTRAN) return CFD_ERR_UNSUPPORTED;
    // K1
    K1Args k1;
    k1.um = a->um; k1.vm = a->vm; k1.ut = a->um_tau; k1.vt = a->vm_tau; k1.p = a->p; k1.T = a->T;
    k1.un = a->um_n; k1.vn = a->vm_n; k1.pn = a->pn; k1.res = (unsigned long long *)a->res_max;
    k1.eps = g->eps; k1.lp = g->lp;
    k1.aun = g->aun; k1.aue = g->aue; k1.avn = g->avn; k1.ave = g->ave;
    k1.rdx = f->rdx; k1.rdy = f->rdy; k1.rdxm = f->rdxm; k1.rdym = f->rdym;
    k1.imax = im; k1.jmax = jm; k1.P = *P;
    k1.nbj = cdiv(jm - 2, W1);
    {
        int nbi = 256 / k1.nbj;
        const int rows = im - 2;
        if (nbi < 1) nbi = 1;
        if (rows / nbi < 24) nbi = rows / 24 > 0 ? rows / 24 : 1;
        k1.hb = cdiv(rows, nbi);
        k1.nbi = cdiv(rows, k1.hb);
    }
    const int G1 = k1.nbi * k1.nbj, grid1 = ((G1 + 7) / 8) * 8;
    if (P->steady) hipLaunchKernelGGL(k_uvp<true>, dim3(grid1), dim3(NT1), 0, s, k1);
    else hipLaunchKernelGGL(k_uvp<false>, dim3(grid1), dim3(NT1), 0, s, k1);
    int r = launched();
    if (r != CFD_OK) return r;
    // K2
    K2Args k2;
    k2.un = a->um_n; k2.vn = a->vm_n; k2.T = a->T; k2.Tt = a->T_tau; k2.C = a->C; k2.Ct = a->C_tau;
    k2.Tn = a->T_n; k2.Cn = a->C_n;
    k2.eps = g->eps; k2.lp = g->lp; k2.y = g->y;
    k2.aue = g->aue; k2.avn = g->avn; k2.rdx = f->rdx; k2.rdy = f->rdy; k2.rdxm = f->rdxm; k2.rdym = f->rdym;
    k2.imax = im; k2.jmax = jm; k2.P = *P;
    k2.nbj = cdiv(jm - 2, W2);
    {
        int nbi = 256 / k2.nbj;
        const int rows = im - 2;
        if (nbi < 1) nbi = 1;
        if (rows / nbi < 16) nbi = rows / 16 > 0 ? rows / 16 : 1;
        k2.hb = cdiv(rows, nbi);
        k2.nbi = cdiv(rows, k2.hb);
    }
    const int G2 = k2.nbi * k2.nbj, grid2 = ((G2 + 7) / 8) * 8;
    if (P->steady) hipLaunchKernelGGL(k_scal<true>, dim3(grid2), dim3(NT2), 0, s, k2);
    else hipLaunchKernelGGL(k_scal<false>, dim3(grid2), dim3(NT2), 0, s, k2);
    r = launched();
    if (r != CFD_OK) return r;
    return cfdk_materialize_bc(g, P, a->um_n, a->vm_n, a->pn, a->T_n, a->C_n, a->vm_tau, s);
}

}  // extern "C"

// ---------------------------------------------------------------------------------------
// support launchers
// ---------------------------------------------------------------------------------------
namespace {
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
    const int nu_r = im + 1, nu_c = jm - 2;
    const int n_u = 2 * nu_r + 4 * nu_c;
    const int n_v = 4 * im + 2 * (jm - 3);
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

}  // extern "C"
