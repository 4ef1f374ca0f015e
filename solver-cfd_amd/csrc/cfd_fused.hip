// cfd_fused.hip — the fused pseudo-step: two row-streaming kernels and their support.
//
// One pseudo-step (main.f90:122-132: solve_U, solve_V, solve_P, solve_Z, solve_C) runs as
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

// Scalars are precomputed on the host so that few values occupy scalar registers.
struct K1Args {
    const double *um, *vm, *ut, *vt, *p, *T, *eps, *lp;   // inputs
    double *un, *vn, *pn;                                  // outputs (interior cells)
    unsigned long long *res;                               // max |r3| bit patterns: u, v, p
    const double *pk;                                      // packed geometry (cfd_devfactors.pack)
    int imax, jmax, nbi, nbj, hb, chan, upwind_all, bouss;
    double dt, dtau, beta, inv_re, inv_reda, cfsd, g, invfr2, ri, t_ref, v_i, u_in;
    double mu[5];
};

// v of the centre row at column j+dj with the BC formulas (boundary.f90:9-12,19-22 /
// channel walls): used where the ring's boundary columns are not yet fixed up.
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

// Reference-set boundary columns of a u ring row (u(jmax) = u(jmax-1), boundary.f90:24-26)
// and of a v ring row (top extrapolation, boundary.f90:19-22), written by the lanes that
// own those columns one iteration after the row was produced and before it is read.
__device__ __forceinline__ void fix_u(double *row, int x, int j, int jm)
{
    if (j == jm) row[PAD + x] = row[PAD + x - 1];
}
__device__ __forceinline__ void fix_v(double *row, int x, int j, int jm)
{
    if (j == jm) {
        row[PAD + x] = 2.0 * row[PAD + x - 2] - row[PAD + x - 1];
    } else if (j == jm + 1) {
        const double a = row[PAD + x - 3], b = row[PAD + x - 2];
        row[PAD + x] = 2.0 * b - (2.0 * a - b);
    }
}

// OCC: waves per SIMD the register budget is sized for (1: no spills; 2: 256 registers, the
// compiler spills the rest to scratch). The autotuner picks the faster one per grid.
template <bool STEADY, bool MUPOLY, int OCC>
__global__ __launch_bounds__(NT1, OCC) void k_uvp(const K1Args a)
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
    const bool chan = a.chan != 0, upwind_all = a.upwind_all != 0, bouss = a.bouss != 0;
    const double dt = a.dt, dtau = a.dtau, vi = a.v_i;
    const int R = im + 2, Cc = jm + 2;
    const double *pr = a.pk, *pc = a.pk + 4 * R;    // rows [aun|avn|rdx|rdxm], cols [aue|ave|rdy|rdym]

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
    const double haue_j = 0.5 * pc[min(max(j, 1), jm)];
    const double aue_j = pc[min(max(j, 1), jm)];
    const double ave_j = pc[Cc + jc], have_j = 0.5 * ave_j;
    const double rdy_j = pc[2 * Cc + jc], rdy_n = pc[2 * Cc + min(jc + 1, jm)];
    const double rdym_j = pc[3 * Cc + jc], rdym_n = pc[3 * Cc + min(jc + 1, jm + 1)];
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
    ie.fill(1.0); ew.fill(1.0); k1w.fill(0.0); k2w.fill(0.0); dpj.fill(0.0); bw.fill(0.0);
    if (MUPOLY) nuw.fill(0.0);
    if (!STEADY) { du.fill(0.0); dv.fill(0.0); }
    w1.fill(0.0); w2.fill(0.0); x1.fill(0.0); x2.fill(0.0); w3.fill(0.0); x3.fill(0.0);
    unsigned long long mu_ = 0, mv_ = 0, mp_ = 0;

    const size_t su = (size_t)(jm + 1), sv = (size_t)(jm + 2);
    // next-row registers: the loads for row t+1 are issued before row t is computed
    double q_ut, q_vt, q_p, q_T, q_e, q_L, q_um, q_vm;
    auto fetch = [&](int t) {
        q_ut = 0.0; q_vt = 0.0; q_p = 0.0; q_T = 1.0; q_e = 1.0; q_L = 0.0; q_um = 0.0; q_vm = 0.0;
        if (jvalid && t >= 1) {
            if (t <= im + 1 && j <= jm) q_ut = a.ut[(size_t)t * su + j];
            if (t <= im) q_vt = a.vt[(size_t)t * sv + j];
        }
        const int rp = t - 2;
        if (jvalid && rp >= 1 && rp <= im) {
            if (j <= jm) {
                const size_t k = (size_t)rp * su + j;
                q_p = a.p[k]; q_T = a.T[k]; q_e = a.eps[k]; q_L = a.lp[k];
                if (!STEADY) q_um = a.um[k];
            }
            if (!STEADY) q_vm = a.vm[(size_t)rp * sv + j];
        }
    };
    __syncthreads();
    fetch(ilo - 6);

    for (int t = ilo - 6; t <= ihi + 6; ++t) {
        // ---------------- phase A: rows t (iterates) and t-2 (pointwise inputs)
        const double ut_new = q_ut, vt_new = q_vt, p_rp = q_p, T_rp = q_T, e = q_e, Lp = q_L;
        const double um_rp = q_um, vm_rp = q_vm;
        fetch(t + 1);
        const int rp = t - 2;
        ut.push(ut_new);
        vt.push(vt_new);
        sUT[s4(t)][PAD + x] = ut_new;
        sVT[s4(t)][PAD + x] = vt_new;
        sP[s2(rp)][PAD + x] = p_rp;
        sT[s2(rp)][PAD + x] = T_rp;
        __syncthreads();

        // ---------------- phase B
        if (!chan) {   // boundary columns of the ring rows produced in the previous iteration
            fix_u(sU1[s3(t - 3)], x, j, jm);
            fix_u(sU2[s3(t - 5)], x, j, jm);
            fix_v(sV1[s3(t - 3)], x, j, jm);
            fix_v(sV2[s3(t - 5)], x, j, jm);
        }
        Su.push(ut_new + sUT[s4(t)][PAD + x - 1]);
        Svn.push(sVT[s4(t - 1)][PAD + x + 1] + sVT[s4(t - 2)][PAD + x + 1]);
        pw.push(p_rp);
        {
            double iev = 1.0, k1 = 0.0, k2 = 0.0;
            if (rp >= ilo - 4 && rp <= ihi + 4) {
                iev = e == 1.0 ? 1.0 : 1.0 / e;
                if (Lp != 0.0) {
                    k1 = e * Lp * a.inv_reda;
                    k2 = Lp * a.cfsd * sqrt(e);
                }
            }
            ie.push(iev); ew.push(e); k1w.push(k1); k2w.push(k2);
            if (MUPOLY) {
                const double tt = T_rp;
                nuw.push(e * a.inv_re * ((((a.mu[4] * tt + a.mu[3]) * tt + a.mu[2]) * tt + a.mu[1]) * tt + a.mu[0]));
            }
            dpj.push(p_rp - sP[s2(rp)][PAD + x - 1]);
            const double Tm = sT[s2(rp)][PAD + x - 1];
            bw.push(bouss ? a.ri * ((T_rp + Tm) * 0.5 - a.t_ref) : a.invfr2 * (1.0 - 2.0 / (T_rp + Tm)));
            if (!STEADY) { du.push(um_rp - ut[2]); dv.push(vm_rp - vt[2]); }
        }

        // ---------------- U stage K at row r = t - 2K (u update: i = 3..imax-1, j = 2..jmax-1)
#define U_STAGE(K, WP, RING_C, WOUT, RING_OUT)                                                   \
        {                                                                                         \
            const int r = t - 2 * K;                                                              \
            const int bk = 2 * K - 2; /* back index of row r in the rows t-6..t-2 windows */      \
            double val = 0.0, rs = 0.0;                                                           \
            bool interior = false;                                                                \
            if (r >= ilo - 2 * (3 - K) && r <= ihi + 2 * (3 - K)) {                               \
                if (r >= 3 && r <= im - 1) {                                                      \
                    if (ucol) {                                                                   \
                        interior = true;                                                          \
                        const double uP = WP[2], uW = WP[3], uWW = WP[4], uE = WP[1], uEE = WP[0]; \
                        const double *c = &RING_C[PAD + x];                                       \
                        const double uS = c[-1], uN = c[1];                                       \
                        const double iev = ie[bk], ev = ew[bk];                                   \
                        const double nu = MUPOLY ? nuw[bk] : ev * a.inv_re;                       \
                        const double aun_r = pr[r], rdx_r = pr[2 * R + r];                        \
                        const double cns = 0.5 * aun_r * iev, cew = haue_j * iev;                 \
                        const double sE = uE + uP, sW = uP + uW, sS = uS + uP, sN = uP + uN;      \
                        const double fn = Svn[2 * K - 1] * cns;                                   \
                        const double fs = (vt[2 * K] + vt[2 * K + 1]) * cns;                      \
                        const double fe = sE * cew, fw = sW * cew;                                \
                        const double diff = (nu * aue_j) * (pr[3 * R + r] * (uW - uP) + pr[3 * R + r + 1] * (uE - uP)) \
                                          + (nu * aun_r) * (rdy_j * (uS - uP) + rdy_n * (uN - uP)); \
                        double pw_, pe_, ps_, pn_;                                                \
                        if (u_upw_col) {                                                          \
                            pw_ = fw > 0.0 ? uW : uP; pe_ = fe > 0.0 ? uP : uE;                   \
                            ps_ = fs > 0.0 ? uS : uP; pn_ = fn > 0.0 ? uP : uN;                   \
                        } else {                                                                  \
                            const double uSS = c[-2], uNN = c[2];                                 \
                            const double cW = uWW - 2.0 * uW + uP, cPi = uW - 2.0 * uP + uE, cE = uP - 2.0 * uE + uEE; \
                            const double cS = uSS - 2.0 * uS + uP, cPj = uS - 2.0 * uP + uN, cN = uP - 2.0 * uN + uNN; \
                            pw_ = quick_face(fw, sW, cW, cPi); pe_ = quick_face(fe, sE, cPi, cE); \
                            ps_ = quick_face(fs, sS, cS, cPj); pn_ = quick_face(fn, sN, cPj, cN); \
                        }                                                                         \
                        const double conv = fw * pw_ - fe * pe_ + fs * ps_ - fn * pn_;            \
                        double Rr = rdx_r * (rdy_j * (diff + conv) - ev * (pw[bk] - pw[bk + 1])) - a.g * ev; \
                        const double k2 = k2w[bk];                                                \
                        if (k2 != 0.0) {                                                          \
                            const double vP = vt[2 * K];                                          \
                            Rr -= uP * (k1w[bk] + k2 * sqrt(uP * uP + vP * vP));                  \
                        }                                                                         \
                        rs = ((STEADY ? 0.0 : du[bk]) + Rr * dt) * dtau;                          \
                        const double tau = ut[2 * K];                                             \
                        if (K == 1) val = tau + rs;                                               \
                        else if (K == 2) val = 0.75 * tau + 0.25 * (uP + rs);                     \
                        else val = 1.0 / 3.0 * tau + 2.0 / 3.0 * (uP + rs);                      \
                    }                                                                             \
                } else if (r <= 2) {                                                              \
                    val = chan ? a.u_in : 0.0;                                                    \
                } else {                                                                          \
                    val = (chan && ucol) ? WOUT[0] : 0.0;                                         \
                }                                                                                 \
                RING_OUT                                                                          \
            }                                                                                     \
            WOUT.push(val);                                                                       \
            if (K == 3 && interior && out_col && r >= ilo && r < ihi) {                           \
                a.un[(size_t)r * su + j] = val;                                                   \
                const unsigned long long bb = dbits(rs); mu_ = bb > mu_ ? bb : mu_;               \
            }                                                                                     \
        }

        // ---------------- V stage K at row r = t - 2K (v update: i = 2..imax-1, j = 3..jmax-1)
#define V_STAGE(K, WP, RING_C, WOUT, RING_OUT)                                                   \
        {                                                                                         \
            const int r = t - 2 * K;                                                              \
            const int bk = 2 * K - 2;                                                             \
            double val = 0.0, rs = 0.0;                                                           \
            bool interior = false;                                                                \
            if (r >= ilo - 2 * (3 - K) && r <= ihi - 1 + 2 * (3 - K)) {                           \
                if (j <= 2) {                                                                     \
                    val = chan ? 0.0 : vi;                                                        \
                } else if (r >= 2 && r <= im - 1) {                                               \
                    if (vcol) {                                                                   \
                        interior = true;                                                          \
                        const double vP = WP[2], vW = WP[3], vWW = WP[4], vE = WP[1], vEE = WP[0]; \
                        const double *c = &RING_C[PAD + x];                                       \
                        const double vS = c[-1], vN = c[1];                                       \
                        const double iev = ie[bk], ev = ew[bk];                                   \
                        const double nu = MUPOLY ? nuw[bk] : ev * a.inv_re;                       \
                        const double avn_r = pr[R + r], rdx_r = pr[2 * R + r];                    \
                        const double cns = 0.5 * avn_r * iev, cew = have_j * iev;                 \
                        const double sN = vP + vN, sS = vP + vS, sE = vE + vP, sW = vP + vW;      \
                        const double fn = sN * cns, fs = sS * cns;                                \
                        const double fe = Su[2 * K - 1] * cew, fw = Su[2 * K] * cew;              \
                        const double diff = (nu * ave_j) * (rdx_r * (vW - vP) + pr[2 * R + r + 1] * (vE - vP)) \
                                          + (nu * avn_r) * (rdym_j * (vS - vP) + rdym_n * (vN - vP)); \
                        double pw_, pe_, ps_, pn_;                                                \
                        if (upwind_all || r == 2 || r == im - 1) {                                \
                            pw_ = fw > 0.0 ? vW : vP; pe_ = fe > 0.0 ? vP : vE;                   \
                            ps_ = fs > 0.0 ? vS : vP; pn_ = fn > 0.0 ? vP : vN;                   \
                        } else {                                                                  \
                            const double vSS = c[-2], vNN = c[2];                                 \
                            const double cW = vWW - 2.0 * vW + vP, cPi = vW - 2.0 * vP + vE, cE = vP - 2.0 * vE + vEE; \
                            const double cS = vSS - 2.0 * vS + vP, cPj = vS - 2.0 * vP + vN, cN = vP - 2.0 * vN + vNN; \
                            pw_ = quick_face(fw, sW, cW, cPi); pe_ = quick_face(fe, sE, cPi, cE); \
                            ps_ = quick_face(fs, sS, cS, cPj); pn_ = quick_face(fn, sN, cPj, cN); \
                        }                                                                         \
                        const double conv = fw * pw_ - fe * pe_ + fs * ps_ - fn * pn_;            \
                        double Rr = rdy_j * (rdx_r * (diff + conv) - ev * dpj[bk]) + bw[bk];      \
                        const double k2 = k2w[bk];                                                \
                        if (k2 != 0.0) {                                                          \
                            const double uP = ut[2 * K];                                          \
                            Rr -= vP * (k1w[bk] + k2 * sqrt(uP * uP + vP * vP));                  \
                        }                                                                         \
                        rs = ((STEADY ? 0.0 : dv[bk]) + Rr * dt) * dtau;                          \
                        const double tau = vt[2 * K];                                             \
                        if (K == 1) val = tau + rs;                                               \
                        else if (K == 2) val = 0.75 * tau + 0.25 * (vP + rs);                     \
                        else val = 1.0 / 3.0 * tau + 2.0 / 3.0 * (vP + rs);                       \
                        if (r == 2 && !chan) WOUT.set(0, val);   /* row 1 = row 2 (boundary.f90:34-36) */ \
                    }                                                                             \
                } else if (r == im) {                                                             \
                    val = vcol ? WOUT[0] : 0.0;                                                   \
                }                                                                                 \
                RING_OUT                                                                          \
            }                                                                                     \
            WOUT.push(val);                                                                       \
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
                const double avn_r = pr[R + r];
                const double rpv = -((u3e * aue_j - u3p * aue_j) + (v3n * avn_r - v3p * avn_r));
                const double dd = dtau * rpv * a.beta;
                const double pp = pw[5];
                double pi = pp + dd;
                pi = 0.75 * pp + 0.25 * (pi + dd);
                a.pn[(size_t)r * su + j] = 1.0 / 3.0 * pp + 2.0 / 3.0 * (pi + dd);
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
    const double *un, *vn, *T, *Tt, *C, *Ct, *eps, *lp, *y;
    double *Tn, *Cn;
    const double *pk;
    int imax, jmax, nbi, nbj, hb, chan;
    double dt, dtau, kt, kc, tinf, t_bottom, t_top, v_i, u_in, slot_lo, slot_hi, slot_value;
};

// T of the centre row at column j+dj (boundary.f90:85-88 / channel wall temperatures).
__device__ __forceinline__ double lds_t(const double *row, int x, int j, int dj, int jm, bool chan, const K2Args &a)
{
    const int jj = j + dj;
    if (jj == 1) return chan ? a.t_bottom : a.tinf;
    if (jj == jm) {
        if (chan) return a.t_top;
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

template <bool STEADY, int OCC>
__global__ __launch_bounds__(NT2, OCC) void k_scal(const K2Args a)
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
    const bool chan = a.chan != 0;
    const double dt = a.dt, dtau = a.dtau, kt = a.kt, kc = a.kc;
    const int R = im + 2, Cc = jm + 2;
    const double *pr = a.pk, *pc = a.pk + 4 * R;

    __shared__ double sTT[3][LROW2], sCT[3][LROW2], sV[2][LROW2];
    __shared__ double sZ1[2][LROW2], sZ2[2][LROW2], sC1[2][LROW2], sC2[2][LROW2];
    for (int k = x; k < LROW2; k += NT2) {
        for (int q = 0; q < 3; ++q) { sTT[q][k] = 0.0; sCT[q][k] = 0.0; }
        for (int q = 0; q < 2; ++q) { sV[q][k] = 0.0; sZ1[q][k] = 0.0; sZ2[q][k] = 0.0; sC1[q][k] = 0.0; sC2[q][k] = 0.0; }
    }
    const bool jvalid = j >= 1 && j <= jm;
    const int jc = min(max(j, 2), jm - 1);
    const double aue_j = pc[min(max(j, 1), jm)];
    const double rdy_j = pc[2 * Cc + jc], rdy_n = pc[2 * Cc + jc + 1];
    const double rdym_n = pc[3 * Cc + jc + 1];
    const bool scol = j >= 2 && j <= jm - 1;
    const bool out_col = j >= jlo && j < jhi;
    const double yj = a.y[min(max(j, 1), jm)];
    const double cslot = chan ? ((yj >= a.slot_lo && yj <= a.slot_hi) ? a.slot_value : 0.0) : 0.0;

    Win<4> tt, ct, uw, vw, vnw, id, oml, dT, dC;
    Win<3> z1, z2, c1, c2, z3, c3;
    tt.fill(0.0); ct.fill(0.0); uw.fill(0.0); vw.fill(0.0); vnw.fill(0.0); id.fill(1.0); oml.fill(1.0);
    if (!STEADY) { dT.fill(0.0); dC.fill(0.0); }
    z1.fill(0.0); z2.fill(0.0); c1.fill(0.0); c2.fill(0.0); z3.fill(0.0); c3.fill(0.0);

    const size_t su = (size_t)(jm + 1), sv = (size_t)(jm + 2);
    double q_t, q_c, q_u, q_v, q_T, q_C, q_e, q_L;
    // next row: u3/v3 from K1's interior cells, boundary rows/columns from the BC formulas
    auto fetch = [&](int t) {
        q_t = 0.0; q_c = 0.0; q_u = 0.0; q_v = 0.0; q_T = 0.0; q_C = 0.0; q_e = 1.0; q_L = 0.0;
        if (jvalid && t >= 1 && t <= im + 1) {
            if (t <= im) {
                const size_t k = (size_t)t * su + j;
                q_t = a.Tt[k]; q_c = a.Ct[k]; q_e = a.eps[k]; q_L = a.lp[k];
                if (!STEADY) { q_T = a.T[k]; q_C = a.C[k]; }
                if (j == 2) q_v = chan ? 0.0 : a.v_i;
                else if (t >= 2 && t <= im - 1 && j <= jm - 1) q_v = a.vn[(size_t)t * sv + j];
            }
            if (t > 2 && t < im) q_u = a.un[(size_t)t * su + j];
        }
    };
    __syncthreads();
    fetch(ilo - 3);

    for (int t = ilo - 3; t <= ihi + 2; ++t) {
        // ---------------- phase A: row t
        const double tnew = q_t, cnew = q_c, vnew = q_v, Tp = q_T, Cp = q_C, e = q_e, Lp = q_L;
        double unew = q_u;
        fetch(t + 1);
        if (t <= 2) unew = chan ? a.u_in : 0.0;             // u3 boundary rows (boundary.f90:29-32,39-42)
        else if (t >= im) unew = chan ? uw[0] : 0.0;
        tt.push(tnew); ct.push(cnew); uw.push(unew); vw.push(vnew);
        sTT[s3(t)][PAD + x] = tnew;
        sCT[s3(t)][PAD + x] = cnew;
        sV[s2(t)][PAD + x] = vnew;
        __syncthreads();

        // ---------------- phase B: per-cell data of row t
        {
            double vn_ = sV[s2(t)][PAD + x + 1];
            if (j + 1 == jm) vn_ = chan ? 0.0 : 2.0 * sV[s2(t)][PAD + x - 1] - sV[s2(t)][PAD + x];
            vnw.push(vn_);
            id.push(Lp != 0.0 ? 1.0 / (Lp * (e - 1.0) + 1.0) : 1.0);
            oml.push(1.0 - Lp);
            if (!STEADY) { dT.push(Tp - tnew); dC.push(Cp - cnew); }
        }

        // stage K of a scalar at row r = t - K; newest own row of WP is t-K+1
#define SCAL_STAGE(K, WP, CROW, ISC, WOUT, RINGW, OUTP, KAPPA, DW, TAUW)                         \
        {                                                                                         \
            const int r = t - K;                                                                  \
            const int bq = K;            /* back index of row r in windows whose newest is t */   \
            double val = 0.0;                                                                     \
            bool interior = false;                                                                \
            if (r >= ilo - (3 - K) && r <= ihi - 1 + (3 - K)) {                                   \
                if (r >= 2 && r <= im - 1) {                                                      \
                    if (scol) {                                                                   \
                        interior = true;                                                          \
                        const double zE = WP[0], zP = WP[1], zW = WP[2];                          \
                        const double *c = CROW;                                                   \
                        const double zS = ISC ? lds_c(c, x, j, -1, jm, chan) : lds_t(c, x, j, -1, jm, chan, a); \
                        const double zN = ISC ? lds_c(c, x, j, 1, jm, chan) : lds_t(c, x, j, 1, jm, chan, a);   \
                        const double avn_r = pr[R + r];                                           \
                        const double diff = (KAPPA * aue_j) * (pr[2 * R + r + 1] * (zE - zP) + pr[2 * R + r] * (zW - zP)) \
                                          + (KAPPA * avn_r) * (rdy_n * (zN - zP) + rdy_j * (zS - zP)); \
                        const double uE = uw[bq - 1], uP = uw[bq], vP = vw[bq], vN = vnw[bq];     \
                        const double conv = 0.5 * ((zE + zP) * uE - (zW + zP) * uP) * aue_j       \
                                          + 0.5 * ((zN + zP) * vN - (zS + zP) * vP) * avn_r;      \
                        const double vols = pr[3 * R + r + 1] * rdym_n;                           \
                        const double Rr = ISC ? vols * (diff - conv * id[bq])                     \
                                              : vols * (diff - oml[bq] * conv) * id[bq];          \
                        const double rs = ((STEADY ? 0.0 : DW[bq]) + Rr * dt) * dtau;             \
                        const double tau = TAUW[bq];                                              \
                        if (K == 1) val = tau + rs;                                               \
                        else if (K == 2) val = 0.75 * tau + 0.25 * (zP + rs);                     \
                        else val = 1.0 / 3.0 * tau + 2.0 / 3.0 * (zP + rs);                       \
                        if (r == 2) WOUT.set(0, (ISC && chan) ? cslot : val);  /* row 1 */        \
                    }                                                                             \
                } else if (r == im) {                                                             \
                    val = scol ? WOUT[0] : 0.0;                                                   \
                } else if (r == 1) {                                                              \
                    val = (ISC && chan) ? cslot : 0.0;                                            \
                }                                                                                 \
                RINGW                                                                             \
            }                                                                                     \
            WOUT.push(val);                                                                       \
            if (OUTP != nullptr && interior && out_col && r >= ilo && r < ihi)                    \
                OUTP[(size_t)r * su + j] = val;                                                   \
        }

        double *const noout = nullptr;
        SCAL_STAGE(1, tt, sTT[s3(t - 1)], false, z1, { sZ1[s2(t - 1)][PAD + x] = val; }, noout, kt, dT, tt)
        SCAL_STAGE(1, ct, sCT[s3(t - 1)], true, c1, { sC1[s2(t - 1)][PAD + x] = val; }, noout, kc, dC, ct)
        SCAL_STAGE(2, z1, sZ1[s2(t - 2)], false, z2, { sZ2[s2(t - 2)][PAD + x] = val; }, noout, kt, dT, tt)
        SCAL_STAGE(2, c1, sC1[s2(t - 2)], true, c2, { sC2[s2(t - 2)][PAD + x] = val; }, noout, kc, dC, ct)
        SCAL_STAGE(3, z2, sZ2[s2(t - 3)], false, z3, {}, a.Tn, kt, dT, tt)
        SCAL_STAGE(3, c2, sC2[s2(t - 3)], true, c3, {}, a.Cn, kc, dC, ct)
#undef SCAL_STAGE
    }
}
}  // namespace

extern "C" {

int cfdk_fused_available(void) { return 1; }

// Row-block height: `want` > 0 (from the autotuner) or one block per CU, at least min_rows.
static int pick_rows(int rows, int nbj, int min_rows, int want, int *hb)
{
    if (want > 0) {
        *hb = want < rows ? want : rows;
        return cdiv(rows, *hb);
    }
    int nbi = 256 / nbj;
    if (nbi < 1) nbi = 1;
    if (rows / nbi < min_rows) nbi = rows / min_rows > 0 ? rows / min_rows : 1;
    *hb = cdiv(rows, nbi);
    return cdiv(rows, *hb);
}

int cfdk_fused_step(const cfd_devgeo *g, const cfd_devfactors *f, const cfd_params *P,
                    const cfd_fused_args *a, int kernel, hipStream_t s)
{
    (void)kernel;
    const int im = g->imax, jm = g->jmax;
    if (P->b_art != 0.0 || P->compat != CFD_COMPAT_FORTRAN) return CFD_ERR_UNSUPPORTED;
    // K1
    K1Args k1;
    k1.um = a->um; k1.vm = a->vm; k1.ut = a->um_tau; k1.vt = a->vm_tau; k1.p = a->p; k1.T = a->T;
    k1.eps = g->eps; k1.lp = g->lp;
    k1.un = a->um_n; k1.vn = a->vm_n; k1.pn = a->pn; k1.res = (unsigned long long *)a->res_max;
    k1.pk = f->pack;
    k1.imax = im; k1.jmax = jm;
    k1.chan = P->bc == CFD_BC_CHANNEL; k1.upwind_all = P->scheme == CFD_SCHEME_UPWIND;
    k1.bouss = P->buoyancy == CFD_BUOY_BOUSSINESQ;
    k1.dt = P->dt; k1.dtau = P->dtau; k1.beta = P->beta; k1.inv_re = 1.0 / P->re;
    k1.inv_reda = 1.0 / (P->re * P->darcy); k1.cfsd = P->cf / sqrt(P->darcy); k1.g = P->g;
    k1.invfr2 = P->invfr2; k1.ri = P->ri; k1.t_ref = P->t_ref; k1.v_i = P->v_i; k1.u_in = P->u_in;
    for (int q = 0; q < 5; ++q) k1.mu[q] = P->mu_coef[q];
    k1.nbj = cdiv(jm - 2, W1);
    k1.nbi = pick_rows(im - 2, k1.nbj, 24, a->hb1, &k1.hb);
    const int G1 = k1.nbi * k1.nbj, grid1 = ((G1 + 7) / 8) * 8;
    const bool mp = P->mu_model == CFD_MU_POLY4;
    const bool timed = a->ev[0] != nullptr;
    if (timed) hipEventRecord(a->ev[0], s);
    const int o1 = a->occ1 == 2 ? 2 : 1;
    const dim3 g1(grid1), b1(NT1);
    if (P->steady) {
        if (mp) {
            if (o1 == 2) hipLaunchKernelGGL((k_uvp<true, true, 2>), g1, b1, 0, s, k1);
            else hipLaunchKernelGGL((k_uvp<true, true, 1>), g1, b1, 0, s, k1);
        } else {
            if (o1 == 2) hipLaunchKernelGGL((k_uvp<true, false, 2>), g1, b1, 0, s, k1);
            else hipLaunchKernelGGL((k_uvp<true, false, 1>), g1, b1, 0, s, k1);
        }
    } else {
        if (mp) {
            if (o1 == 2) hipLaunchKernelGGL((k_uvp<false, true, 2>), g1, b1, 0, s, k1);
            else hipLaunchKernelGGL((k_uvp<false, true, 1>), g1, b1, 0, s, k1);
        } else {
            if (o1 == 2) hipLaunchKernelGGL((k_uvp<false, false, 2>), g1, b1, 0, s, k1);
            else hipLaunchKernelGGL((k_uvp<false, false, 1>), g1, b1, 0, s, k1);
        }
    }
    int r = launched();
    if (r != CFD_OK) return r;
    if (timed) hipEventRecord(a->ev[1], s);
    // K2
    K2Args k2;
    k2.un = a->um_n; k2.vn = a->vm_n; k2.T = a->T; k2.Tt = a->T_tau; k2.C = a->C; k2.Ct = a->C_tau;
    k2.eps = g->eps; k2.lp = g->lp; k2.y = g->y;
    k2.Tn = a->T_n; k2.Cn = a->C_n;
    k2.pk = f->pack;
    k2.imax = im; k2.jmax = jm; k2.chan = P->bc == CFD_BC_CHANNEL;
    k2.dt = P->dt; k2.dtau = P->dtau; k2.kt = 1.0 / P->pe; k2.kc = 1.0 / P->re / P->sc;
    k2.tinf = P->tinf; k2.t_bottom = P->t_bottom; k2.t_top = P->t_top; k2.v_i = P->v_i; k2.u_in = P->u_in;
    k2.slot_lo = P->slot_lo; k2.slot_hi = P->slot_hi; k2.slot_value = P->slot_value;
    k2.nbj = cdiv(jm - 2, W2);
    k2.nbi = pick_rows(im - 2, k2.nbj, 16, a->hb2, &k2.hb);
    const int G2 = k2.nbi * k2.nbj, grid2 = ((G2 + 7) / 8) * 8;
    const dim3 g2(grid2), b2(NT2);
    if (a->occ2 == 4) {
        if (P->steady) hipLaunchKernelGGL((k_scal<true, 4>), g2, b2, 0, s, k2);
        else hipLaunchKernelGGL((k_scal<false, 4>), g2, b2, 0, s, k2);
    } else {
        if (P->steady) hipLaunchKernelGGL((k_scal<true, 1>), g2, b2, 0, s, k2);
        else hipLaunchKernelGGL((k_scal<false, 1>), g2, b2, 0, s, k2);
    }
    r = launched();
    if (r != CFD_OK) return r;
    if (timed) hipEventRecord(a->ev[2], s);
    r = cfdk_materialize_bc(g, P, a->um_n, a->vm_n, a->pn, a->T_n, a->C_n, a->vm_tau, s);
    if (timed) hipEventRecord(a->ev[3], s);
    return r;
}

}  // extern "C"

// ---------------------------------------------------------------------------------------
// support launchers
// ---------------------------------------------------------------------------------------
namespace {
__global__ void k_factors(Geo g, double *rdx, double *rdy, double *rdxm, double *rdym, double *pack)
{
    const int t = blockIdx.x * blockDim.x + threadIdx.x;
    const int R = g.imax + 2, Cc = g.jmax + 2;
    double *pr = pack, *pc = pack + 4 * R;
    if (t < R) {
        const double a = (t >= 2 && t <= g.imax) ? 1.0 / (g.x[t] - g.x[t - 1]) : 0.0;
        const double b = (t >= 2 && t <= g.imax + 1) ? 1.0 / (g.xm[t] - g.xm[t - 1]) : 0.0;
        if (t >= 2 && t <= g.imax) rdx[t] = a;
        if (t >= 2 && t <= g.imax + 1) rdxm[t] = b;
        pr[t] = (t >= 2 && t <= g.imax) ? g.aun[t] : 0.0;
        pr[R + t] = (t >= 1 && t <= g.imax) ? g.avn[t] : 0.0;
        pr[2 * R + t] = a;
        pr[3 * R + t] = b;
    }
    if (t < Cc) {
        const double a = (t >= 2 && t <= g.jmax) ? 1.0 / (g.y[t] - g.y[t - 1]) : 0.0;
        const double b = (t >= 2 && t <= g.jmax + 1) ? 1.0 / (g.ym[t] - g.ym[t - 1]) : 0.0;
        if (t >= 2 && t <= g.jmax) rdy[t] = a;
        if (t >= 2 && t <= g.jmax + 1) rdym[t] = b;
        pc[t] = (t >= 1 && t <= g.jmax) ? g.aue[t] : 0.0;
        pc[Cc + t] = (t >= 2 && t <= g.jmax) ? g.ave[t] : 0.0;
        pc[2 * Cc + t] = a;
        pc[3 * Cc + t] = b;
    }
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
    hipLaunchKernelGGL(k_factors, dim3(cdiv(n, 256)), dim3(256), 0, s, mk(g), f->rdx, f->rdy, f->rdxm, f->rdym, f->pack);
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
