// cfd_fused.hip — the fused pseudo-step: two row-streaming kernels and their support.
//
// One pseudo-step (main.f90:122-132: solve_U, solve_V, solve_P, solve_Z, solve_C) runs as
//   K1  k_uvp    the three SSP-RK3 stages of u and v and the pressure update,
//   K2  k_scal   the three stages of T and of C,
//   k_materialize  the boundary cells of the five outputs (the kernels themselves read
//                  boundary values of in-flight stages through the BC formulas).
//
// Decomposition. Lanes walk j (the contiguous axis): every global access of a wavefront is
// one coalesced row segment. A workgroup owns a column strip and a block of rows (a tile)
// and marches down its rows; stage k of a chain is computed a fixed number of rows behind
// stage k-1 (QUICK reaches two rows), so all three RK stages and the pressure update of a
// row happen in one pass over the inputs. i-direction neighbours come from per-lane
// register windows (Win<N>), j-direction neighbours of the centre row from small LDS rings.
// Each input row is read from HBM once per pass (halo columns/rows of neighbouring tiles
// hit L2/MALL).
//
// Precision. The tile routines are templated on the field type R: double (the FP64
// drop-in) or float (FP32 storage and stencil arithmetic). Pressure (p, pn), the
// continuity update and the residual maxima stay FP64 in both.
//
// Two launch forms share the tile routines: one grid per launch (k_uvp / k_scal, tile
// from blockIdx with an XCD-aware order) and a persistent batch (k_uvp_batch /
// k_scal_batch: a fixed grid walks a tile list that spans many independent cases, each
// with its own sizes, parameters and buffers, skipping cases that have converged).
//
// Numerics: flux-form QUICK (face value 1/2 (U+D) - 1/8 curvature at the upstream cell,
// which equals the reference's coefficient assembly of equations.f90:51-72), reciprocals of
// the 1-D spacings precomputed, FMA contraction allowed: FP64 results differ from the
// literal reference arithmetic by rounding only (tested at 1e-10 relative).
#include <type_traits>

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
template <int N, typename T>
struct Win {
    T v[N];
    __device__ __forceinline__ void fill(T x)
    {
#pragma unroll
        for (int k = 0; k < N; ++k) v[k] = x;
    }
    __device__ __forceinline__ void push(T x)
    {
#pragma unroll
        for (int k = 0; k < N - 1; ++k) v[k] = v[k + 1];
        v[N - 1] = x;
    }
    __device__ __forceinline__ T operator[](int back) const { return v[N - 1 - back]; }
    __device__ __forceinline__ void set(int back, T x) { v[N - 1 - back] = x; }
};

__device__ __forceinline__ int s2(int r) { return (r + 24) & 1; }
__device__ __forceinline__ int s3(int r) { return (r + 24) % 3; }
__device__ __forceinline__ int s4(int r) { return (r + 24) & 3; }
__device__ __forceinline__ int s8(int r) { return (r + 24) & 7; }
// ring slot of row t + d (d a compile-time offset, -10 < d < 10) from m5 = (t + 40) % 5
__device__ __forceinline__ int s5(int m5, int d)
{
    const int q = m5 + (d + 10) % 5;
    return q >= 5 ? q - 5 : q;
}

__device__ __forceinline__ unsigned long long dbits(double v)
{
    return (unsigned long long)__double_as_longlong(fabs(v));
}

// Sum over the wavefront, one atomic add per wave (skipped when the sum is zero).
__device__ __forceinline__ void fold_sum(double *out, double v)
{
    for (int o = 32; o > 0; o >>= 1) v += __shfl_xor(v, o, 64);
    if ((threadIdx.x & 63) == 0 && v != 0.0) atomicAdd(out, v);
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
template <typename R>
__device__ __forceinline__ R quick_face(R f, R s, R ca, R cb)
{
    return R(0.5) * s - R(0.125) * (f > R(0) ? ca : cb);
}

// Square root of x >= 0 (a sum of squares, a porosity): the hardware reciprocal square root
// r, then the coupled Newton iteration for g ~ sqrt(x) and h ~ 1/(2 sqrt(x)) (g = x r,
// h = r / 2; e = 1/2 - h g; g += g e, h += h e; then twice d = x - g^2, g += d h). This is the
// operation sequence of the compiler's FP64 square root for x >= 2^-767, which it takes without
// rescaling the argument, so the result is the same there; for smaller x (velocities below
// 1e-115) it can differ by a few ulps. x = 0 and x = +inf return x, NaN stays NaN.
__device__ __forceinline__ double csqrt(double x)
{
    const double r = __builtin_amdgcn_rsq(x);
    double g = x * r, h = 0.5 * r;
    const double e = fma(-h, g, 0.5);
    g = fma(g, e, g);
    h = fma(h, e, h);
    double d = fma(-g, g, x);
    g = fma(d, h, g);
    d = fma(-g, g, x);
    g = fma(d, h, g);
    return __builtin_amdgcn_class(x, 0x260) ? x : g;   // +-0, +inf
}
__device__ __forceinline__ float csqrt(float x) { return sqrtf(x); }

// One tile of a persistent batch: case index, row block, column strip.
struct BatchTile {
    int c, bi, bj, pad;
};

// ---------------------------------------------------------------------------------------
// K1: u, v (three RK stages each) and p.
// ---------------------------------------------------------------------------------------
// A workgroup (or, in the split-chain kernels, each chain) spans NT columns, one per lane:
// NT1 = 256 for a single grid; NTB = 64 (one wave) for batch cases, whose grids are narrow:
// a 64..256-wide case fills 24..50 % of the lanes of 256-column strips (a strip carries 13
// halo columns, so 254 interior columns already need two), and 48..79 % of 64-column ones.
constexpr int NT1 = 256;
constexpr int NTB = 64;
constexpr int HL1 = 6, HR1 = 7;                 // column halo of the U/V/P chain
constexpr int PAD = 3;                          // LDS padding for j-3 .. j+2 reads
__host__ __device__ constexpr int w1(int nt) { return nt - HL1 - HR1; }   // output columns per workgroup
__host__ __device__ constexpr int lrow(int nt) { return nt + 2 * PAD; }

// Scalars are precomputed on the host so that few values occupy scalar registers.
template <typename R>
struct K1Args {
    const R *um, *vm, *ut, *vt;                            // inputs
    const double *p;
    const R *T, *eps, *lp;
    R *un, *vn;                                            // outputs (interior cells)
    double *pn;
    unsigned long long *res;                               // max |r3| bit patterns: u, v, p
    double *l2;                                            // sums of r3^2 (u, v, p) or null
    unsigned long long *res_next;                          // next step's slots, cleared here, or null
    const double *pk;                                      // packed geometry (cfd_devfactors.pack)
    int imax, jmax, nbi, nbj, hb, chan, upwind_all, bouss;
    double dt, dtau, beta, inv_re, inv_reda, cfsd, g, invfr2, ri, t_ref, v_i, u_in;
    double mu[5];
};

// Two layouts of the stage values of u and v. RINGS = false (the one-wave kernel): each lane
// keeps its own column's recent stage rows in register windows, and 4-row LDS rings serve
// only the j-neighbours. RINGS = true (the two-wave kernel, 256 registers): the stage rows
// live only in LDS, 5-row rings (slot = row mod 5) that hold rows r-2..r+2 for the next
// stage, and the u_tau ring keeps rows t..t-6 so that the v stages' u_tau(j) + u_tau(j-1)
// sums are read rather than kept; the workgroup then needs just under 80 KB, so that two fit
// on a CU.
template <typename R, int NT, bool RINGS>
struct Smem1 {
    static constexpr int UTR = RINGS ? 8 : 4, SR = RINGS ? 5 : 4;
    R sUT[UTR][lrow(NT)], sVT[4][lrow(NT)], sT[2][lrow(NT)];
    R sU1[SR][lrow(NT)], sU2[SR][lrow(NT)], sV1[SR][lrow(NT)], sV2[SR][lrow(NT)], sV3[2][lrow(NT)];
    double sP[2][lrow(NT)];
    // per-row constants, read from here once per row rather than held in scalar registers
    // for the whole march
    double cst[16];
};
enum { CST_INV_REDA, CST_CFSD, CST_MU0, CST_RI = CST_MU0 + 5, CST_T_REF, CST_INVFR2, CST_U_IN, CST_BETA, CST_N };

// v of the centre row at column j+dj with the BC formulas (boundary.f90:9-12,19-22 /
// channel walls): used where the ring's boundary columns are not yet fixed up.
template <typename R>
__device__ __forceinline__ R lds_v(const R *row, int x, int j, int dj, int jm, bool chan, R vi)
{
    const int jj = j + dj;
    if (jj <= 2) return chan ? R(0) : vi;
    if (jj == jm) return chan ? R(0) : R(2) * row[PAD + x + dj - 2] - row[PAD + x + dj - 1];
    if (jj == jm + 1) {
        if (chan) return R(0);
        const R a = row[PAD + x + dj - 3], b = row[PAD + x + dj - 2];
        return R(2) * b - (R(2) * a - b);
    }
    return row[PAD + x + dj];
}

// Reference-set boundary columns of a u ring row (u(jmax) = u(jmax-1), boundary.f90:24-26)
// and of a v ring row (top extrapolation, boundary.f90:19-22), written by the lanes that
// own those columns one iteration after the row was produced and before it is read.
template <typename R>
__device__ __forceinline__ void fix_u(R *row, int x, int j, int jm)
{
    if (j == jm) row[PAD + x] = row[PAD + x - 1];
}
template <typename R>
__device__ __forceinline__ void fix_v(R *row, int x, int j, int jm)
{
    if (j == jm) {
        row[PAD + x] = R(2) * row[PAD + x - 2] - row[PAD + x - 1];
    } else if (j == jm + 1) {
        const R a = row[PAD + x - 3], b = row[PAD + x - 2];
        row[PAD + x] = R(2) * b - (R(2) * a - b);
    }
}

// L2: also fold the sums of r3^2 (single grids; batches report the maxima only, and their
// register budget stays as it was)
template <typename R, bool STEADY, bool MUPOLY, int NT, bool RINGS, bool L2>
__device__ __forceinline__ void uvp_tile(const K1Args<R> &a, int bi, int bj, Smem1<R, NT, RINGS> &sm)
{
    const int im = a.imax, jm = a.jmax;
    const int ilo = 2 + bi * a.hb, ihi = min(ilo + a.hb, im);        // output rows [ilo, ihi)
    const int jlo = 2 + bj * w1(NT), jhi = min(jlo + w1(NT), jm);            // output cols [jlo, jhi)
    const int x = threadIdx.x;
    const int j = jlo - HL1 + x;
    const bool chan = a.chan != 0, upwind_all = a.upwind_all != 0, bouss = a.bouss != 0;
    const R dt = R(a.dt), dtau = R(a.dtau), vi = R(a.v_i), inv_re = R(a.inv_re), gg = R(a.g);
    const int RR = im + 2, Cc = jm + 2;
    const double *pr = a.pk, *pc = a.pk + 4 * RR;    // rows (aun avn rdx rdxm) x i, cols [aue|ave|rdy|rdym]
    auto sUT = sm.sUT; auto sVT = sm.sVT; auto sT = sm.sT; auto sP = sm.sP;
    auto sU1 = sm.sU1; auto sU2 = sm.sU2; auto sV1 = sm.sV1; auto sV2 = sm.sV2; auto sV3 = sm.sV3;

    if (x < CST_N) {
        const double cv[CST_N] = {a.inv_reda, a.cfsd, a.mu[0], a.mu[1], a.mu[2], a.mu[3], a.mu[4], a.ri, a.t_ref,
                                  a.invfr2, a.u_in, a.beta};
        sm.cst[x] = cv[x];
    }
    const double *cst = sm.cst;
    for (int k = x; k < lrow(NT); k += NT) {
        for (int q = 0; q < sm.UTR; ++q) sUT[q][k] = R(0);
        for (int q = 0; q < 4; ++q) sVT[q][k] = R(0);
        for (int q = 0; q < sm.SR; ++q) { sU1[q][k] = R(0); sU2[q][k] = R(0); sV1[q][k] = R(0); sV2[q][k] = R(0); }
        for (int q = 0; q < 2; ++q) { sP[q][k] = 0.0; sT[q][k] = R(1); sV3[q][k] = R(0); }
    }

    // per-column geometry (clamped indices; values outside the ranges are never used)
    const bool jvalid = j >= 1 && j <= jm + 1;
    const int jc = min(max(j, 2), jm);
    const double aue_jd = pc[min(max(j, 1), jm)];
    const R aue_j = R(aue_jd);
    const R ave_j = R(pc[Cc + jc]);
    // (the one-wave form keeps the half face areas; the two-wave one forms them at use)
    const R haue_j = RINGS ? R(0) : R(0.5) * aue_j, have_j = RINGS ? R(0) : R(0.5) * ave_j;
    const R rdy_j = R(pc[2 * Cc + jc]), rdy_n = R(pc[2 * Cc + min(jc + 1, jm)]);
    const R rdym_j = R(pc[3 * Cc + jc]), rdym_n = R(pc[3 * Cc + min(jc + 1, jm + 1)]);
    const bool ucol = j >= 2 && j <= jm - 1;           // u / p interior column
    const bool vcol = j >= 3 && j <= jm - 1;           // v interior column
    const bool u_upw_col = upwind_all || j == 2 || j == jm - 1;
    const bool out_col = j >= jlo && j < jhi;

    Win<7, R> ut;         // u_tau rows t-6..t
    Win<7, R> Su;         // one-wave form: u_tau(r,j) + u_tau(r,j-1), rows t-6..t
    Win<5, R> w1, w2, x1, x2;   // one-wave form: u1, u2, v1, v2 own windows
    Win<8, R> vt;         // v_tau rows t-7..t
    Win<6, R> Svn;        // v_tau(r,j+1) + v_tau(r-1,j+1), rows t-6..t-1
    Win<6, double> pw;    // p rows t-7..t-2
    Win<5, R> ie, ew, nuw, k1w, k2w, dpj, bw, du, dv;   // per-cell data, rows t-6..t-2
    // The unsteady register-window form keeps four per-cell inputs of rows t-6..t-2 in LDS
    // (slot r mod 5, each lane its own column) rather than in registers: the source terms
    // um - u_tau, vm - v_tau, the pressure difference p(j) - p(j-1) and the buoyancy term.
    constexpr bool DLDS = !RINGS && !STEADY && sizeof(R) == 8;
    __shared__ R sDUV[DLDS ? 20 : 1][DLDS ? NT : 1];
    Win<2, R> w3, x3;                                    // u3, v3
    ut.fill(R(0)); vt.fill(R(0)); Svn.fill(R(0)); pw.fill(0.0);
    if (!RINGS) { Su.fill(R(0)); w1.fill(R(0)); w2.fill(R(0)); x1.fill(R(0)); x2.fill(R(0)); }
    ie.fill(R(1)); ew.fill(R(1)); k1w.fill(R(0)); k2w.fill(R(0)); dpj.fill(R(0)); bw.fill(R(0));
    if (MUPOLY) nuw.fill(R(0));
    if (!STEADY) { du.fill(R(0)); dv.fill(R(0)); }
    w3.fill(R(0)); x3.fill(R(0));
    unsigned long long mu_ = 0, mv_ = 0, mp_ = 0;
    double qu_ = 0.0, qv_ = 0.0, qp_ = 0.0;        // sums of r3^2 (FP64 in both precisions)

    const size_t su = (size_t)(jm + 1), sv = (size_t)(jm + 2);
    // next-row registers: the loads for row t+1 are issued before row t is computed
    R q_ut, q_vt, q_T, q_e, q_L, q_um, q_vm;
    double q_p;
    // Loads at clamped (always valid) indices, one site per field, defaults selected for
    // rows and columns outside the arrays: no divergent branches around the loads.
    const int ju = min(max(j, 1), jm), jv = min(max(j, 1), jm + 1);
    const bool lu = jvalid && j <= jm;
    auto fetch = [&](int t) {
        const int rp = t - 2;
        const int tu = min(max(t, 1), im + 1), tv = min(max(t, 1), im), rc = min(max(rp, 1), im);
        const bool in_u = t >= 1 && t <= im + 1, in_v = t >= 1 && t <= im, in_p = rp >= 1 && rp <= im;
        const size_t k = (size_t)rc * su + ju;
        const R ut_ = a.ut[(size_t)tu * su + ju], vt_ = a.vt[(size_t)tv * sv + jv];
        const double p_ = a.p[k];
        const R T_ = a.T[k], e_ = a.eps[k], L_ = a.lp[k];
        q_ut = ut_; q_vt = vt_; q_p = p_; q_T = T_; q_e = e_; q_L = L_;
        if (!STEADY) { q_um = a.um[k]; q_vm = a.vm[(size_t)rc * sv + jv]; }
        else { q_um = R(0); q_vm = R(0); }
        (void)in_u; (void)in_v; (void)in_p;
    };
    // Rows run two at a time, so that half of the window shifts are register renames rather
    // than moves. Groups start at even rows: a row is always computed by the same position of
    // the group, whatever the block, so blocks of different heights (a batch case against a
    // single context) give bit-identical results. The loop may thus start one row before
    // ilo - 6 and end one row after ihi + 6; in those rows every stage, the P update and
    // every global store are outside their row ranges (only ring slots that no computed row
    // reads are written).
    constexpr int UNR = 2;
    // Register-window form, FP64 or steady: the interior groups run with the row guards
    // compiled out (below). Not in the ring form or in unsteady FP32, whose register budgets
    // (two waves per SIMD, or two u/v/p waves and a T/C wave) the extra live values of the
    // straight-line code would exceed.
    constexpr bool PEEL = !RINGS && (sizeof(R) == 8 || STEADY);
    const int t0 = (ilo - 6) - ((ilo - 6) & (UNR - 1));
    __syncthreads();
    fetch(t0);

    auto row = [&](const int t, auto fullc) __attribute__((always_inline)) {
        // FULL: every row guard below holds (the rows of the group lie in [ilo + 7, min(im, ihi + 5)])
        constexpr bool FULL = decltype(fullc)::value;
        const int m5 = (t + 40) % 5;     // slot of row t in the 5-row stage rings
        // slots of row t + d in the u_tau ring and in the stage rings
#define UTS(d) (RINGS ? s8(t + (d)) : s4(t + (d)))
#define STS(d) (RINGS ? s5(m5, (d)) : s4(t + (d)))
        // ---------------- phase A: rows t (iterates) and t-2 (pointwise inputs)
        // the defaults of rows and columns outside the arrays, applied where the values are
        // consumed (a select at the load would wait for it)
        const bool in_u = FULL || (t >= 1 && t <= im + 1), in_v = FULL || (t >= 1 && t <= im),
                   in_p = FULL || (t - 2 >= 1 && t - 2 <= im);
        const bool pu = in_p && lu, pv = in_p && jvalid;
        const R ut_new = (in_u && lu) ? q_ut : R(0), vt_new = (in_v && jvalid) ? q_vt : R(0);
        const R T_rp = pu ? q_T : R(1), e = pu ? q_e : R(1), Lp = pu ? q_L : R(0);
        const double p_rp = pu ? q_p : 0.0;
        const R um_rp = pu ? q_um : R(0), vm_rp = pv ? q_vm : R(0);
        fetch(t + 1);
        const int rp = t - 2;
        ut.push(ut_new);
        vt.push(vt_new);
        sUT[UTS(0)][PAD + x] = ut_new;
        sVT[s4(t)][PAD + x] = vt_new;
        sP[s2(rp)][PAD + x] = p_rp;
        sT[s2(rp)][PAD + x] = T_rp;
        __syncthreads();

        // ---------------- phase B
        if (!chan) {   // boundary columns of the ring rows produced in the previous iteration
            fix_u(sU1[STS(-3)], x, j, jm);
            fix_u(sU2[STS(-5)], x, j, jm);
            fix_v(sV1[STS(-3)], x, j, jm);
            fix_v(sV2[STS(-5)], x, j, jm);
        }
        if (!RINGS) Su.push(ut_new + sUT[UTS(0)][PAD + x - 1]);
        Svn.push(sVT[s4(t - 1)][PAD + x + 1] + sVT[s4(t - 2)][PAD + x + 1]);
        pw.push(p_rp);
        {
            R iev = R(1), k1 = R(0), k2 = R(0);
            if (FULL || (rp >= ilo - 4 && rp <= ihi + 4)) {
                iev = e == R(1) ? R(1) : R(1) / e;
                if (Lp != R(0)) {
                    k1 = e * Lp * R(cst[CST_INV_REDA]);
                    k2 = Lp * R(cst[CST_CFSD]) * csqrt(e);
                }
            }
            ie.push(iev); ew.push(e); k1w.push(k1); k2w.push(k2);
            if (MUPOLY) {
                const R tt = T_rp;
                nuw.push(e * inv_re * ((((R(cst[CST_MU0 + 4]) * tt + R(cst[CST_MU0 + 3])) * tt + R(cst[CST_MU0 + 2])) * tt
                                         + R(cst[CST_MU0 + 1])) * tt + R(cst[CST_MU0])));
            }
            if (DLDS) sDUV[10 + s5(m5, -2)][x] = R(p_rp - sP[s2(rp)][PAD + x - 1]);
            else dpj.push(R(p_rp - sP[s2(rp)][PAD + x - 1]));
            const R Tm = sT[s2(rp)][PAD + x - 1];
            const R bv = bouss ? R(cst[CST_RI]) * ((T_rp + Tm) * R(0.5) - R(cst[CST_T_REF]))
                               : R(cst[CST_INVFR2]) * (R(1) - R(2) / (T_rp + Tm));
            if (DLDS) sDUV[15 + s5(m5, -2)][x] = bv;
            else bw.push(bv);
            if (!STEADY) {
                if (DLDS) { sDUV[s5(m5, -2)][x] = um_rp - ut[2]; sDUV[5 + s5(m5, -2)][x] = vm_rp - vt[2]; }
                else { du.push(um_rp - ut[2]); dv.push(vm_rp - vt[2]); }
            }
        }

        // ---------------- U stage K at row r = t - 2K (u update: i = 3..imax-1, j = 2..jmax-1)
#define U_STAGE(K, IN, RING_C, PREV, RING_OUT, PUSH)                                             \
        {                                                                                         \
            const int r = t - 2 * K;                                                              \
            const int bk = 2 * K - 2; /* back index of row r in the rows t-6..t-2 windows */      \
            R val = R(0), rs = R(0);                                                              \
            bool interior = false;                                                                \
            if (FULL || (r >= ilo - 2 * (3 - K) && r <= ihi + 2 * (3 - K))) {                     \
                if (FULL || (r >= 3 && r <= im - 1)) {                                            \
                    /* every lane computes (inputs outside the arrays hold finite defaults); */   \
                    /* columns outside the u update keep 0 */                                     \
                    {                                                                             \
                        interior = ucol;                                                          \
                        const R uP = IN(0), uW = IN(-1), uWW = IN(-2), uE = IN(1), uEE = IN(2);  \
                        const R *c = &RING_C[PAD + x];                                            \
                        const R uS = c[-1], uN = c[1];                                            \
                        const R iev = ie[bk], ev = ew[bk];                                        \
                        const R nu = MUPOLY ? nuw[bk] : ev * inv_re;                              \
                        const R aun_r = R(pr[4 * r]), rdx_r = R(pr[4 * r + 2]);                      \
                        const R cns = R(0.5) * aun_r * iev, cew = (RINGS ? R(0.5) * aue_j : haue_j) * iev; \
                        const R sE = uE + uP, sW = uP + uW, sS = uS + uP, sN = uP + uN;           \
                        const R fn = Svn[2 * K - 1] * cns;                                        \
                        const R fs = (vt[2 * K] + vt[2 * K + 1]) * cns;                           \
                        const R fe = sE * cew, fw = sW * cew;                                     \
                        const R diff = (nu * aue_j) * (R(pr[4 * r + 3]) * (uW - uP) + R(pr[4 * r + 7]) * (uE - uP)) \
                                     + (nu * aun_r) * (rdy_j * (uS - uP) + rdy_n * (uN - uP));    \
                        R pw_, pe_, ps_, pn_;                                                     \
                        if (u_upw_col) {                                                          \
                            pw_ = fw > R(0) ? uW : uP; pe_ = fe > R(0) ? uP : uE;                 \
                            ps_ = fs > R(0) ? uS : uP; pn_ = fn > R(0) ? uP : uN;                 \
                        } else {                                                                  \
                            const R uSS = c[-2], uNN = c[2];                                      \
                            const R cW = uWW - R(2) * uW + uP, cPi = uW - R(2) * uP + uE, cE = uP - R(2) * uE + uEE; \
                            const R cS = uSS - R(2) * uS + uP, cPj = uS - R(2) * uP + uN, cN = uP - R(2) * uN + uNN; \
                            pw_ = quick_face(fw, sW, cW, cPi); pe_ = quick_face(fe, sE, cPi, cE); \
                            ps_ = quick_face(fs, sS, cS, cPj); pn_ = quick_face(fn, sN, cPj, cN); \
                        }                                                                         \
                        const R conv = fw * pw_ - fe * pe_ + fs * ps_ - fn * pn_;                 \
                        R Rr = rdx_r * (rdy_j * (diff + conv) - ev * R(pw[bk] - pw[bk + 1])) - gg * ev; \
                        const R k1 = k1w[bk], k2 = k2w[bk];                                       \
                        if (k1 != R(0) || k2 != R(0)) {                                           \
                            const R vP = vt[2 * K];                                               \
                            Rr -= uP * (k1 + k2 * csqrt(uP * uP + vP * vP));                   \
                        }                                                                         \
                        rs = ((STEADY ? R(0) : DLDS ? sDUV[s5(m5, -2 * K)][x] : du[bk]) + Rr * dt) * dtau; \
                        const R tau = ut[2 * K];                                                  \
                        R cv;                                                                     \
                        if (K == 1) cv = tau + rs;                                                \
                        else if (K == 2) cv = R(0.75) * tau + R(0.25) * (uP + rs);                \
                        else cv = R(1.0 / 3.0) * tau + R(2.0 / 3.0) * (uP + rs);                  \
                        val = ucol ? cv : R(0);                                                   \
                    }                                                                             \
                } else if (r <= 2) {                                                              \
                    val = chan ? R(cst[CST_U_IN]) : R(0);                                                \
                } else {                                                                          \
                    val = (chan && ucol) ? PREV : R(0);                                           \
                }                                                                                 \
                RING_OUT                                                                          \
            }                                                                                     \
            PUSH                                                                                  \
            if (K == 3 && interior && out_col && (FULL || (r >= ilo && r < ihi))) {               \
                a.un[(size_t)r * su + j] = val;                                                   \
                const unsigned long long bb = dbits((double)rs); mu_ = bb > mu_ ? bb : mu_; if (L2) qu_ += (double)rs * (double)rs;       \
            }                                                                                     \
        }

        // ---------------- V stage K at row r = t - 2K (v update: i = 2..imax-1, j = 3..jmax-1)
#define V_STAGE(K, IN, RING_C, PREV, RING_OUT, PUSH, SET_ROW1)                                   \
        {                                                                                         \
            const int r = t - 2 * K;                                                              \
            const int bk = 2 * K - 2;                                                             \
            R val = R(0), rs = R(0);                                                              \
            bool interior = false;                                                                \
            if (FULL || (r >= ilo - 2 * (3 - K) && r <= ihi - 1 + 2 * (3 - K))) {                 \
                if (FULL || (r >= 2 && r <= im - 1)) {                                            \
                    /* every lane computes; columns outside the v update keep 0, the inflow */    \
                    /* columns j <= 2 take their boundary value below */                         \
                    {                                                                             \
                        interior = vcol;                                                          \
                        const R vP = IN(0), vW = IN(-1), vWW = IN(-2), vE = IN(1), vEE = IN(2);  \
                        const R *c = &RING_C[PAD + x];                                            \
                        const R vS = c[-1], vN = c[1];                                            \
                        const R iev = ie[bk], ev = ew[bk];                                        \
                        const R nu = MUPOLY ? nuw[bk] : ev * inv_re;                              \
                        const R avn_r = R(pr[4 * r + 1]), rdx_r = R(pr[4 * r + 2]);                 \
                        const R cns = R(0.5) * avn_r * iev, cew = (RINGS ? R(0.5) * ave_j : have_j) * iev; \
                        const R sN = vP + vN, sS = vP + vS, sE = vE + vP, sW = vP + vW;           \
                        const R fn = sN * cns, fs = sS * cns;                                     \
                        const R fe = SU(2 * K - 1) * cew, fw = SU(2 * K) * cew;                   \
                        const R diff = (nu * ave_j) * (rdx_r * (vW - vP) + R(pr[4 * r + 6]) * (vE - vP)) \
                                     + (nu * avn_r) * (rdym_j * (vS - vP) + rdym_n * (vN - vP));  \
                        R pw_, pe_, ps_, pn_;                                                     \
                        if (upwind_all || (!FULL && (r == 2 || r == im - 1))) {                   \
                            pw_ = fw > R(0) ? vW : vP; pe_ = fe > R(0) ? vP : vE;                 \
                            ps_ = fs > R(0) ? vS : vP; pn_ = fn > R(0) ? vP : vN;                 \
                        } else {                                                                  \
                            const R vSS = c[-2], vNN = c[2];                                      \
                            const R cW = vWW - R(2) * vW + vP, cPi = vW - R(2) * vP + vE, cE = vP - R(2) * vE + vEE; \
                            const R cS = vSS - R(2) * vS + vP, cPj = vS - R(2) * vP + vN, cN = vP - R(2) * vN + vNN; \
                            pw_ = quick_face(fw, sW, cW, cPi); pe_ = quick_face(fe, sE, cPi, cE); \
                            ps_ = quick_face(fs, sS, cS, cPj); pn_ = quick_face(fn, sN, cPj, cN); \
                        }                                                                         \
                        const R conv = fw * pw_ - fe * pe_ + fs * ps_ - fn * pn_;                 \
                        R Rr = rdy_j * (rdx_r * (diff + conv) - ev * (DLDS ? sDUV[10 + s5(m5, -2 * K)][x] : dpj[bk])) \
                                 + (DLDS ? sDUV[15 + s5(m5, -2 * K)][x] : bw[bk]);                       \
                        const R k1 = k1w[bk], k2 = k2w[bk];                                       \
                        if (k1 != R(0) || k2 != R(0)) {                                           \
                            const R uP = ut[2 * K];                                               \
                            Rr -= vP * (k1 + k2 * csqrt(uP * uP + vP * vP));                   \
                        }                                                                         \
                        rs = ((STEADY ? R(0) : DLDS ? sDUV[5 + s5(m5, -2 * K)][x] : dv[bk]) + Rr * dt) * dtau; \
                        const R tau = vt[2 * K];                                                  \
                        R cv;                                                                     \
                        if (K == 1) cv = tau + rs;                                                \
                        else if (K == 2) cv = R(0.75) * tau + R(0.25) * (vP + rs);                \
                        else cv = R(1.0 / 3.0) * tau + R(2.0 / 3.0) * (vP + rs);                  \
                        val = vcol ? cv : R(0);                                                   \
                        /* row 1 = row 2 (boundary.f90:34-36), in the v update's columns */       \
                        if (!FULL && r == 2 && !chan) { SET_ROW1; }                               \
                    }                                                                             \
                } else if (r == im) {                                                             \
                    val = vcol ? PREV : R(0);                                                     \
                }                                                                                 \
                if (j <= 2) val = chan ? R(0) : vi;                                               \
                RING_OUT                                                                          \
            }                                                                                     \
            PUSH                                                                                  \
            if (K == 3 && interior && out_col && (FULL || (r >= ilo && r < ihi))) {               \
                a.vn[(size_t)r * sv + j] = val;                                                   \
                const unsigned long long bb = dbits((double)rs); mv_ = bb > mv_ ? bb : mv_; if (L2) qv_ += (double)rs * (double)rs;       \
            }                                                                                     \
        }

        // u_tau(row t - q, j) + u_tau(row t - q, j - 1)
#define SU(q) (RINGS ? ut[q] + sUT[UTS(-(q))][PAD + x - 1] : Su[q])
        if (RINGS) {
            // stage 1 reads the tau windows (rows t..t-4 = r+2..r-2), stages 2 and 3 the
            // rings of the stage before, in the lane's own column; PREV is the stage's own
            // previous row
#define TAU_U(q) ut[2 - (q)]
#define TAU_V(q) vt[2 - (q)]
#define RU1(q) sU1[STS(-4 + (q))][PAD + x]
#define RU2(q) sU2[STS(-6 + (q))][PAD + x]
#define RV1(q) sV1[STS(-4 + (q))][PAD + x]
#define RV2(q) sV2[STS(-6 + (q))][PAD + x]
            U_STAGE(1, TAU_U, sUT[UTS(-2)], sU1[STS(-3)][PAD + x], { sU1[STS(-2)][PAD + x] = val; }, )
            V_STAGE(1, TAU_V, sVT[s4(t - 2)], sV1[STS(-3)][PAD + x], { sV1[STS(-2)][PAD + x] = val; }, ,
                    (sV1[STS(-3)][PAD + x] = vcol ? cv : sV1[STS(-3)][PAD + x]))
            U_STAGE(2, RU1, sU1[STS(-4)], sU2[STS(-5)][PAD + x], { sU2[STS(-4)][PAD + x] = val; }, )
            V_STAGE(2, RV1, sV1[STS(-4)], sV2[STS(-5)][PAD + x], { sV2[STS(-4)][PAD + x] = val; }, ,
                    (sV2[STS(-5)][PAD + x] = vcol ? cv : sV2[STS(-5)][PAD + x]))
            U_STAGE(3, RU2, sU2[STS(-6)], w3[0], {}, w3.push(val);)
            V_STAGE(3, RV2, sV2[STS(-6)], x3[0], { sV3[s2(t - 6)][PAD + x] = val; }, x3.push(val);,
                    (x3.set(0, vcol ? cv : x3[0])))
#undef RU1
#undef RU2
#undef RV1
#undef RV2
        } else {
            // register windows: the lane's own column of the stage before comes from them
#define RW1(q) w1[2 - (q)]
#define RW2(q) w2[2 - (q)]
#define RX1(q) x1[2 - (q)]
#define RX2(q) x2[2 - (q)]
            U_STAGE(1, TAU_U, sUT[UTS(-2)], w1[0], { sU1[STS(-2)][PAD + x] = val; }, w1.push(val);)
            V_STAGE(1, TAU_V, sVT[s4(t - 2)], x1[0], { sV1[STS(-2)][PAD + x] = val; }, x1.push(val);,
                    (x1.set(0, vcol ? cv : x1[0])))
            U_STAGE(2, RW1, sU1[STS(-4)], w2[0], { sU2[STS(-4)][PAD + x] = val; }, w2.push(val);)
            V_STAGE(2, RX1, sV1[STS(-4)], x2[0], { sV2[STS(-4)][PAD + x] = val; }, x2.push(val);,
                    (x2.set(0, vcol ? cv : x2[0])))
            U_STAGE(3, RW2, sU2[STS(-6)], w3[0], {}, w3.push(val);)
            V_STAGE(3, RX2, sV2[STS(-6)], x3[0], { sV3[s2(t - 6)][PAD + x] = val; }, x3.push(val);,
                    (x3.set(0, vcol ? cv : x3[0])))
#undef RW1
#undef RW2
#undef RX1
#undef RX2
        }
#undef TAU_U
#undef TAU_V
#undef SU
#undef UTS
#undef STS
#undef U_STAGE
#undef V_STAGE

        // ---------------- P at row r = t - 7 (equations.f90:620-658), FP64 accumulation
        {
            const int r = t - 7;
            if ((FULL || (r >= 2 && r <= im - 1 && r >= ilo && r < ihi)) && ucol && out_col) {
                const double u3p = (double)w3[1], u3e = (double)w3[0];
                const double v3p = (double)x3[1];
                const double v3n = (double)lds_v(sV3[s2(r)], x, j, 1, jm, chan, vi);
                const double avn_r = pr[4 * r + 1];
                const double rpv = -((u3e * aue_jd - u3p * aue_jd) + (v3n * avn_r - v3p * avn_r));
                const double dd = a.dtau * rpv * cst[CST_BETA];
                const double pp = pw[5];
                double pi = pp + dd;
                pi = 0.75 * pp + 0.25 * (pi + dd);
                a.pn[(size_t)r * su + j] = 1.0 / 3.0 * pp + 2.0 / 3.0 * (pi + dd);
                const unsigned long long bb = dbits(dd); mp_ = bb > mp_ ? bb : mp_; if (L2) qp_ += dd * dd;
            }
        }
    };
    // Groups whose rows all lie in [ilo + 7, min(im, ihi + 5)] run the form with the row
    // guards compiled out (every stage, the P update and the stores are in range there); the
    // first and last groups of a block keep them. The arithmetic of a row is the same
    // expressions in both forms (the file is compiled with -ffp-contract=on, so each
    // expression's FMA contraction is fixed by the source), so a row gives the same bits
    // whichever form computes it.
    const int tfull_lo = ilo + 7, tfull_hi = min(im, ihi + 5);
    using full_t = std::integral_constant<bool, true>;
    using guarded_t = std::integral_constant<bool, false>;
    static_assert(UNR == 2, "a group is a row pair");
    auto group = [&](const int t, auto fullc) __attribute__((always_inline)) {
        row(t, fullc);
        row(t + 1, fullc);
    };
    if (PEEL) {
        int t = t0;
        for (; t <= ihi + 6 && !(t >= tfull_lo && t + UNR - 1 <= tfull_hi); t += UNR) group(t, guarded_t{});
        for (; t + UNR - 1 <= tfull_hi; t += UNR) group(t, full_t{});
        for (; t <= ihi + 6; t += UNR) group(t, guarded_t{});
    } else {
        for (int t = t0; t <= ihi + 6; t += UNR) group(t, guarded_t{});
    }
    fold_max(a.res + 0, mu_);
    fold_max(a.res + 1, mv_);
    fold_max(a.res + 2, mp_);
    if (L2 && a.l2) {
        fold_sum(a.l2 + 0, qu_);
        fold_sum(a.l2 + 1, qv_);
        fold_sum(a.l2 + 2, qp_);
    }
}

// OCC: waves per SIMD the register budget is sized for (1: no spills for FP64; 2: 256
// registers). The autotuner picks the faster one per grid. L2: the form that also folds the
// sums of squares (contexts that asked for the L2 norms; it costs the two-wave form spills).
template <typename R, bool STEADY, bool MUPOLY, int OCC, int NT, bool L2>
__global__ __launch_bounds__(NT, OCC) void k_uvp(const K1Args<R> a)
{
    __shared__ Smem1<R, NT, (OCC > 1)> sm;
    const int G = a.nbi * a.nbj, per = (G + 7) / 8;
    const int b = blockIdx.x, L = (b % 8) * per + b / 8;   // blocks of one XCD are neighbours
    if (a.res_next && b == 0 && threadIdx.x < 3) a.res_next[threadIdx.x] = 0ull;
    if (L >= G) return;
    // issue priority over co-resident waves: in the two-stream step the T/C kernel of the
    // previous pseudo-step shares the SIMDs, and this kernel is the critical path
    __builtin_amdgcn_s_setprio(3);
    uvp_tile<R, STEADY, MUPOLY, NT, (OCC > 1), L2>(a, L / a.nbj, L % a.nbj, sm);
}

// Persistent batch: `cases` holds two argument sets per case (even and odd tau parity);
// ctl[c] bit 0 = active, bit 1 = parity. Every block walks the tile list with a stride of
// the grid size; inactive cases are skipped.
template <typename R, bool STEADY, bool MUPOLY, int OCC, int NT>
__global__ __launch_bounds__(NT, OCC) void k_uvp_batch(const K1Args<R> *cases, const BatchTile *tiles,
                                                       int ntiles, const int *ctl)
{
    __shared__ Smem1<R, NT, (OCC > 1)> sm;
    for (int w = blockIdx.x; w < ntiles; w += gridDim.x) {
        const BatchTile tl = tiles[w];
        const int cs = ctl[tl.c];
        if (!(cs & 1)) continue;
        const K1Args<R> a = cases[2 * tl.c + ((cs >> 1) & 1)];
        uvp_tile<R, STEADY, MUPOLY, NT, (OCC > 1), false>(a, tl.bi, tl.bj, sm);
        __syncthreads();   // the next tile re-initialises the LDS rings
    }
}

// ---------------------------------------------------------------------------------------
// K1, split chains: the u chain and the v chain of a tile in separate waves.
//
// The three u stages read v only at the tau level (and v likewise u), so the two chains are
// independent until the pressure update. A 2 NT-thread workgroup runs the u chain in the
// first NT lanes and the v chain in the second NT over the same NT columns; each chain holds
// only its own register windows (no spills, two waves per SIMD), loads its own inputs, and
// reads the other field's tau rows from deeper LDS rings (u: 8 rows, v: 16 rows). The
// pressure update runs in the u chain with v3 from the v chain's ring. The expressions are
// those of uvp_tile; with FMA contraction the compiler pairs some products differently, so
// the two forms agree to a few ulps.
// ---------------------------------------------------------------------------------------

template <typename R, int NT>
struct Smem1S {
    R sUT[8][lrow(NT)], sVT[16][lrow(NT)], sT[2][lrow(NT)];
    R sU1[3][lrow(NT)], sU2[3][lrow(NT)], sV1[3][lrow(NT)], sV2[3][lrow(NT)], sV3[2][lrow(NT)];
    double sP[2][lrow(NT)];
};

// The v ring is read down to row t-7 by the u stages while a v wave may already write row
// t+1, so it holds 16 rows; the u ring is read down to row t-6 and holds 8.
__device__ __forceinline__ int s16(int r) { return (r + 32) & 15; }

template <typename R, bool STEADY, bool MUPOLY, int NT, bool L2>
__device__ __forceinline__ void uvp_tile_split(const K1Args<R> &a, int bi, int bj, Smem1S<R, NT> &sm)
{
    const int im = a.imax, jm = a.jmax;
    const int ilo = 2 + bi * a.hb, ihi = min(ilo + a.hb, im);
    const int jlo = 2 + bj * w1(NT), jhi = min(jlo + w1(NT), jm);
    const int x = threadIdx.x & (NT - 1);
    const bool uchain = threadIdx.x < NT;
    const int j = jlo - HL1 + x;
    const bool chan = a.chan != 0, upwind_all = a.upwind_all != 0, bouss = a.bouss != 0;
    const R dt = R(a.dt), dtau = R(a.dtau), vi = R(a.v_i), inv_re = R(a.inv_re), gg = R(a.g);
    const int RR = im + 2, Cc = jm + 2;
    const double *pr = a.pk, *pc = a.pk + 4 * RR;
    auto sUT = sm.sUT; auto sVT = sm.sVT; auto sT = sm.sT; auto sP = sm.sP;
    auto sU1 = sm.sU1; auto sU2 = sm.sU2; auto sV1 = sm.sV1; auto sV2 = sm.sV2; auto sV3 = sm.sV3;

    for (int k = threadIdx.x; k < lrow(NT); k += 2 * NT) {
        for (int q = 0; q < 8; ++q) sUT[q][k] = R(0);
        for (int q = 0; q < 16; ++q) sVT[q][k] = R(0);
        for (int q = 0; q < 3; ++q) { sU1[q][k] = R(0); sU2[q][k] = R(0); sV1[q][k] = R(0); sV2[q][k] = R(0); }
        for (int q = 0; q < 2; ++q) { sP[q][k] = 0.0; sT[q][k] = R(1); sV3[q][k] = R(0); }
    }

    const bool jvalid = j >= 1 && j <= jm + 1;
    const int jc = min(max(j, 2), jm);
    const double aue_jd = pc[min(max(j, 1), jm)];
    const bool ucol = j >= 2 && j <= jm - 1;
    const bool vcol = j >= 3 && j <= jm - 1;
    const bool u_upw_col = upwind_all || j == 2 || j == jm - 1;
    const bool out_col = j >= jlo && j < jhi;

    Win<7, R> own;                      // own tau field, rows t-6..t
    Win<6, double> aux;                 // u chain: p rows t-7..t-2; v chain: dp/dj rows t-6..t-2
    Win<5, R> ie, ew, nuw, k1w, k2w, dd;   // per-cell data rows t-6..t-2 (dd: du or dv)
    Win<5, R> bw;                       // buoyancy (v chain)
    Win<5, R> w1, w2;                   // own stage 1 and 2 windows
    Win<2, R> w3;                       // own stage 3
    own.fill(R(0)); aux.fill(0.0);
    ie.fill(R(1)); ew.fill(R(1)); k1w.fill(R(0)); k2w.fill(R(0)); bw.fill(R(0));
    if (MUPOLY) nuw.fill(R(0));
    if (!STEADY) dd.fill(R(0));
    w1.fill(R(0)); w2.fill(R(0)); w3.fill(R(0));
    unsigned long long mu_ = 0, mv_ = 0, mp_ = 0;
    double qu_ = 0.0, qv_ = 0.0, qp_ = 0.0;        // sums of r3^2 (FP64 in both precisions)

    const size_t su = (size_t)(jm + 1), sv = (size_t)(jm + 2);
    R q_t, q_T, q_e, q_L, q_m;
    double q_p;
    auto fetch = [&](int t) {
        q_t = R(0); q_T = R(1); q_e = R(1); q_L = R(0); q_m = R(0); q_p = 0.0;
        if (jvalid && t >= 1) {
            if (uchain) { if (t <= im + 1 && j <= jm) q_t = a.ut[(size_t)t * su + j]; }
            else if (t <= im) q_t = a.vt[(size_t)t * sv + j];
        }
        const int rp = t - 2;
        if (jvalid && rp >= 1 && rp <= im) {
            if (j <= jm) {
                const size_t k = (size_t)rp * su + j;
                q_e = a.eps[k]; q_L = a.lp[k];
                if (uchain) { q_p = a.p[k]; if (MUPOLY) q_T = a.T[k]; }
                else q_T = a.T[k];
                if (!STEADY && uchain) q_m = a.um[k];
            }
            if (!STEADY && !uchain) q_m = a.vm[(size_t)rp * sv + j];
        }
    };
    __syncthreads();
    fetch(ilo - 6);

    for (int t = ilo - 6; t <= ihi + 6; ++t) {
        const R t_new = q_t, T_own = q_T, e = q_e, Lp = q_L, m_rp = q_m;
        const double p_rp = q_p;
        fetch(t + 1);
        const int rp = t - 2;
        own.push(t_new);
        if (uchain) {
            sUT[s8(t)][PAD + x] = t_new;
            sP[s2(rp)][PAD + x] = p_rp;
        } else {
            sVT[s16(t)][PAD + x] = t_new;
            sT[s2(rp)][PAD + x] = T_own;
        }
        __syncthreads();

        if (!chan) {
            if (uchain) { fix_u(sU1[s3(t - 3)], x, j, jm); fix_u(sU2[s3(t - 5)], x, j, jm); }
            else { fix_v(sV1[s3(t - 3)], x, j, jm); fix_v(sV2[s3(t - 5)], x, j, jm); }
        }
        {
            const R T_rp = uchain ? (MUPOLY ? T_own : R(1)) : T_own;
            R iev = R(1), k1 = R(0), k2 = R(0);
            if (rp >= ilo - 4 && rp <= ihi + 4) {
                iev = e == R(1) ? R(1) : R(1) / e;
                if (Lp != R(0)) {
                    k1 = e * Lp * R(a.inv_reda);
                    k2 = Lp * R(a.cfsd) * csqrt(e);
                }
            }
            ie.push(iev); ew.push(e); k1w.push(k1); k2w.push(k2);
            if (MUPOLY) {
                const R tt = T_rp;
                nuw.push(e * inv_re * ((((R(a.mu[4]) * tt + R(a.mu[3])) * tt + R(a.mu[2])) * tt + R(a.mu[1])) * tt + R(a.mu[0])));
            }
            if (uchain) {
                aux.push(p_rp);
            } else {
                const double pj = sP[s2(rp)][PAD + x];
                aux.push((double)R(pj - sP[s2(rp)][PAD + x - 1]));
                const R Tm = sT[s2(rp)][PAD + x - 1];
                bw.push(bouss ? R(a.ri) * ((T_rp + Tm) * R(0.5) - R(a.t_ref)) : R(a.invfr2) * (R(1) - R(2) / (T_rp + Tm)));
            }
            if (!STEADY) dd.push(m_rp - own[2]);
        }

        if (uchain) {
            const R haue_j = R(0.5 * pc[min(max(j, 1), jm)]);
            const R aue_j = R(aue_jd);
            const R rdy_j = R(pc[2 * Cc + jc]), rdy_n = R(pc[2 * Cc + min(jc + 1, jm)]);
#define US_STAGE(K, WP, RING_C, WOUT, RING_OUT)                                                  \
            {                                                                                     \
                const int r = t - 2 * K;                                                          \
                const int bk = 2 * K - 2;                                                         \
                R val = R(0), rs = R(0);                                                          \
                bool interior = false;                                                            \
                if (r >= ilo - 2 * (3 - K) && r <= ihi + 2 * (3 - K)) {                           \
                    if (r >= 3 && r <= im - 1) {                                                  \
                        if (ucol) {                                                               \
                            interior = true;                                                      \
                            const R uP = WP[2], uW = WP[3], uWW = WP[4], uE = WP[1], uEE = WP[0]; \
                            const R *c = &RING_C[PAD + x];                                        \
                            const R uS = c[-1], uN = c[1];                                        \
                            const R iev = ie[bk], ev = ew[bk];                                    \
                            const R nu = MUPOLY ? nuw[bk] : ev * inv_re;                          \
                            const R aun_r = R(pr[4 * r]), rdx_r = R(pr[4 * r + 2]);                  \
                            const R cns = R(0.5) * aun_r * iev, cew = haue_j * iev;               \
                            const R sE = uE + uP, sW = uP + uW, sS = uS + uP, sN = uP + uN;       \
                            const R vr = sVT[s16(r)][PAD + x], vr1 = sVT[s16(r - 1)][PAD + x];    \
                            const R fn = (sVT[s16(r)][PAD + x + 1] + sVT[s16(r - 1)][PAD + x + 1]) * cns; \
                            const R fs = (vr + vr1) * cns;                                        \
                            const R fe = sE * cew, fw = sW * cew;                                 \
                            const R diff = (nu * aue_j) * (R(pr[4 * r + 3]) * (uW - uP) + R(pr[4 * r + 7]) * (uE - uP)) \
                                         + (nu * aun_r) * (rdy_j * (uS - uP) + rdy_n * (uN - uP)); \
                            R pw_, pe_, ps_, pn_;                                                 \
                            if (u_upw_col) {                                                      \
                                pw_ = fw > R(0) ? uW : uP; pe_ = fe > R(0) ? uP : uE;             \
                                ps_ = fs > R(0) ? uS : uP; pn_ = fn > R(0) ? uP : uN;             \
                            } else {                                                              \
                                const R uSS = c[-2], uNN = c[2];                                  \
                                const R cW = uWW - R(2) * uW + uP, cPi = uW - R(2) * uP + uE, cE = uP - R(2) * uE + uEE; \
                                const R cS = uSS - R(2) * uS + uP, cPj = uS - R(2) * uP + uN, cN = uP - R(2) * uN + uNN; \
                                pw_ = quick_face(fw, sW, cW, cPi); pe_ = quick_face(fe, sE, cPi, cE); \
                                ps_ = quick_face(fs, sS, cS, cPj); pn_ = quick_face(fn, sN, cPj, cN); \
                            }                                                                     \
                            const R conv = fw * pw_ - fe * pe_ + fs * ps_ - fn * pn_;             \
                            R Rr = rdx_r * (rdy_j * (diff + conv) - ev * R(aux[bk] - aux[bk + 1])) - gg * ev; \
                            const R k1 = k1w[bk], k2 = k2w[bk];                                   \
                            if (k1 != R(0) || k2 != R(0))                                         \
                                Rr -= uP * (k1 + k2 * csqrt(uP * uP + vr * vr));                   \
                            rs = ((STEADY ? R(0) : dd[bk]) + Rr * dt) * dtau;                     \
                            const R tau = own[2 * K];                                             \
                            if (K == 1) val = tau + rs;                                           \
                            else if (K == 2) val = R(0.75) * tau + R(0.25) * (uP + rs);           \
                            else val = R(1.0 / 3.0) * tau + R(2.0 / 3.0) * (uP + rs);             \
                        }                                                                         \
                    } else if (r <= 2) {                                                          \
                        val = chan ? R(a.u_in) : R(0);                                            \
                    } else {                                                                      \
                        val = (chan && ucol) ? WOUT[0] : R(0);                                    \
                    }                                                                             \
                    RING_OUT                                                                      \
                }                                                                                 \
                WOUT.push(val);                                                                   \
                if (K == 3 && interior && out_col && r >= ilo && r < ihi) {                       \
                    a.un[(size_t)r * su + j] = val;                                               \
                    const unsigned long long bb = dbits((double)rs); mu_ = bb > mu_ ? bb : mu_; if (L2) qu_ += (double)rs * (double)rs;   \
                }                                                                                 \
            }
            US_STAGE(1, own, sUT[s8(t - 2)], w1, { sU1[s3(t - 2)][PAD + x] = val; })
            US_STAGE(2, w1, sU1[s3(t - 4)], w2, { sU2[s3(t - 4)][PAD + x] = val; })
            US_STAGE(3, w2, sU2[s3(t - 6)], w3, {})
#undef US_STAGE
            // P at row r = t - 7 (equations.f90:620-658), FP64 accumulation, v3 from the v chain
            const int r = t - 7;
            if (r >= 2 && r <= im - 1 && ucol && out_col && r >= ilo && r < ihi) {
                const double u3p = (double)w3[1], u3e = (double)w3[0];
                const double v3p = (double)sV3[s2(r)][PAD + x];
                const double v3n = (double)lds_v(sV3[s2(r)], x, j, 1, jm, chan, vi);
                const double avn_r = pr[4 * r + 1];
                const double rpv = -((u3e * aue_jd - u3p * aue_jd) + (v3n * avn_r - v3p * avn_r));
                const double ddp = a.dtau * rpv * a.beta;
                const double pp = aux[5];
                double pi = pp + ddp;
                pi = 0.75 * pp + 0.25 * (pi + ddp);
                a.pn[(size_t)r * su + j] = 1.0 / 3.0 * pp + 2.0 / 3.0 * (pi + ddp);
                const unsigned long long bb = dbits(ddp); mp_ = bb > mp_ ? bb : mp_; if (L2) qp_ += ddp * ddp;
            }
        } else {
            const R ave_j = R(pc[Cc + jc]), have_j = R(0.5 * pc[Cc + jc]);
            const R rdy_j = R(pc[2 * Cc + jc]);
            const R rdym_j = R(pc[3 * Cc + jc]), rdym_n = R(pc[3 * Cc + min(jc + 1, jm + 1)]);
#define VS_STAGE(K, WP, RING_C, WOUT, RING_OUT)                                                  \
            {                                                                                     \
                const int r = t - 2 * K;                                                          \
                const int bk = 2 * K - 2;                                                         \
                R val = R(0), rs = R(0);                                                          \
                bool interior = false;                                                            \
                if (r >= ilo - 2 * (3 - K) && r <= ihi - 1 + 2 * (3 - K)) {                       \
                    if (j <= 2) {                                                                 \
                        val = chan ? R(0) : vi;                                                   \
                    } else if (r >= 2 && r <= im - 1) {                                           \
                        if (vcol) {                                                               \
                            interior = true;                                                      \
                            const R vP = WP[2], vW = WP[3], vWW = WP[4], vE = WP[1], vEE = WP[0]; \
                            const R *c = &RING_C[PAD + x];                                        \
                            const R vS = c[-1], vN = c[1];                                        \
                            const R iev = ie[bk], ev = ew[bk];                                    \
                            const R nu = MUPOLY ? nuw[bk] : ev * inv_re;                          \
                            const R avn_r = R(pr[4 * r + 1]), rdx_r = R(pr[4 * r + 2]);             \
                            const R cns = R(0.5) * avn_r * iev, cew = have_j * iev;               \
                            const R sN = vP + vN, sS = vP + vS, sE = vE + vP, sW = vP + vW;       \
                            const R fn = sN * cns, fs = sS * cns;                                 \
                            const R ur = sUT[s8(r)][PAD + x];                                     \
                            const R fe = (sUT[s8(r + 1)][PAD + x] + sUT[s8(r + 1)][PAD + x - 1]) * cew; \
                            const R fw = (ur + sUT[s8(r)][PAD + x - 1]) * cew;                    \
                            const R diff = (nu * ave_j) * (rdx_r * (vW - vP) + R(pr[4 * r + 6]) * (vE - vP)) \
                                         + (nu * avn_r) * (rdym_j * (vS - vP) + rdym_n * (vN - vP)); \
                            R pw_, pe_, ps_, pn_;                                                 \
                            if (upwind_all || r == 2 || r == im - 1) {                            \
                                pw_ = fw > R(0) ? vW : vP; pe_ = fe > R(0) ? vP : vE;             \
                                ps_ = fs > R(0) ? vS : vP; pn_ = fn > R(0) ? vP : vN;             \
                            } else {                                                              \
                                const R vSS = c[-2], vNN = c[2];                                  \
                                const R cW = vWW - R(2) * vW + vP, cPi = vW - R(2) * vP + vE, cE = vP - R(2) * vE + vEE; \
                                const R cS = vSS - R(2) * vS + vP, cPj = vS - R(2) * vP + vN, cN = vP - R(2) * vN + vNN; \
                                pw_ = quick_face(fw, sW, cW, cPi); pe_ = quick_face(fe, sE, cPi, cE); \
                                ps_ = quick_face(fs, sS, cS, cPj); pn_ = quick_face(fn, sN, cPj, cN); \
                            }                                                                     \
                            const R conv = fw * pw_ - fe * pe_ + fs * ps_ - fn * pn_;             \
                            R Rr = rdy_j * (rdx_r * (diff + conv) - ev * R(aux[bk])) + bw[bk];        \
                            const R k1 = k1w[bk], k2 = k2w[bk];                                   \
                            if (k1 != R(0) || k2 != R(0))                                         \
                                Rr -= vP * (k1 + k2 * csqrt(ur * ur + vP * vP));                   \
                            rs = ((STEADY ? R(0) : dd[bk]) + Rr * dt) * dtau;                     \
                            const R tau = own[2 * K];                                             \
                            if (K == 1) val = tau + rs;                                           \
                            else if (K == 2) val = R(0.75) * tau + R(0.25) * (vP + rs);           \
                            else val = R(1.0 / 3.0) * tau + R(2.0 / 3.0) * (vP + rs);             \
                            if (r == 2 && !chan) WOUT.set(0, val);                                \
                        }                                                                         \
                    } else if (r == im) {                                                         \
                        val = vcol ? WOUT[0] : R(0);                                              \
                    }                                                                             \
                    RING_OUT                                                                      \
                }                                                                                 \
                WOUT.push(val);                                                                   \
                if (K == 3 && interior && out_col && r >= ilo && r < ihi) {                       \
                    a.vn[(size_t)r * sv + j] = val;                                               \
                    const unsigned long long bb = dbits((double)rs); mv_ = bb > mv_ ? bb : mv_; if (L2) qv_ += (double)rs * (double)rs;   \
                }                                                                                 \
            }
            VS_STAGE(1, own, sVT[s16(t - 2)], w1, { sV1[s3(t - 2)][PAD + x] = val; })
            VS_STAGE(2, w1, sV1[s3(t - 4)], w2, { sV2[s3(t - 4)][PAD + x] = val; })
            VS_STAGE(3, w2, sV2[s3(t - 6)], w3, { sV3[s2(t - 6)][PAD + x] = val; })
#undef VS_STAGE
        }
    }
    fold_max(a.res + 0, mu_);
    fold_max(a.res + 1, mv_);
    fold_max(a.res + 2, mp_);
    if (L2 && a.l2) {
        fold_sum(a.l2 + 0, qu_);
        fold_sum(a.l2 + 1, qv_);
        fold_sum(a.l2 + 2, qp_);
    }
}

template <typename R, bool STEADY, bool MUPOLY, int NT, bool L2>
__global__ __launch_bounds__(2 * NT, 1) void k_uvp_split(const K1Args<R> a)
{
    __shared__ Smem1S<R, NT> sm;
    const int G = a.nbi * a.nbj, per = (G + 7) / 8;
    const int b = blockIdx.x, L = (b % 8) * per + b / 8;
    if (a.res_next && b == 0 && threadIdx.x < 3) a.res_next[threadIdx.x] = 0ull;
    if (L >= G) return;
    __builtin_amdgcn_s_setprio(3);      // as k_uvp
    uvp_tile_split<R, STEADY, MUPOLY, NT, L2>(a, L / a.nbj, L % a.nbj, sm);
}

template <typename R, bool STEADY, bool MUPOLY, int NT>
__global__ __launch_bounds__(2 * NT, 1) void k_uvp_split_batch(const K1Args<R> *cases, const BatchTile *tiles,
                                                               int ntiles, const int *ctl)
{
    __shared__ Smem1S<R, NT> sm;
    for (int w = blockIdx.x; w < ntiles; w += gridDim.x) {
        const BatchTile tl = tiles[w];
        const int cs = ctl[tl.c];
        if (!(cs & 1)) continue;
        const K1Args<R> a = cases[2 * tl.c + ((cs >> 1) & 1)];
        uvp_tile_split<R, STEADY, MUPOLY, NT, false>(a, tl.bi, tl.bj, sm);
        __syncthreads();
    }
}

// ---------------------------------------------------------------------------------------
// K2: T (energy, 1/Pe) and C (species, 1/(Re Sc)), three RK stages each.
// ---------------------------------------------------------------------------------------
constexpr int NT2 = 256;
constexpr int HL2 = 3, HR2 = 3;
__host__ __device__ constexpr int w2(int nt) { return nt - HL2 - HR2; }

template <typename R>
struct K2Args {
    const R *un, *vn, *T, *Tt, *C, *Ct, *eps, *lp;
    const double *y;
    R *Tn, *Cn;
    const double *pk;
    int imax, jmax, nbi, nbj, hb, chan;
    double dt, dtau, kt, kc, tinf, t_bottom, t_top, v_i, u_in, slot_lo, slot_hi, slot_value;
    double kcoef[5];             // k(T) polynomial (k_model CFD_K_POLY4): kappa = kt * k(T)
};

// k(T) of the energy diffusivity at the physical-level temperature t (the conductivity over
// its reference value; the reference's RESZ has the constant 1/Pe)
template <typename R>
__device__ __forceinline__ R kpoly(const double *c, R t)
{
    return (((R(c[4]) * t + R(c[3])) * t + R(c[2])) * t + R(c[1])) * t + R(c[0]);
}

template <typename R, int NT>
struct Smem2 {
    R sTT[4][lrow(NT)], sCT[4][lrow(NT)], sV[2][lrow(NT)];
    R sZ1[2][lrow(NT)], sZ2[2][lrow(NT)], sC1[2][lrow(NT)], sC2[2][lrow(NT)];
};

// T of the centre row at column j+dj (boundary.f90:85-88 / channel wall temperatures).
template <typename R>
__device__ __forceinline__ R lds_t(const R *row, int x, int j, int dj, int jm, bool chan, R tinf, R tb, R ttop)
{
    const int jj = j + dj;
    if (jj == 1) return chan ? tb : tinf;
    if (jj == jm) {
        if (chan) return ttop;
        const R a1 = row[PAD + x + dj - 1], a2 = row[PAD + x + dj - 2];
        return a1 + R(1) * (a1 - a2);
    }
    return row[PAD + x + dj];
}

// C of the centre row (boundary.f90:128-131 / channel: zero-gradient walls).
template <typename R>
__device__ __forceinline__ R lds_c(const R *row, int x, int j, int dj, int jm, bool chan)
{
    const int jj = j + dj;
    if (jj == 1) return chan ? row[PAD + x + dj + 1] : R(0);
    if (jj == jm) {
        if (chan) return row[PAD + x + dj - 1];
        const R a1 = row[PAD + x + dj - 1], a2 = row[PAD + x + dj - 2];
        return a1 + R(1) * (a1 - a2);
    }
    return row[PAD + x + dj];
}

// Neighbour values of the centre row with the boundary formulas as selects: zs/zn the ring
// values at j-1/j+1, zp the lane's own value (the extrapolation a1 + (a1 - a2) at jj = jm
// uses a1 = zp and a2 = the raw j-1 value).
template <typename R>
__device__ __forceinline__ void nb_t(R &zs, R &zn, R zp, int j, int jm, bool chan, R tinf, R tb, R ttop)
{
    const R ext = chan ? ttop : zp + R(1) * (zp - zs);
    const R s1 = chan ? tb : tinf;
    zn = (j + 1 == jm) ? ext : zn;
    zs = (j - 1 == 1) ? s1 : zs;
}
template <typename R>
__device__ __forceinline__ void nb_c(R &zs, R &zn, R zp, int j, int jm, bool chan)
{
    const R ext = chan ? zp : zp + R(1) * (zp - zs);
    const R s1 = chan ? zp : R(0);
    zn = (j + 1 == jm) ? ext : zn;
    zs = (j - 1 == 1) ? s1 : zs;
}

// KPOLY: the T chain's diffusivity is kt * k(T) per cell, T of the physical level (the
// iterate in steady mode), fixed through the three stages
template <typename R, bool STEADY, int NT, bool KPOLY>
__device__ __forceinline__ void scal_tile(const K2Args<R> &a, int bi, int bj, Smem2<R, NT> &sm)
{
    const int im = a.imax, jm = a.jmax;
    const int ilo = 2 + bi * a.hb, ihi = min(ilo + a.hb, im);
    const int jlo = 2 + bj * w2(NT), jhi = min(jlo + w2(NT), jm);
    const int x = threadIdx.x;
    const int j = jlo - HL2 + x;
    const bool chan = a.chan != 0;
    const R dt = R(a.dt), dtau = R(a.dtau), kt = R(a.kt), kc = R(a.kc);
    const R tinf = R(a.tinf), tb = R(a.t_bottom), ttop = R(a.t_top);
    const int RR = im + 2, Cc = jm + 2;
    const double *pr = a.pk, *pc = a.pk + 4 * RR;
    auto sTT = sm.sTT; auto sCT = sm.sCT; auto sV = sm.sV;
    auto sZ1 = sm.sZ1; auto sZ2 = sm.sZ2; auto sC1 = sm.sC1; auto sC2 = sm.sC2;

    for (int k = x; k < lrow(NT); k += NT) {
        for (int q = 0; q < 4; ++q) { sTT[q][k] = R(0); sCT[q][k] = R(0); }
        for (int q = 0; q < 2; ++q) { sV[q][k] = R(0); sZ1[q][k] = R(0); sZ2[q][k] = R(0); sC1[q][k] = R(0); sC2[q][k] = R(0); }
    }
    const bool jvalid = j >= 1 && j <= jm;
    const int jc = min(max(j, 2), jm - 1);
    const R aue_j = R(pc[min(max(j, 1), jm)]);
    const R rdy_j = R(pc[2 * Cc + jc]), rdy_n = R(pc[2 * Cc + jc + 1]);
    const R rdym_n = R(pc[3 * Cc + jc + 1]);
    const bool scol = j >= 2 && j <= jm - 1;
    const bool out_col = j >= jlo && j < jhi;
    const double yj = a.y[min(max(j, 1), jm)];
    const R cslot = chan ? ((yj >= a.slot_lo && yj <= a.slot_hi) ? R(a.slot_value) : R(0)) : R(0);

    Win<4, R> tt, ct, uw, vw, vnw, id, oml, dT, dC, kw;
    Win<3, R> z1, z2, c1, c2, z3, c3;
    tt.fill(R(0)); ct.fill(R(0)); uw.fill(R(0)); vw.fill(R(0)); vnw.fill(R(0)); id.fill(R(1)); oml.fill(R(1));
    if (KPOLY) kw.fill(kt);
    if (!STEADY) { dT.fill(R(0)); dC.fill(R(0)); }
    z1.fill(R(0)); z2.fill(R(0)); c1.fill(R(0)); c2.fill(R(0)); z3.fill(R(0)); c3.fill(R(0));

    const size_t su = (size_t)(jm + 1), sv = (size_t)(jm + 2);
    // two prefetch sets: the loads of row t go out two rows ahead (set t & 1)
    R qa_t, qa_c, qa_u, qa_v, qa_T, qa_C, qa_e, qa_L;
    R qb_t, qb_c, qb_u, qb_v, qb_T, qb_C, qb_e, qb_L;
    // next row: u3/v3 from K1's interior cells, boundary rows/columns from the BC formulas
    // Loads at clamped (always valid) indices, one site per field; the defaults of rows and
    // columns outside the arrays are applied where the values are consumed (a select at the
    // load would wait for it). u and v come from interior cells only: their boundary cells may
    // still be being written (the two-stream step runs K2 beside the u/v/p boundary pass).
    const int jc1 = min(max(j, 1), jm), jvi = min(max(j, 3), jm - 1), jui = min(max(j, 2), jm - 1);
#define K2_FETCH(Q, t)                                                                            \
    {                                                                                             \
        const size_t k_ = (size_t)min(max((t), 1), im) * su + jc1;                                \
        Q##_t = a.Tt[k_]; Q##_c = a.Ct[k_]; Q##_e = a.eps[k_]; Q##_L = a.lp[k_];                  \
        if (!STEADY) { Q##_T = a.T[k_]; Q##_C = a.C[k_]; } else { Q##_T = R(0); Q##_C = R(0); }  \
        Q##_v = a.vn[(size_t)min(max((t), 2), im - 1) * sv + jvi];                               \
        Q##_u = a.un[(size_t)min(max((t), 3), im - 1) * su + jui];                               \
    }
    // Rows run two at a time from an even row (as in K1: a row is computed by the same half of
    // the pair in every block); the extra row at either end has no stage in range.
    const int t0 = (ilo - 3) - ((ilo - 3) & 1);
    __syncthreads();
    K2_FETCH(qa, t0)
    K2_FETCH(qb, t0 + 1)

    // argument fields used in the row body, as locals (the row closure then holds no
    // reference into the argument block)
    const R v_i = R(a.v_i), u_in = R(a.u_in);
    R *const Tn_out = a.Tn, *const Cn_out = a.Cn;
    auto row = [&](const int t, R &q_t, R &q_c, R &q_u, R &q_v, R &q_T, R &q_C, R &q_e, R &q_L, auto fullc)
                   __attribute__((always_inline)) {
        // FULL: every row guard below holds (the rows of the pair lie in
        // [max(ilo + 3, 6), min(im - 1, ihi + 2)])
        constexpr bool FULL = decltype(fullc)::value;
        // ---------------- phase A: row t
        const bool inr = jvalid && (FULL || (t >= 1 && t <= im));
        const R tnew = inr ? q_t : R(0), cnew = inr ? q_c : R(0), e = inr ? q_e : R(1), Lp = inr ? q_L : R(0);
        const R Tp = (!STEADY && inr) ? q_T : R(0), Cp = (!STEADY && inr) ? q_C : R(0);
        const R vnew = !inr ? R(0) : (j == 2 ? (chan ? R(0) : v_i)
                                             : (((FULL || (t >= 2 && t <= im - 1)) && j >= 3 && j <= jm - 1) ? q_v : R(0)));
        R unew = (jvalid && (FULL || (t > 2 && t < im)) && j >= 2 && j <= jm - 1) ? q_u : R(0);
        {   // row t + 2 into the set just consumed
            const int tn = t + 2;
            const size_t k_ = (size_t)min(max(tn, 1), im) * su + jc1;
            q_t = a.Tt[k_]; q_c = a.Ct[k_]; q_e = a.eps[k_]; q_L = a.lp[k_];
            if (!STEADY) { q_T = a.T[k_]; q_C = a.C[k_]; }
            q_v = a.vn[(size_t)min(max(tn, 2), im - 1) * sv + jvi];
            q_u = a.un[(size_t)min(max(tn, 3), im - 1) * su + jui];
        }
        if (!FULL) {                                   // u3 boundary rows (boundary.f90:29-32,39-42)
            if (t <= 2) unew = chan ? u_in : R(0);
            else if (t >= im) unew = chan ? uw[0] : R(0);
        }
        tt.push(tnew); ct.push(cnew); uw.push(unew); vw.push(vnew);
        sTT[s4(t)][PAD + x] = tnew;
        sCT[s4(t)][PAD + x] = cnew;
        sV[s2(t)][PAD + x] = vnew;
        __syncthreads();

        // ---------------- phase B: per-cell data of row t
        {
            R vn_ = sV[s2(t)][PAD + x + 1];
            if (j + 1 == jm) vn_ = chan ? R(0) : R(2) * sV[s2(t)][PAD + x - 1] - sV[s2(t)][PAD + x];
            vnw.push(vn_);
            id.push(Lp != R(0) ? R(1) / (Lp * (e - R(1)) + R(1)) : R(1));
            oml.push(R(1) - Lp);
            if (!STEADY) { dT.push(Tp - tnew); dC.push(Cp - cnew); }
            if (KPOLY) kw.push(kt * kpoly(a.kcoef, STEADY ? tnew : Tp));
        }

        // stage K of a scalar at row r = t - K; newest own row of WP is t-K+1
#define SCAL_STAGE(K, WP, CROW, ISC, WOUT, RINGW, OUTP, KAPPA, DW, TAUW)                         \
        {                                                                                         \
            const int r = t - K;                                                                  \
            const int bq = K;            /* back index of row r in windows whose newest is t */   \
            R val = R(0);                                                                         \
            bool interior = false;                                                                \
            if (FULL || (r >= ilo - (3 - K) && r <= ihi - 1 + (3 - K))) {                         \
                if (FULL || (r >= 2 && r <= im - 1)) {                                            \
                    /* every lane computes; columns outside the update keep 0 */                 \
                    {                                                                             \
                        interior = scol;                                                          \
                        const R zE = WP[0], zP = WP[1], zW = WP[2];                               \
                        const R *c = CROW;                                                        \
                        R zS = c[PAD + x - 1], zN = c[PAD + x + 1];                               \
                        if (ISC) nb_c(zS, zN, zP, j, jm, chan);                                   \
                        else nb_t(zS, zN, zP, j, jm, chan, tinf, tb, ttop);                       \
                        const R avn_r = R(pr[4 * r + 1]);                                            \
                        const R diff = (KAPPA * aue_j) * (R(pr[4 * r + 6]) * (zE - zP) + R(pr[4 * r + 2]) * (zW - zP)) \
                                     + (KAPPA * avn_r) * (rdy_n * (zN - zP) + rdy_j * (zS - zP)); \
                        const R uE = uw[bq - 1], uP = uw[bq], vP = vw[bq], vN = vnw[bq];          \
                        const R conv = R(0.5) * ((zE + zP) * uE - (zW + zP) * uP) * aue_j         \
                                     + R(0.5) * ((zN + zP) * vN - (zS + zP) * vP) * avn_r;        \
                        const R vols = R(pr[4 * r + 7]) * rdym_n;                            \
                        const R Rr = ISC ? vols * (diff - conv * id[bq])                          \
                                         : vols * (diff - oml[bq] * conv) * id[bq];               \
                        const R rs = ((STEADY ? R(0) : DW[bq]) + Rr * dt) * dtau;                 \
                        const R tau = TAUW[bq];                                                   \
                        R cv;                                                                     \
                        if (K == 1) cv = tau + rs;                                                \
                        else if (K == 2) cv = R(0.75) * tau + R(0.25) * (zP + rs);                \
                        else cv = R(1.0 / 3.0) * tau + R(2.0 / 3.0) * (zP + rs);                  \
                        val = scol ? cv : R(0);                                                   \
                        if (!FULL && r == 2) WOUT.set(0, scol ? ((ISC && chan) ? cslot : cv) : WOUT[0]); /* row 1 */ \
                    }                                                                             \
                } else if (r == im) {                                                             \
                    val = scol ? WOUT[0] : R(0);                                                  \
                } else if (r == 1) {                                                              \
                    val = (ISC && chan) ? cslot : R(0);                                           \
                }                                                                                 \
                RINGW                                                                             \
            }                                                                                     \
            WOUT.push(val);                                                                       \
            if (OUTP != nullptr && interior && out_col && (FULL || (r >= ilo && r < ihi)))        \
                OUTP[(size_t)r * su + j] = val;                                                   \
        }

        R *const noout = nullptr;
        SCAL_STAGE(1, tt, sTT[s4(t - 1)], false, z1, { sZ1[s2(t - 1)][PAD + x] = val; }, noout, (KPOLY ? kw[bq] : kt), dT, tt)
        SCAL_STAGE(1, ct, sCT[s4(t - 1)], true, c1, { sC1[s2(t - 1)][PAD + x] = val; }, noout, kc, dC, ct)
        SCAL_STAGE(2, z1, sZ1[s2(t - 2)], false, z2, { sZ2[s2(t - 2)][PAD + x] = val; }, noout, (KPOLY ? kw[bq] : kt), dT, tt)
        SCAL_STAGE(2, c1, sC1[s2(t - 2)], true, c2, { sC2[s2(t - 2)][PAD + x] = val; }, noout, kc, dC, ct)
        SCAL_STAGE(3, z2, sZ2[s2(t - 3)], false, z3, {}, Tn_out, (KPOLY ? kw[bq] : kt), dT, tt)
        SCAL_STAGE(3, c2, sC2[s2(t - 3)], true, c3, {}, Cn_out, kc, dC, ct)
#undef SCAL_STAGE
    };
    // Steady problems: pairs whose rows all lie in [max(ilo + 3, 6), min(im - 1, ihi + 2)] run
    // the form with the row guards compiled out; a row gives the same bits in both forms
    // (same expressions, contraction fixed per expression). Unsteady problems keep the
    // guarded form throughout: the straight-line code needs more registers than the room a
    // T/C wave has beside a u/v/p wave in the two-stream step.
    const int tfull_lo = max(ilo + 3, 6), tfull_hi = min(im - 1, ihi + 2);
    auto pair = [&](const int t, auto fullc) __attribute__((always_inline)) {
        row(t, qa_t, qa_c, qa_u, qa_v, qa_T, qa_C, qa_e, qa_L, fullc);
        row(t + 1, qb_t, qb_c, qb_u, qb_v, qb_T, qb_C, qb_e, qb_L, fullc);
    };
    using full_t = std::integral_constant<bool, true>;
    using guarded_t = std::integral_constant<bool, false>;
    if (STEADY) {
        int t = t0;
        for (; t <= ihi + 2 && !(t >= tfull_lo && t + 1 <= tfull_hi); t += 2) pair(t, guarded_t{});
        for (; t + 1 <= tfull_hi; t += 2) pair(t, full_t{});
        for (; t <= ihi + 2; t += 2) pair(t, guarded_t{});
    } else {
        for (int t = t0; t <= ihi + 2; t += 2) pair(t, guarded_t{});
    }
#undef K2_FETCH
}

// ---------------------------------------------------------------------------------------
// K2, split chains: T in the first NT lanes, C in the second NT of a 2 NT-lane workgroup
// over the same NT columns. T and C do not interact, so each chain keeps only its own windows (about half
// the registers of scal_tile: more waves per SIMD, and room beside K1 in the two-stream
// step). Both chains read u, v, eps and L; the T chain publishes the v row for the j+1
// neighbour. Same expressions as scal_tile.
// ---------------------------------------------------------------------------------------

template <typename R, bool STEADY, int NT, bool KPOLY>
__device__ __forceinline__ void scal_tile_split(const K2Args<R> &a, int bi, int bj, Smem2<R, NT> &sm)
{
    const int im = a.imax, jm = a.jmax;
    const int ilo = 2 + bi * a.hb, ihi = min(ilo + a.hb, im);
    const int jlo = 2 + bj * w2(NT), jhi = min(jlo + w2(NT), jm);
    const int x = threadIdx.x & (NT - 1);
    const bool isC = threadIdx.x >= NT;
    const int j = jlo - HL2 + x;
    const bool chan = a.chan != 0;
    const R dt = R(a.dt), dtau = R(a.dtau), kt = R(a.kt), kc = R(a.kc);
    const R tinf = R(a.tinf), tb = R(a.t_bottom), ttop = R(a.t_top);
    const int RR = im + 2, Cc = jm + 2;
    const double *pr = a.pk, *pc = a.pk + 4 * RR;
    auto sTT = sm.sTT; auto sCT = sm.sCT; auto sV = sm.sV;
    auto sZ1 = sm.sZ1; auto sZ2 = sm.sZ2; auto sC1 = sm.sC1; auto sC2 = sm.sC2;

    for (int k = threadIdx.x; k < lrow(NT); k += 2 * NT) {
        for (int q = 0; q < 3; ++q) { sTT[q][k] = R(0); sCT[q][k] = R(0); }
        for (int q = 0; q < 2; ++q) { sV[q][k] = R(0); sZ1[q][k] = R(0); sZ2[q][k] = R(0); sC1[q][k] = R(0); sC2[q][k] = R(0); }
    }
    const bool jvalid = j >= 1 && j <= jm;
    const int jc = min(max(j, 2), jm - 1);
    const R aue_j = R(pc[min(max(j, 1), jm)]);
    const R rdy_j = R(pc[2 * Cc + jc]), rdy_n = R(pc[2 * Cc + jc + 1]);
    const R rdym_n = R(pc[3 * Cc + jc + 1]);
    const bool scol = j >= 2 && j <= jm - 1;
    const bool out_col = j >= jlo && j < jhi;
    const double yj = a.y[min(max(j, 1), jm)];
    const R cslot = chan ? ((yj >= a.slot_lo && yj <= a.slot_hi) ? R(a.slot_value) : R(0)) : R(0);
    const R kap = isC ? kc : kt;
    R *const outp = isC ? a.Cn : a.Tn;
    auto own_ring = isC ? sCT : sTT;
    auto ring1 = isC ? sC1 : sZ1;
    auto ring2 = isC ? sC2 : sZ2;

    Win<4, R> zt, uw, vw, vnw, id, oml, dZ, kw;
    Win<3, R> y1, y2, y3;
    zt.fill(R(0)); uw.fill(R(0)); vw.fill(R(0)); vnw.fill(R(0)); id.fill(R(1)); oml.fill(R(1));
    if (KPOLY) kw.fill(kap);
    if (!STEADY) dZ.fill(R(0));
    y1.fill(R(0)); y2.fill(R(0)); y3.fill(R(0));

    const size_t su = (size_t)(jm + 1), sv = (size_t)(jm + 2);
    R q_t, q_u, q_v, q_Z, q_e, q_L;
    auto fetch = [&](int t) {
        q_t = R(0); q_u = R(0); q_v = R(0); q_Z = R(0); q_e = R(1); q_L = R(0);
        if (jvalid && t >= 1 && t <= im + 1) {
            if (t <= im) {
                const size_t k = (size_t)t * su + j;
                q_t = isC ? a.Ct[k] : a.Tt[k];
                q_e = a.eps[k]; q_L = a.lp[k];
                if (!STEADY) q_Z = isC ? a.C[k] : a.T[k];
                if (j == 2) q_v = chan ? R(0) : R(a.v_i);
                else if (t >= 2 && t <= im - 1 && j >= 3 && j <= jm - 1) q_v = a.vn[(size_t)t * sv + j];
            }
            // interior cells only (see scal_tile)
            if (t > 2 && t < im && j >= 2 && j <= jm - 1) q_u = a.un[(size_t)t * su + j];
        }
    };
    __syncthreads();
    fetch(ilo - 3);

    for (int t = ilo - 3; t <= ihi + 2; ++t) {
        const R znew = q_t, vnew = q_v, Zp = q_Z, e = q_e, Lp = q_L;
        R unew = q_u;
        fetch(t + 1);
        if (t <= 2) unew = chan ? R(a.u_in) : R(0);
        else if (t >= im) unew = chan ? uw[0] : R(0);
        zt.push(znew); uw.push(unew); vw.push(vnew);
        own_ring[s3(t)][PAD + x] = znew;
        if (!isC) sV[s2(t)][PAD + x] = vnew;
        __syncthreads();
        {
            R vn_ = sV[s2(t)][PAD + x + 1];
            if (j + 1 == jm) vn_ = chan ? R(0) : R(2) * sV[s2(t)][PAD + x - 1] - sV[s2(t)][PAD + x];
            vnw.push(vn_);
            id.push(Lp != R(0) ? R(1) / (Lp * (e - R(1)) + R(1)) : R(1));
            oml.push(R(1) - Lp);
            if (!STEADY) dZ.push(Zp - znew);
            if (KPOLY) kw.push(isC ? kap : kap * kpoly(a.kcoef, STEADY ? znew : Zp));
        }
#define SCS_STAGE(K, WP, CROW, WOUT, RINGW, OUTP)                                                \
        {                                                                                         \
            const int r = t - K;                                                                  \
            const int bq = K;                                                                     \
            R val = R(0);                                                                         \
            bool interior = false;                                                                \
            if (r >= ilo - (3 - K) && r <= ihi - 1 + (3 - K)) {                                   \
                if (r >= 2 && r <= im - 1) {                                                      \
                    if (scol) {                                                                   \
                        interior = true;                                                          \
                        const R zE = WP[0], zP = WP[1], zW = WP[2];                               \
                        const R *c = CROW;                                                        \
                        const R zS = isC ? lds_c(c, x, j, -1, jm, chan) : lds_t(c, x, j, -1, jm, chan, tinf, tb, ttop); \
                        const R zN = isC ? lds_c(c, x, j, 1, jm, chan) : lds_t(c, x, j, 1, jm, chan, tinf, tb, ttop);   \
                        const R avn_r = R(pr[4 * r + 1]);                                            \
                        const R kq = KPOLY ? kw[bq] : kap;                                        \
                        const R diff = (kq * aue_j) * (R(pr[4 * r + 6]) * (zE - zP) + R(pr[4 * r + 2]) * (zW - zP)) \
                                     + (kq * avn_r) * (rdy_n * (zN - zP) + rdy_j * (zS - zP));    \
                        const R uE = uw[bq - 1], uP = uw[bq], vP = vw[bq], vN = vnw[bq];          \
                        const R conv = R(0.5) * ((zE + zP) * uE - (zW + zP) * uP) * aue_j         \
                                     + R(0.5) * ((zN + zP) * vN - (zS + zP) * vP) * avn_r;        \
                        const R vols = R(pr[4 * r + 7]) * rdym_n;                            \
                        const R Rr = isC ? vols * (diff - conv * id[bq])                          \
                                         : vols * (diff - oml[bq] * conv) * id[bq];               \
                        const R rs = ((STEADY ? R(0) : dZ[bq]) + Rr * dt) * dtau;                 \
                        const R tau = zt[bq];                                                     \
                        if (K == 1) val = tau + rs;                                               \
                        else if (K == 2) val = R(0.75) * tau + R(0.25) * (zP + rs);               \
                        else val = R(1.0 / 3.0) * tau + R(2.0 / 3.0) * (zP + rs);                 \
                        if (r == 2) WOUT.set(0, (isC && chan) ? cslot : val);                     \
                    }                                                                             \
                } else if (r == im) {                                                             \
                    val = scol ? WOUT[0] : R(0);                                                  \
                } else if (r == 1) {                                                              \
                    val = (isC && chan) ? cslot : R(0);                                           \
                }                                                                                 \
                RINGW                                                                             \
            }                                                                                     \
            WOUT.push(val);                                                                       \
            if (OUTP != nullptr && interior && out_col && r >= ilo && r < ihi)                    \
                OUTP[(size_t)r * su + j] = val;                                                   \
        }
        R *const noout = nullptr;
        SCS_STAGE(1, zt, own_ring[s3(t - 1)], y1, { ring1[s2(t - 1)][PAD + x] = val; }, noout)
        SCS_STAGE(2, y1, ring1[s2(t - 2)], y2, { ring2[s2(t - 2)][PAD + x] = val; }, noout)
        SCS_STAGE(3, y2, ring2[s2(t - 3)], y3, {}, outp)
#undef SCS_STAGE
    }
}

template <typename R, bool STEADY, int NT, bool KPOLY>
__global__ __launch_bounds__(2 * NT, 4) void k_scal_split(const K2Args<R> a)
{
    __shared__ Smem2<R, NT> sm;
    const int G = a.nbi * a.nbj, per = (G + 7) / 8;
    const int b = blockIdx.x, L = (b % 8) * per + b / 8;
    if (L >= G) return;
    scal_tile_split<R, STEADY, NT, KPOLY>(a, L / a.nbj, L % a.nbj, sm);
}

template <typename R, bool STEADY, int NT>
__global__ __launch_bounds__(2 * NT, 4) void k_scal_split_batch(const K2Args<R> *cases, const BatchTile *tiles,
                                                                int ntiles, const int *ctl)
{
    __shared__ Smem2<R, NT> sm;
    for (int w = blockIdx.x; w < ntiles; w += gridDim.x) {
        const BatchTile tl = tiles[w];
        const int cs = ctl[tl.c];
        if (!(cs & 1)) continue;
        const K2Args<R> a = cases[2 * tl.c + ((cs >> 1) & 1)];
        scal_tile_split<R, STEADY, NT, false>(a, tl.bi, tl.bj, sm);
        __syncthreads();
    }
}

template <typename R, bool STEADY, int OCC, int NT, bool KPOLY>
__global__ __launch_bounds__(NT, OCC) void k_scal(const K2Args<R> a)
{
    __shared__ Smem2<R, NT> sm;
    const int G = a.nbi * a.nbj, per = (G + 7) / 8;
    const int b = blockIdx.x, L = (b % 8) * per + b / 8;
    if (L >= G) return;
    scal_tile<R, STEADY, NT, KPOLY>(a, L / a.nbj, L % a.nbj, sm);
}

template <typename R, bool STEADY, int OCC, int NT>
__global__ __launch_bounds__(NT, OCC) void k_scal_batch(const K2Args<R> *cases, const BatchTile *tiles,
                                                        int ntiles, const int *ctl)
{
    __shared__ Smem2<R, NT> sm;
    for (int w = blockIdx.x; w < ntiles; w += gridDim.x) {
        const BatchTile tl = tiles[w];
        const int cs = ctl[tl.c];
        if (!(cs & 1)) continue;
        const K2Args<R> a = cases[2 * tl.c + ((cs >> 1) & 1)];
        scal_tile<R, STEADY, NT, false>(a, tl.bi, tl.bj, sm);
        __syncthreads();
    }
}

// ---------------------------------------------------------------------------------------
// boundary cells of the five outputs
// ---------------------------------------------------------------------------------------
__host__ __device__ inline int materialize_count(int im, int jm)
{
    return 2 * (im + 1) + 4 * (jm - 2) + 4 * im + 2 * (jm - 3) + 3 * (2 * im + 2 * (jm - 2));
}

// Boundary cell number t of a fused step's outputs, from their interior values. The
// reference v corner (imax, jmax+1) reads the pre-BC column imax, which in the reference
// holds the previous step's output: that is the step's input iterate vprev.
template <typename F>
__device__ __forceinline__ void materialize_cell(const Geo &g, const cfd_params &P, F *u, F *v, double *p, F *T,
                                                 F *C, const F *vprev, int t)
{
    const int im = g.imax, jm = g.jmax;
    const bool chan = P.bc == CFD_BC_CHANNEL;
    const int nu_r = im + 1, nu_c = jm - 2;
    const int n_u = 2 * nu_r + 4 * nu_c;
    const int n_v = 4 * im + 2 * (jm - 3);
    const int n_s = 2 * im + 2 * (jm - 2);
    if (t < n_u) {
        int i, j;
        if (t < 2 * nu_r) { i = 1 + t % nu_r; j = t < nu_r ? 1 : jm; }
        else { const int q = t - 2 * nu_r, c = q / nu_c; j = 2 + q % nu_c; i = c == 0 ? 1 : (c == 1 ? 2 : (c == 2 ? im : im + 1)); }
        u[iu(g, i, j)] = (F)(chan ? bc_u_chan(g, P, u, i, j) : bc_u_ref(g, u, i, j));
        return;
    }
    int q = t - n_u;
    if (q < n_v) {
        int i, j;
        if (q < 4 * im) { const int r = q / im; i = 1 + q % im; j = r == 0 ? 1 : (r == 1 ? 2 : (r == 2 ? jm : jm + 1)); }
        else { const int c = q - 4 * im, nc = jm - 3; i = c < nc ? 1 : im; j = 3 + c % nc; }
        double val;
        if (!chan && i == im && j == jm + 1) {
            const double a = (double)vprev[iv(g, im, jm - 2)], b = (double)vprev[iv(g, im, jm - 1)];
            val = 2.0 * b - (2.0 * a - b);
        } else {
            val = chan ? bc_v_chan(g, v, i, j) : bc_v_ref(g, P, v, i, j);
        }
        v[iv(g, i, j)] = (F)val;
        return;
    }
    q -= n_v;
    if (q < 3 * n_s) {
        const int kind = q / n_s, r = q % n_s;
        int i, j;
        if (r < 2 * im) { i = 1 + r % im; j = r < im ? 1 : jm; }
        else { const int c = r - 2 * im, nc = jm - 2; i = c < nc ? 1 : im; j = 2 + c % nc; }
        if (kind == 0) {
            p[iu(g, i, j)] = chan ? bc_s_chan(g, P, 0, p, i, j) : bc_s_ref(g, P, 0, p, i, j);
        } else {
            F *a = kind == 1 ? T : C;
            a[iu(g, i, j)] = (F)(chan ? bc_s_chan(g, P, kind, a, i, j) : bc_s_ref(g, P, kind, a, i, j));
        }
    }
}

// boundary cells t0 .. t1-1 of the materialize_cell enumeration (u, v, p first, then T, C)
template <typename F>
__global__ void k_materialize(Geo g, cfd_params P, F *u, F *v, double *p, F *T, F *C, const F *vprev, int t0, int t1)
{
    const int t = t0 + blockIdx.x * blockDim.x + threadIdx.x;
    if (t < t1) materialize_cell(g, P, u, v, p, T, C, vprev, t);
}

// first T/C cell of the materialize_cell enumeration
__host__ __device__ inline int materialize_tc_start(int im, int jm)
{
    return 2 * (im + 1) + 4 * (jm - 2) + 4 * im + 2 * (jm - 3) + (2 * im + 2 * (jm - 2));
}

}  // namespace

// ---------------------------------------------------------------------------------------
// launchers
// ---------------------------------------------------------------------------------------
namespace {

// Row-block height: `want` > 0 (from the autotuner) or `per_cu` blocks per CU (the blocks
// that fit on one CU at once), at least min_rows.
int pick_rows(int rows, int nbj, int min_rows, int want, int per_cu, int *hb)
{
    if (want > 0) {
        *hb = want < rows ? want : rows;
        return cdiv(rows, *hb);
    }
    int nbi = 256 * per_cu / nbj;
    if (nbi < 1) nbi = 1;
    if (rows / nbi < min_rows) nbi = rows / min_rows > 0 ? rows / min_rows : 1;
    *hb = cdiv(rows, nbi);
    return cdiv(rows, *hb);
}

template <typename R>
void fill_k1(K1Args<R> &k1, int im, int jm, const double *pack, const cfd_params *P, const cfd_fused_args *a, int nt)
{
    k1.um = (const R *)a->um; k1.vm = (const R *)a->vm; k1.ut = (const R *)a->um_tau; k1.vt = (const R *)a->vm_tau;
    k1.p = a->p; k1.T = (const R *)a->T;
    k1.eps = (const R *)a->epsr; k1.lp = (const R *)a->lpr;
    k1.un = (R *)a->um_n; k1.vn = (R *)a->vm_n; k1.pn = a->pn; k1.res = (unsigned long long *)a->res_max;
    k1.l2 = a->res_l2;
    k1.res_next = (unsigned long long *)a->res_next;
    k1.pk = pack;
    k1.imax = im; k1.jmax = jm;
    k1.chan = P->bc == CFD_BC_CHANNEL; k1.upwind_all = P->scheme == CFD_SCHEME_UPWIND;
    k1.bouss = P->buoyancy == CFD_BUOY_BOUSSINESQ;
    k1.dt = P->dt; k1.dtau = P->dtau; k1.beta = P->beta; k1.inv_re = 1.0 / P->re;
    k1.inv_reda = 1.0 / (P->re * P->darcy); k1.cfsd = P->cf / sqrt(P->darcy); k1.g = P->g;
    k1.invfr2 = P->invfr2; k1.ri = P->ri; k1.t_ref = P->t_ref; k1.v_i = P->v_i; k1.u_in = P->u_in;
    for (int q = 0; q < 5; ++q) k1.mu[q] = P->mu_coef[q];
    k1.nbj = cdiv(jm - 2, w1(nt));
    // the two-wave kernel fits two workgroups on a CU
    k1.nbi = pick_rows(im - 2, k1.nbj, a->lone ? 4 : 24, a->hb1, a->occ1 == 2 ? 2 : 1, &k1.hb);
}

template <typename R>
void fill_k2(K2Args<R> &k2, int im, int jm, const double *y, const double *pack, const cfd_params *P,
             const cfd_fused_args *a, int nt)
{
    k2.un = (const R *)a->um_n; k2.vn = (const R *)a->vm_n; k2.T = (const R *)a->T; k2.Tt = (const R *)a->T_tau;
    k2.C = (const R *)a->C; k2.Ct = (const R *)a->C_tau;
    k2.eps = (const R *)a->epsr; k2.lp = (const R *)a->lpr; k2.y = y;
    k2.Tn = (R *)a->T_n; k2.Cn = (R *)a->C_n;
    k2.pk = pack;
    k2.imax = im; k2.jmax = jm; k2.chan = P->bc == CFD_BC_CHANNEL;
    k2.dt = P->dt; k2.dtau = P->dtau; k2.kt = 1.0 / P->pe; k2.kc = 1.0 / P->re / P->sc;
    k2.tinf = P->tinf; k2.t_bottom = P->t_bottom; k2.t_top = P->t_top; k2.v_i = P->v_i; k2.u_in = P->u_in;
    k2.slot_lo = P->slot_lo; k2.slot_hi = P->slot_hi; k2.slot_value = P->slot_value;
    for (int q = 0; q < 5; ++q) k2.kcoef[q] = P->k_coef[q];
    k2.nbj = cdiv(jm - 2, w2(nt));
    // beside the one-wave u/v/p kernel (the two-stream step) one T/C workgroup per CU fits;
    // after the two-wave one, which fills the CUs, two
    k2.nbi = pick_rows(im - 2, k2.nbj, a->lone ? 4 : 16, a->hb2, a->occ1 == 2 ? 2 : 1, &k2.hb);
}

// Column span of a single grid's workgroups: 64 when 64-column strips need no more lanes
// than 256-column ones (jmax 64, 128, 204, 256, 512 among others; not 1024 or 2048), else
// 256. The per-lane cost of the two is about the same and one-wave workgroups have no
// cross-wave barrier. Measured per pseudo-step, 256 -> 64 columns: 64x64 0.073 -> 0.064 ms,
// 204x204 0.076 -> 0.064 ms, 512x256 0.078 -> 0.071 ms, 1024x512 0.118 -> 0.113 ms; at
// 2048x2048 (2624 against 2304 lanes) 0.47 against 0.37 ms for K1.
int lone_span(int jm)
{
    const int lanes_b = cdiv(jm - 2, w1(NTB)) * NTB, lanes_1 = cdiv(jm - 2, w1(NT1)) * NT1;
    return lanes_b <= lanes_1 ? NTB : NT1;
}

// occ: 1 or 2 select k_uvp's register budget, 3 the split-chain kernel.
template <typename R, bool STEADY, bool MUPOLY, int NT, bool L2>
void launch_k1_l2(const K1Args<R> &k1, int occ, int grid1, hipStream_t s)
{
    const dim3 g1(grid1);
    if (occ == 3) hipLaunchKernelGGL((k_uvp_split<R, STEADY, MUPOLY, NT, L2>), g1, dim3(2 * NT), 0, s, k1);
    else if (occ == 2) hipLaunchKernelGGL((k_uvp<R, STEADY, MUPOLY, 2, NT, L2>), g1, dim3(NT), 0, s, k1);
    else hipLaunchKernelGGL((k_uvp<R, STEADY, MUPOLY, 1, NT, L2>), g1, dim3(NT), 0, s, k1);
}

template <typename R, bool STEADY, bool MUPOLY, int NT>
void launch_k1_v(const K1Args<R> &k1, int occ, int grid1, hipStream_t s)
{
    if (k1.l2) launch_k1_l2<R, STEADY, MUPOLY, NT, true>(k1, occ, grid1, s);
    else launch_k1_l2<R, STEADY, MUPOLY, NT, false>(k1, occ, grid1, s);
}

template <typename R, int NT>
void launch_k1_nt(const K1Args<R> &k1, const cfd_params *P, int occ, hipStream_t s)
{
    const int G1 = k1.nbi * k1.nbj, grid1 = ((G1 + 7) / 8) * 8;
    const bool mp = P->mu_model == CFD_MU_POLY4;
    if (P->steady) {
        if (mp) launch_k1_v<R, true, true, NT>(k1, occ, grid1, s);
        else launch_k1_v<R, true, false, NT>(k1, occ, grid1, s);
    } else {
        if (mp) launch_k1_v<R, false, true, NT>(k1, occ, grid1, s);
        else launch_k1_v<R, false, false, NT>(k1, occ, grid1, s);
    }
}

template <typename R>
void launch_k1(const K1Args<R> &k1, const cfd_params *P, int occ, hipStream_t s, int nt)
{
    if (nt == NTB) launch_k1_nt<R, NTB>(k1, P, occ, s);
    else launch_k1_nt<R, NT1>(k1, P, occ, s);
}

template <typename R, int NT, bool KP>
void launch_k2_kp(const K2Args<R> &k2, const cfd_params *P, int occ, hipStream_t s)
{
    const int G2 = k2.nbi * k2.nbj, grid2 = ((G2 + 7) / 8) * 8;
    const dim3 g2(grid2), b2(NT);
    if (occ == 2) {           // split chains: T and C in separate waves
        if (P->steady) hipLaunchKernelGGL((k_scal_split<R, true, NT, KP>), g2, dim3(2 * NT), 0, s, k2);
        else hipLaunchKernelGGL((k_scal_split<R, false, NT, KP>), g2, dim3(2 * NT), 0, s, k2);
    } else if (occ == 4) {
        if (P->steady) hipLaunchKernelGGL((k_scal<R, true, 4, NT, KP>), g2, b2, 0, s, k2);
        else hipLaunchKernelGGL((k_scal<R, false, 4, NT, KP>), g2, b2, 0, s, k2);
    } else {
        if (P->steady) hipLaunchKernelGGL((k_scal<R, true, 1, NT, KP>), g2, b2, 0, s, k2);
        else hipLaunchKernelGGL((k_scal<R, false, 1, NT, KP>), g2, b2, 0, s, k2);
    }
}

template <typename R, int NT>
void launch_k2_nt(const K2Args<R> &k2, const cfd_params *P, int occ, hipStream_t s)
{
    if (P->k_model == CFD_K_POLY4) launch_k2_kp<R, NT, true>(k2, P, occ, s);
    else launch_k2_kp<R, NT, false>(k2, P, occ, s);
}

template <typename R>
void launch_k2(const K2Args<R> &k2, const cfd_params *P, int occ, hipStream_t s, int nt)
{
    if (nt == NTB) launch_k2_nt<R, NTB>(k2, P, occ, s);
    else launch_k2_nt<R, NT2>(k2, P, occ, s);
}

template <typename R>
int fused_step(const cfd_devgeo *g, const cfd_devfactors *f, const cfd_params *P, const cfd_fused_args *a,
               hipStream_t s)
{
    const bool timed = a->ev[0] != nullptr;
    // Two-stream form (a->s2): K2 and the T/C boundary cells run on s2 after K1, so that the
    // next step's K1 (which reads only u, v, p and the physical level) overlaps them. The
    // caller passes the event of the K2 two steps back, whose inputs this K1 overwrites.
    const bool two = a->s2 != nullptr && !timed;
    const int im = g->imax, jm = g->jmax;
    const int n = materialize_count(im, jm), ntc = materialize_tc_start(im, jm);
    K1Args<R> k1;
    const int nt = lone_span(jm);
    fill_k1<R>(k1, im, jm, f->pack, P, a, nt);
    if (two && a->ev_k2_wait) (void)hipStreamWaitEvent(s, a->ev_k2_wait, 0);
    if (timed) (void)hipEventRecord(a->ev[0], s);
    launch_k1<R>(k1, P, a->occ1, s, nt);
    int r = launched();
    if (r != CFD_OK) return r;
    if (timed) (void)hipEventRecord(a->ev[1], s);
    const hipStream_t s2 = two ? a->s2 : s;
    if (two) {
        (void)hipEventRecord(a->ev_k1, s);
        hipLaunchKernelGGL((k_materialize<R>), dim3(cdiv(ntc, 256)), dim3(256), 0, s, mk(g), *P, (R *)a->um_n,
                           (R *)a->vm_n, a->pn, (R *)a->T_n, (R *)a->C_n, (const R *)a->vm_tau, 0, ntc);
        r = launched();
        if (r != CFD_OK) return r;
        (void)hipStreamWaitEvent(s2, a->ev_k1, 0);
    }
    K2Args<R> k2;
    fill_k2<R>(k2, im, jm, g->y, f->pack, P, a, nt);
    launch_k2<R>(k2, P, a->occ2, s2, nt);
    r = launched();
    if (r != CFD_OK) return r;
    if (timed) (void)hipEventRecord(a->ev[2], s);
    const int t0 = two ? ntc : 0;
    hipLaunchKernelGGL((k_materialize<R>), dim3(cdiv(n - t0, 256)), dim3(256), 0, s2, mk(g), *P, (R *)a->um_n,
                       (R *)a->vm_n, a->pn, (R *)a->T_n, (R *)a->C_n, (const R *)a->vm_tau, t0, n);
    r = launched();
    if (r != CFD_OK) return r;
    if (two) (void)hipEventRecord(a->ev_k2_rec, s2);
    if (timed) (void)hipEventRecord(a->ev[3], s);
    return r;
}

}  // namespace

extern "C" {

int cfdk_fused_available(void) { return 1; }

int cfdk_fused_step(const cfd_devgeo *g, const cfd_devfactors *f, const cfd_params *P,
                    const cfd_fused_args *a, int kernel, hipStream_t s)
{
    (void)kernel;
    if (P->b_art != 0.0 || P->compat != CFD_COMPAT_FORTRAN) return CFD_ERR_UNSUPPORTED;
    if (a->prec == CFD_PREC_FP32) return fused_step<float>(g, f, P, a, s);
    return fused_step<double>(g, f, P, a, s);
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
        pr[4 * t] = (t >= 2 && t <= g.imax) ? g.aun[t] : 0.0;
        pr[4 * t + 1] = (t >= 1 && t <= g.imax) ? g.avn[t] : 0.0;
        pr[4 * t + 2] = a;
        pr[4 * t + 3] = b;
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

__global__ void k_d2f(float *dst, const double *src, long n)
{
    for (long t = (long)blockIdx.x * blockDim.x + threadIdx.x; t < n; t += (long)gridDim.x * blockDim.x)
        dst[t] = (float)src[t];
}

__global__ void k_f2d(double *dst, const float *src, long n)
{
    for (long t = (long)blockIdx.x * blockDim.x + threadIdx.x; t < n; t += (long)gridDim.x * blockDim.x)
        dst[t] = (double)src[t];
}

inline int conv_blocks(long n)
{
    const long b = (n + 255) / 256;
    return (int)(b < 4096 ? (b > 0 ? b : 1) : 4096);
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
    const int n = materialize_count(g->imax, g->jmax);
    hipLaunchKernelGGL((k_materialize<double>), dim3(cdiv(n, 256)), dim3(256), 0, s, mk(g), *P, um_n, vm_n, pn,
                       T_n, C_n, vm_prev, 0, n);
    return launched();
}

int cfdk_d2f(float *dst, const double *src, long n, hipStream_t s)
{
    hipLaunchKernelGGL(k_d2f, dim3(conv_blocks(n)), dim3(256), 0, s, dst, src, n);
    return launched();
}

int cfdk_f2d(double *dst, const float *src, long n, hipStream_t s)
{
    hipLaunchKernelGGL(k_f2d, dim3(conv_blocks(n)), dim3(256), 0, s, dst, src, n);
    return launched();
}

}  // extern "C"

// ---------------------------------------------------------------------------------------
// batched persistent path
// ---------------------------------------------------------------------------------------
namespace {

template <typename F>
struct MatArgs {
    Geo g;
    cfd_params P;
    F *u, *v, *T, *C;
    double *p;
    const F *vprev;
};

// Boundary cells of every active case; blockIdx.y is the case.
template <typename F>
__global__ void k_materialize_batch(const MatArgs<F> *mats, const int *ctl)
{
    const int c = blockIdx.y;
    const int cs = ctl[c];
    if (!(cs & 1)) return;
    const MatArgs<F> m = mats[2 * c + ((cs >> 1) & 1)];
    const int t = blockIdx.x * blockDim.x + threadIdx.x;
    if (t < materialize_count(m.g.imax, m.g.jmax)) materialize_cell(m.g, m.P, m.u, m.v, m.p, m.T, m.C, m.vprev, t);
}

__global__ void k_batch_zero(unsigned long long *res, const int *ctl, int n)
{
    const int c = blockIdx.x * blockDim.x + threadIdx.x;
    if (c < n && (ctl[c] & 1)) { res[3 * c] = 0; res[3 * c + 1] = 0; res[3 * c + 2] = 0; }
}

// Exit test of main.f90:139-150 per case: err = max(u, v, p); stop without the tau update
// when itc != 1 and err < tol; otherwise swap the tau roles (parity bit) and go on.
// The residual slots of a case that stays active are cleared here, after they are read, for
// its next pseudo-step (a run clears the slots of its active cases once before its first step).
__global__ void k_batch_check(int *ctl, unsigned long long *res, int *status, long long *itc, double *best,
                              double *last, int n, double tol, double diverge, long long itc_max)
{
    const int c = blockIdx.x * blockDim.x + threadIdx.x;
    if (c >= n) return;
    int cs = ctl[c];
    if (!(cs & 1)) return;
    const long long k = itc[c] + 1;
    itc[c] = k;
    double r[3];
    for (int q = 0; q < 3; ++q) {
        r[q] = __longlong_as_double((long long)res[3 * c + q]);
        last[3 * c + q] = r[q];
    }
    const double err = fmax(fmax(r[0], r[1]), r[2]);
    if (!(r[0] <= 1.0e308 && r[1] <= 1.0e308 && r[2] <= 1.0e308)) {       // NaN or Inf
        status[c] = 2;
        cs &= ~1;
    } else if (k != 1 && err < tol) {
        status[c] = 1;
        cs &= ~1;
    } else {
        const double b = fmin(best[c], err);
        best[c] = b;
        cs ^= 2;
        if (diverge > 0.0 && err > diverge * b) { status[c] = 3; cs &= ~1; }
        else if (k >= itc_max) { status[c] = 4; cs &= ~1; }
    }
    ctl[c] = cs;
    if (cs & 1) { res[3 * c] = 0; res[3 * c + 1] = 0; res[3 * c + 2] = 0; }
}

__global__ void k_batch_count(const int *ctl, int n, int *out)
{
    __shared__ int part[256];
    int s = 0;
    for (int c = threadIdx.x; c < n; c += blockDim.x) s += ctl[c] & 1;
    part[threadIdx.x] = s;
    __syncthreads();
    for (int o = blockDim.x / 2; o > 0; o >>= 1) {
        if ((int)threadIdx.x < o) part[threadIdx.x] += part[threadIdx.x + o];
        __syncthreads();
    }
    if (threadIdx.x == 0) *out = part[0];
}

template <typename R, bool STEADY, bool MUPOLY, int NT>
void batch_launch_k1_nt(const cfd_batch_launch *L, hipStream_t s)
{
    const dim3 g1(L->grid1), b1(NT);
    const K1Args<R> *a = (const K1Args<R> *)L->k1;
    const BatchTile *t = (const BatchTile *)L->tiles1;
    if (L->occ1 == 3)
        hipLaunchKernelGGL((k_uvp_split_batch<R, STEADY, MUPOLY, NT>), g1, dim3(2 * NT), 0, s, a, t, L->ntiles1, L->ctl);
    else if (L->occ1 == 2)
        hipLaunchKernelGGL((k_uvp_batch<R, STEADY, MUPOLY, 2, NT>), g1, b1, 0, s, a, t, L->ntiles1, L->ctl);
    else hipLaunchKernelGGL((k_uvp_batch<R, STEADY, MUPOLY, 1, NT>), g1, b1, 0, s, a, t, L->ntiles1, L->ctl);
}

template <typename R, bool STEADY, bool MUPOLY>
void batch_launch_k1(const cfd_batch_launch *L, hipStream_t s)
{
    if (L->span == NT1) batch_launch_k1_nt<R, STEADY, MUPOLY, NT1>(L, s);
    else batch_launch_k1_nt<R, STEADY, MUPOLY, NTB>(L, s);
}

template <typename R, bool STEADY, int NT>
void batch_launch_k2_nt(const cfd_batch_launch *L, hipStream_t s)
{
    const dim3 g2(L->grid2), b2(NT);
    const K2Args<R> *a = (const K2Args<R> *)L->k2;
    const BatchTile *t = (const BatchTile *)L->tiles2;
    if (L->occ2 == 2)
        hipLaunchKernelGGL((k_scal_split_batch<R, STEADY, NT>), g2, dim3(2 * NT), 0, s, a, t, L->ntiles2, L->ctl);
    else if (L->occ2 == 4) hipLaunchKernelGGL((k_scal_batch<R, STEADY, 4, NT>), g2, b2, 0, s, a, t, L->ntiles2, L->ctl);
    else hipLaunchKernelGGL((k_scal_batch<R, STEADY, 1, NT>), g2, b2, 0, s, a, t, L->ntiles2, L->ctl);
}

template <typename R, bool STEADY>
void batch_launch_k2(const cfd_batch_launch *L, hipStream_t s)
{
    if (L->span == NT1) batch_launch_k2_nt<R, STEADY, NT1>(L, s);
    else batch_launch_k2_nt<R, STEADY, NTB>(L, s);
}

template <typename R>
int batch_step(const cfd_batch_launch *L, hipStream_t s)
{
    const int nb = cdiv(L->ncases, 256);
    if (L->steady) {
        if (L->mupoly) batch_launch_k1<R, true, true>(L, s);
        else batch_launch_k1<R, true, false>(L, s);
        batch_launch_k2<R, true>(L, s);
    } else {
        if (L->mupoly) batch_launch_k1<R, false, true>(L, s);
        else batch_launch_k1<R, false, false>(L, s);
        batch_launch_k2<R, false>(L, s);
    }
    hipLaunchKernelGGL((k_materialize_batch<R>), dim3(cdiv(L->max_mat, 256), L->ncases), dim3(256), 0, s,
                       (const MatArgs<R> *)L->mat, L->ctl);
    hipLaunchKernelGGL(k_batch_check, dim3(nb), dim3(256), 0, s, L->ctl, (unsigned long long *)L->res,
                       L->status, (long long *)L->itc, L->best, L->last, L->ncases, L->tol, L->diverge,
                       (long long)L->itc_max);
    return launched();
}

// Boundary cells of the tau-level iterates of every case. The argument set of the other
// parity writes its n buffers, which are this parity's tau buffers. Channel BC set only:
// the reference set's corner (imax, jmax+1) reads the pre-BC column of the previous step.
template <typename F>
__global__ void k_apply_bc_batch(const MatArgs<F> *mats, const int *ctl)
{
    const int c = blockIdx.y;
    const int cs = ctl[c];
    const MatArgs<F> m = mats[2 * c + (1 - ((cs >> 1) & 1))];
    const int t = blockIdx.x * blockDim.x + threadIdx.x;
    if (t < materialize_count(m.g.imax, m.g.jmax)) materialize_cell(m.g, m.P, m.u, m.v, m.p, m.T, m.C, m.v, t);
}

template <typename R>
void build_args(int32_t ncases, const cfd_batch_case *cases, const int32_t *hb1, const int32_t *hb2, void *k1v,
                void *k2v, void *matv, int span)
{
    K1Args<R> *k1 = (K1Args<R> *)k1v;
    K2Args<R> *k2 = (K2Args<R> *)k2v;
    MatArgs<R> *mat = (MatArgs<R> *)matv;
    for (int c = 0; c < ncases; ++c) {
        const cfd_batch_case &b = cases[c];
        const int im = b.P.imax, jm = b.P.jmax;
        for (int q = 0; q < 2; ++q) {
            cfd_fused_args a = b.a[q];
            a.hb1 = hb1 ? hb1[c] : 0;
            a.hb2 = hb2 ? hb2[c] : 0;
            fill_k1<R>(k1[2 * c + q], im, jm, b.pack, &b.P, &a, span);
            fill_k2<R>(k2[2 * c + q], im, jm, b.y, b.pack, &b.P, &a, span);
            MatArgs<R> &m = mat[2 * c + q];
            m.g = Geo{};
            m.g.y = b.y; m.g.imax = im; m.g.jmax = jm;
            m.P = b.P;
            m.u = (R *)a.um_n; m.v = (R *)a.vm_n; m.p = a.pn; m.T = (R *)a.T_n; m.C = (R *)a.C_n;
            m.vprev = (const R *)a.vm_tau;
        }
    }
}

}  // namespace

extern "C" {

int cfdk_batch_arg_sizes(int32_t prec, size_t *k1, size_t *k2, size_t *mat)
{
    if (prec == CFD_PREC_FP32) {
        *k1 = sizeof(K1Args<float>); *k2 = sizeof(K2Args<float>); *mat = sizeof(MatArgs<float>);
    } else {
        *k1 = sizeof(K1Args<double>); *k2 = sizeof(K2Args<double>); *mat = sizeof(MatArgs<double>);
    }
    return CFD_OK;
}

int cfdk_batch_build_args(int32_t prec, int32_t ncases, const cfd_batch_case *cases, const int32_t *hb1,
                          const int32_t *hb2, void *k1, void *k2, void *mat, int32_t span)
{
    if (span != NTB && span != NT1) return CFD_ERR_ARG;
    if (prec == CFD_PREC_FP32) build_args<float>(ncases, cases, hb1, hb2, k1, k2, mat, span);
    else build_args<double>(ncases, cases, hb1, hb2, k1, k2, mat, span);
    return CFD_OK;
}

// Column span of the batch's workgroups: 64 unless 256-column strips need fewer lanes over
// all cases (wide grids: a 2048-column case takes 2304 lanes in 256-column strips, 2624 in
// 64-column ones), as for a single grid.
int32_t cfdk_batch_span(int32_t ncases, const int32_t *jmax)
{
    long lanes_b = 0, lanes_1 = 0;
    for (int32_t c = 0; c < ncases; ++c) {
        lanes_b += (long)cdiv(jmax[c] - 2, w1(NTB)) * NTB;
        lanes_1 += (long)cdiv(jmax[c] - 2, w1(NT1)) * NT1;
    }
    return lanes_b <= lanes_1 ? NTB : NT1;
}

int cfdk_batch_case_tiles(int kernel, int32_t span, int32_t imax, int32_t jmax, int32_t hb, int32_t c, int32_t *out)
{
    const int W = kernel == 1 ? w1(span) : w2(span);
    const int nbj = cdiv(jmax - 2, W);
    int h = 0;
    const int nbi = pick_rows(imax - 2, nbj, kernel == 1 ? 24 : 16, hb, 1, &h);
    if (out) {
        int k = 0;
        for (int bi = 0; bi < nbi; ++bi)
            for (int bj = 0; bj < nbj; ++bj, ++k) {
                out[4 * k] = c; out[4 * k + 1] = bi; out[4 * k + 2] = bj; out[4 * k + 3] = 0;
            }
    }
    return nbi * nbj;
}

}  // extern "C"

namespace {

template <int NT>
hipError_t batch_occupancy(int32_t prec, int32_t steady, int32_t mupoly, int32_t occ1, int32_t occ2, int *b1, int *b2)
{
    hipError_t e1, e2;
#define OCC1(R, S, M)                                                                                      \
    (occ1 == 3 ? hipOccupancyMaxActiveBlocksPerMultiprocessor(b1, k_uvp_split_batch<R, S, M, NT>, 2 * NT, 0) \
     : occ1 == 2 ? hipOccupancyMaxActiveBlocksPerMultiprocessor(b1, k_uvp_batch<R, S, M, 2, NT>, NT, 0)    \
                 : hipOccupancyMaxActiveBlocksPerMultiprocessor(b1, k_uvp_batch<R, S, M, 1, NT>, NT, 0))
#define OCC2(R, S)                                                                                         \
    (occ2 == 2 ? hipOccupancyMaxActiveBlocksPerMultiprocessor(b2, k_scal_split_batch<R, S, NT>, 2 * NT, 0) \
     : occ2 == 4 ? hipOccupancyMaxActiveBlocksPerMultiprocessor(b2, k_scal_batch<R, S, 4, NT>, NT, 0)      \
                 : hipOccupancyMaxActiveBlocksPerMultiprocessor(b2, k_scal_batch<R, S, 1, NT>, NT, 0))
    if (prec == CFD_PREC_FP32) {
        if (steady) { e1 = mupoly ? OCC1(float, true, true) : OCC1(float, true, false); e2 = OCC2(float, true); }
        else { e1 = mupoly ? OCC1(float, false, true) : OCC1(float, false, false); e2 = OCC2(float, false); }
    } else {
        if (steady) { e1 = mupoly ? OCC1(double, true, true) : OCC1(double, true, false); e2 = OCC2(double, true); }
        else { e1 = mupoly ? OCC1(double, false, true) : OCC1(double, false, false); e2 = OCC2(double, false); }
    }
#undef OCC1
#undef OCC2
    return e1 != hipSuccess ? e1 : e2;
}

}  // namespace

extern "C" {

int cfdk_batch_occupancy(int32_t prec, int32_t steady, int32_t mupoly, int32_t occ1, int32_t occ2, int32_t span,
                         int *k1_blocks_per_cu, int *k2_blocks_per_cu)
{
    int b1 = 0, b2 = 0;
    const hipError_t e = span == NT1 ? batch_occupancy<NT1>(prec, steady, mupoly, occ1, occ2, &b1, &b2)
                                     : batch_occupancy<NTB>(prec, steady, mupoly, occ1, occ2, &b1, &b2);
    if (e != hipSuccess) return CFD_ERR_HIP;
    *k1_blocks_per_cu = b1 > 0 ? b1 : 1;
    *k2_blocks_per_cu = b2 > 0 ? b2 : 1;
    return CFD_OK;
}

int cfdk_batch_step(const cfd_batch_launch *L, hipStream_t s)
{
    if (L->prec == CFD_PREC_FP32) return batch_step<float>(L, s);
    return batch_step<double>(L, s);
}

int cfdk_batch_clear(const cfd_batch_launch *L, hipStream_t s)
{
    hipLaunchKernelGGL(k_batch_zero, dim3(cdiv(L->ncases, 256)), dim3(256), 0, s, (unsigned long long *)L->res,
                       L->ctl, L->ncases);
    return launched();
}

int cfdk_batch_count(const cfd_batch_launch *L, hipStream_t s)
{
    hipLaunchKernelGGL(k_batch_count, dim3(1), dim3(256), 0, s, L->ctl, L->ncases, L->nactive);
    return launched();
}

int cfdk_batch_apply_bc(const cfd_batch_launch *L, hipStream_t s)
{
    if (L->prec == CFD_PREC_FP32)
        hipLaunchKernelGGL((k_apply_bc_batch<float>), dim3(cdiv(L->max_mat, 256), L->ncases), dim3(256), 0, s,
                           (const MatArgs<float> *)L->mat, L->ctl);
    else
        hipLaunchKernelGGL((k_apply_bc_batch<double>), dim3(cdiv(L->max_mat, 256), L->ncases), dim3(256), 0, s,
                           (const MatArgs<double> *)L->mat, L->ctl);
    return launched();
}

}  // extern "C"
