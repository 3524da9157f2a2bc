// Device body of the fused barotropic step, shared by the single-basin
// kernel (bt_fused.hip) and the batched-ensemble kernel (bt_ensemble.hip).
// The algorithm and its modes are described at the top of bt_fused.hip.
#pragma once

#include "bt_internal.h"

namespace ibt {

constexpr int WAVE = 64;
constexpr int SW = 60;   // output columns per wave, one substep per launch
constexpr int SW2 = 58;  // output columns per wave, two substeps per launch
constexpr int NXCD = 8;
// cells (summed over an ensemble's members) from which the fused step runs
// two substeps per launch by default
constexpr long long TWO_SUBSTEP_CELLS = 1LL << 20;

struct FusedParams {
    int nx, ny, strips, chunk, nchunks, total, per_xcd;
    const uint8_t *mb;
    const void *H, *rd;     // stencil type
    const void *u, *v, *s;  // level n, stencil type
    const double *up, *vp, *sp;
    void *uo, *vo, *so;
    double *upo, *vpo, *spo;
    const double *row_f, *row_sx, *row_sy, *row_wf;  // per row, index n-1
    double g, rho0, ts, tau, cb, eb, nu, ffs, d;
    // sub-cycling: substep `sub` of `nsub` (tau = tau_inner = outer tau / nsub)
    int sub, nsub;
    double dn;             // dfloat(nstep)
    void *frx, *fry;       // forcing frozen over the substeps of one call (stencil type), written
    const void *frx_in, *fry_in;  // the same, read (a separate buffer when a launch forms new forcing)
    double *hup_c, *hvp_c; // filtered U/V depths carried between substeps (FP64_EXACT only)
};

// Lane shifts by DPP (no LDS traffic).  wave_shl:1 gives lane L the value
// of lane L+1 (column m+1), wave_shr:1 the value of lane L-1 (column m-1);
// the lanes at the wave's ends receive 0 and only ever feed halo lanes.
__device__ __forceinline__ int dpp_from_right(int x)
{
    return __builtin_amdgcn_update_dpp(0, x, 0x130, 0xF, 0xF, true);
}
__device__ __forceinline__ int dpp_from_left(int x)
{
    return __builtin_amdgcn_update_dpp(0, x, 0x138, 0xF, 0xF, true);
}
__device__ __forceinline__ double xr(double x)  // value of column m+1
{
    return __hiloint2double(dpp_from_right(__double2hiint(x)), dpp_from_right(__double2loint(x)));
}
__device__ __forceinline__ double xl(double x)  // value of column m-1
{
    return __hiloint2double(dpp_from_left(__double2hiint(x)), dpp_from_left(__double2loint(x)));
}
__device__ __forceinline__ float xr(float x) { return __int_as_float(dpp_from_right(__float_as_int(x))); }
__device__ __forceinline__ float xl(float x) { return __int_as_float(dpp_from_left(__float_as_int(x))); }

// Stores of the step's outputs.  Measured on MI355X at 2048 x 1536 (us per
// launch, FP64 / FP32 / 64-member FP32 ensemble): all cached 100.0 / 71.5 /
// 577; non-temporal level-n and filtered-level stores in FP64, filtered-level
// stores in FP32 (FP64 arrays), frozen forcing cached: 94.4 / 66.5 / 572.
// Non-temporal FP32 level-n stores (69.0 / 612 FP32 / ensemble) and
// non-temporal loads of the filtered level (74.3 / 603) are slower.
template <bool NT, class T>
__device__ __forceinline__ void st_out(void *base, size_t i, T v)
{
    T *q = static_cast<T *>(base) + i;
    if constexpr (NT)
        __builtin_nontemporal_store(v, q);
    else
        *q = v;
}

template <class R>
__device__ __forceinline__ R bitf(int mb, int bit)
{
    return (mb & bit) ? R(1) : R(0);
}

template <typename R, bool EXACT>
struct Ops {
    const FusedParams &p;
    R d, inv_d;
    __device__ Ops(const FusedParams &pp) : p(pp), d(R(pp.d)), inv_d(R(1.0 / pp.d)) {}

    // depth on a U or V point from its two T points (Function/depth.f90:29-49)
    __device__ R avg2(R qa, R qb, R la, R lb) const
    {
        R slu = la + lb;
        if constexpr (EXACT) {
            return (((qa * d) * d) * la + ((qb * d) * d) * lb) / slu / d / d;
        } else {
#pragma clang fp contract(fast)
            return (la * qa + lb * qb) * (slu > R(1.5) ? R(0.5) : R(1));
        }
    }
    // depth on an H point from its four T points (:51-66)
    __device__ R avg4(R q1, R q2, R q3, R q4, R l1, R l2, R l3, R l4) const
    {
        R slu = l1 + l2 + l3 + l4;
        if constexpr (EXACT) {
            return (((q1 * d) * d) * l1 + ((q2 * d) * d) * l2 + ((q3 * d) * d) * l3 +
                    ((q4 * d) * d) * l4) /
                   slu / d / d;
        } else {
#pragma clang fp contract(fast)
            R rs = slu > R(3.5) ? R(0.25) : slu > R(2.5) ? R(1.0 / 3.0) : slu > R(1.5) ? R(0.5) : R(1);
            return (l1 * q1 + l2 * q2 + l3 * q3 + l4 * q4) * rs;
        }
    }
};

// Loads are branch-free: the address is clamped into the array and the
// value is zeroed afterwards when the cell lies outside it.  Straight-line
// loads let the compiler wait for exactly the older (prefetched) loads.
struct RowAddr {
    int nx, ny, mc;  // mc = column clamped to 1..nx
    __device__ __forceinline__ size_t at(int row) const
    {
        const int rc = min(max(row, 1), ny);
        return (size_t)(rc - 1) * nx + (mc - 1);
    }
};

// Rows per wave when the caller does not choose one, for a grid of `rows`
// rows cut into `strips` strips on a chip that holds `slots` waves of the
// kernel at once.  An item marches its chunk plus `warm` warm-up rows, and
// the launch takes as many rounds as it has items per slot, so the time
// goes as rounds x (chunk + warm): pick the number of rounds k (1..8) that
// minimises it, with the chunk that fits k rounds exactly.  Measured on
// MI355X at 2048 x 1536 (1024 SIMDs): one item more than a round costs up to
// 40 % (two substeps per launch in FP32: 28 rows, 1980 items in 2048 slots,
// 54 us per substep; 27 rows, 2088 items, 77 us).  Items keep at least `lo`
// rows (512 x 384, one substep per launch: 2 rows 8.7 us, 1 row 10.5 us).
inline int fit_chunk(int rows, int strips, long long slots, int warm, int lo)
{
    int best = rows;
    long long best_cost = -1;
    for (int k = 1; k <= 8; ++k) {
        const long long per = (long long)k * slots / strips;  // chunks per strip that fit k rounds
        if (per < 1) continue;
        int chunk = (int)((rows + per - 1) / per);
        if (chunk < lo) chunk = lo;
        const long long items = (long long)strips * ((rows + chunk - 1) / chunk);
        const long long rounds = (items + slots - 1) / slots;
        const long long cost = rounds * (chunk + warm);
        if (best_cost < 0 || cost < best_cost) {
            best_cost = cost;
            best = chunk;
        }
    }
    return best;
}

// The older rule, kept for the ensemble's persistent launch: the measured
// best for large grids (FP64 32, FP32 20 rows), fewer when the work items
// (strips x row chunks) would not give the chip about 4 per CU.
inline int auto_chunk(bool fp32, long long strip_rows)
{
    const long long big = fp32 ? 20 : 32;
    const long long c = strip_rows / 1024;
    return (int)(c < 4 ? 4 : c > big ? big : c);
}

// What one substep consumes at iteration n of its march (already zeroed
// outside the array): surface height, depth at rest and mask bits of row
// n+2; level-n velocity, filtered level and Rayleigh coefficient of row
// n+1; the frozen forcing and carried U/V depths of row n; the row tables.
template <typename R>
struct StageIn {
    R s2, H2;
    int mb2;
    R u1, v1, rd1;
    double up1, vp1, sp1;
    R fx0, fy0;
    double hupc0, hvpc0;
    R f0, sx0, sy0, wf1;
};

// What one substep produces for output row n: the new level-n fields and
// filtered level after the Robert-Asselin shift, the forcing it used
// (formed at the call's first substep, frozen afterwards) and, in
// FP64_EXACT, the hh_shift-filtered U/V depths for the next substep.
template <typename R>
struct StageOut {
    R s, u, v;
    double sp, up, vp;
    R rhs_u, rhs_v;
    double hu, hv;
};

// One leap-frog substep marching down the rows of one strip: the rolling
// register window (suffix m1 = row n-1, 0 = row n, p1 = n+1, p2 = n+2) of
// input rows and of the T- and H-point intermediates.  row(n) takes the
// inputs of iteration n and, when asked, computes the outputs of row n.
// FORM: the call's first substep, which forms the wind + drag forcing;
// the later substeps take it frozen (StageIn::fx0, fy0).
template <typename R, bool EXACT, bool FORM>
struct Stage {
    static constexpr bool frozen = !FORM;
    const FusedParams &p;
    const Ops<R, EXACT> op;
    const bool inx, inx1;

    R u_m1 = 0, u_0 = 0, u_p1 = 0, v_m1 = 0, v_0 = 0, v_p1 = 0;
    R s_0 = 0, s_p1 = 0, s_p2 = 0, H_p1 = 0, H_p2 = 0;
    double up_0 = 0, up_p1 = 0, vp_0 = 0, vp_p1 = 0, sp_0 = 0, sp_p1 = 0;
    R rd_0 = 0, rd_p1 = 0;
    int mb_m1 = 0, mb_0 = 0, mb_p1 = 0, mb_p2 = 0;
    R hq_0 = 0, hq_p1 = 0, hq_p2 = 0, hqp_0 = 0, hqp_p1 = 0, hqn_0 = 0, hqn_p1 = 0;
    R hu_0 = 0, hu_p1 = 0, hup_0 = 0, hup_p1 = 0;
    R hv_m1 = 0, hv_0 = 0, hv_p1 = 0, hvp_0 = 0, hvn_0 = 0;
    R hh_m1 = 0, hh_0 = 0;
    double sn_0 = 0, sn_p1 = 0;
    R hun_0 = 0, hun_p1 = 0, stt_0 = 0, stt_p1 = 0, bf_0 = 0, bf_p1 = 0;
    R sts_m1 = 0, sts_0 = 0, vo_m1 = 0, vo_0 = 0;
    R f_m1 = 0, f_0 = 0;

    __device__ Stage(const FusedParams &pp, int m)
        : p(pp), op(pp), inx(m >= 3 && m <= pp.nx - 2), inx1(m >= 2 && m <= pp.nx - 1)
    {
    }

    __device__ __forceinline__ void row(int n, const StageIn<R> &q, bool produce, StageOut<R> &o)
    {
        const R d = op.d;
        const R tau = R(p.tau);
        const R G = R(p.g);
        const R nu = R(p.nu);
        const int k1 = n + 1;
        s_p2 = q.s2;
        H_p2 = q.H2;
        mb_p2 = q.mb2;
        u_p1 = q.u1;
        v_p1 = q.v1;
        up_p1 = q.up1;
        vp_p1 = q.vp1;
        sp_p1 = q.sp1;
        rd_p1 = q.rd1;
        f_m1 = f_0;
        f_0 = q.f0;
        const R sx_0 = q.sx0, sy_0 = q.sy0, wf_p1 = q.wf1;

        const R lu_0 = bitf<R>(mb_0, MB_LU), lu_p1 = bitf<R>(mb_p1, MB_LU), lu_p2 = bitf<R>(mb_p2, MB_LU);
        const R lu_0r = xr(lu_0), lu_p1r = xr(lu_p1);
        const bool iny1 = (k1 >= 3 && k1 <= p.ny - 2);

        // ---- depths (hh_init / hh_update, Function/depth.f90)
        hq_p2 = H_p2 + s_p2;
        hqp_p1 = H_p1 + R(sp_p1);
        {
            const R hq1r = xr(hq_p1), hqp1r = xr(hqp_p1), hqp0 = hqp_0;
            hu_p1 = (mb_p1 & MB_LLU) ? op.avg2(hq_p1, hq1r, lu_p1, lu_p1r) : R(0);
            hup_p1 = (mb_p1 & MB_LLU) ? op.avg2(hqp_p1, hqp1r, lu_p1, lu_p1r) : R(0);
            hv_p1 = (mb_p1 & MB_LLV) ? op.avg2(hq_p1, hq_p2, lu_p1, lu_p2) : R(0);
            hvp_0 = (mb_0 & MB_LLV) ? op.avg2(hqp0, hqp_p1, lu_0, lu_p1) : R(0);
            const R hq0r = xr(hq_0);
            hh_0 = (mb_0 & MB_LUH) ? op.avg4(hq_0, hq0r, hq_p1, hq1r, lu_0, lu_0r, lu_p1, lu_p1r) : R(0);
        }

        // ---- T-point stage of row n+1: continuity, new depths, tension, drag
        const R u1l = xl(u_p1), hu1l = xl(hu_p1), up1l = xl(R(up_p1));
        {
            const bool tw = (mb_p1 & MB_LU) && inx && iny1;
            const R wf = wf_p1;
            double sn;
            if constexpr (EXACT) {
                sn = sp_p1 + 2.0 * p.tau *
                                 (wf / p.rho0 * p.ffs -
                                  (u_p1 * hu_p1 * d - u1l * hu1l * d + v_p1 * hv_p1 * d -
                                   v_0 * hv_0 * d) /
                                      (d * d));
            } else {
#pragma clang fp contract(fast)
                const R inc = R(2.0 * p.tau) * (wf * R(p.ffs / p.rho0) -
                                                (u_p1 * hu_p1 - u1l * hu1l + v_p1 * hv_p1 - v_0 * hv_0) *
                                                    op.inv_d);
                sn = sp_p1 + double(inc);
            }
            sn_p1 = tw ? sn : 0.0;
        }
        hqn_p1 = H_p1 + R(sn_p1);
        {
            const R hqn1r = xr(hqn_p1);
            hun_p1 = (mb_p1 & MB_LLU) ? op.avg2(hqn_p1, hqn1r, lu_p1, lu_p1r) : R(0);
            hvn_0 = (mb_0 & MB_LLV) ? op.avg2(hqn_0, hqn_p1, lu_0, lu_p1) : R(0);
        }
        {
            const bool tw = (mb_p1 & MB_LU) && inx && iny1;
            R st, bf;
            const R upr = R(up_p1), vpr = R(vp_p1), vp0r = R(vp_0);
            bf = R(0);
            if constexpr (EXACT) {
                st = d / d * (upr / d - up1l / d) - d / d * (vpr / d - vp0r / d);
                if constexpr (FORM) {
                    R ut = (u_p1 * d * d * hu_p1 + u1l * d * d * hu1l) / 2.0 / d / d / hq_p1;
                    R vt = (v_p1 * d * d * hv_p1 + v_0 * d * d * hv_0) / 2.0 / d / d / hq_p1;
                    bf = R(p.cb) * sqrt(ut * ut + vt * vt + R(p.eb));
                }
            } else {
#pragma clang fp contract(fast)
                st = op.inv_d * ((upr - up1l) - (vpr - vp0r));
                if constexpr (FORM) {
                    const R rh = R(0.5) / hq_p1;
                    R ut = (u_p1 * hu_p1 + u1l * hu1l) * rh;
                    R vt = (v_p1 * hv_p1 + v_0 * hv_0) * rh;
                    bf = R(p.cb) * sqrt(ut * ut + vt * vt + R(p.eb));
                }
            }
            stt_p1 = tw ? st : R(0);
            bf_p1 = (tw && FORM) ? bf : R(0);
        }

        // ---- H-point stage of row n: shear and metric vorticity
        {
            const bool hw = (mb_0 & MB_LUU) && inx && (n >= 3 && n <= p.ny - 2);
            const R vp0 = R(vp_0), vp0r = xr(vp0);
            const R v0r = xr(v_0);
            R ss, vo;
            if constexpr (EXACT) {
                ss = d / d * (R(up_p1) / d - R(up_0) / d) + d / d * (vp0r / d - vp0 / d);
                vo = (v0r * d - v_0 * d) - (u_p1 * d - u_0 * d) - ((v0r - v_0) * d - (u_p1 - u_0) * d);
            } else {
#pragma clang fp contract(fast)
                ss = op.inv_d * ((R(up_p1) - R(up_0)) + (vp0r - vp0));
                vo = R(0);  // identically zero for uniform metrics (vel_ssh.f90:35-37)
            }
            sts_0 = hw ? ss : R(0);
            vo_0 = hw ? vo : R(0);
        }

        // ---- outputs of row n
        if (produce) {
            const bool iny = (n >= 3 && n <= p.ny - 2);
            const R lu_m1 = bitf<R>(mb_m1, MB_LU);
            const R mu_m1 = lu_m1 * nu, mu_0 = lu_0 * nu, mu_p1 = lu_p1 * nu;
            const R mu_m1r = xr(mu_m1), mu_0r = xr(mu_0), mu_p1r = xr(mu_p1);
            const R mu_0l = xl(mu_0), mu_p1l = xl(mu_p1);
            const R u0r = xr(u_0), u0l = xl(u_0), hu0r = xr(hu_0), hu0l = xl(hu_0);
            const R v0r = xr(v_0), v0l = xl(v_0), hv0r = xr(hv_0);
            const R vm1r = xr(v_m1), hvm1r = xr(hv_m1);
            const R up1l_ = xl(u_p1), hup1l = xl(hu_p1);
            const R s0r = xr(s_0), rd0r = xr(rd_0), bf0r = xr(bf_0);
            const R hq0r = xr(hq_0), stt0r = xr(stt_0);
            const R hh0l = xl(hh_0), sts0l = xl(sts_0), vo0l = xl(vo_0);
            const R up0 = R(up_0), vp0 = R(vp_0);
            const R luu0 = bitf<R>(mb_0, MB_LUU), luum1 = bitf<R>(mb_m1, MB_LUU);
            const R sx = sx_0, sy = sy_0;
            // the RHS of Control/dyncall.f90:114-126 is formed once per call
            // (first substep) and frozen over the call's substeps
            R rhs_u = q.fx0, rhs_v = q.fy0;
            if constexpr (FORM) {
                if constexpr (EXACT) {
                    rhs_u = 0.0 + (sx + (-(bf_0 + bf0r) / 2.0 * up0)) * d * d;
                    rhs_v = 0.0 + (sy + (-(bf_0 + bf_p1) / 2.0 * vp0)) * d * d;
                } else {
#pragma clang fp contract(fast)
                    rhs_u = sx - R(0.5) * (bf_0 + bf0r) * up0;
                    rhs_v = sy - R(0.5) * (bf_0 + bf_p1) * vp0;
                }
            }
            o.rhs_u = rhs_u;
            o.rhs_v = rhs_v;
            // level n-1 U/V depths: hh_init values at the first substep,
            // the hh_shift-filtered carry afterwards (depth.f90:195-203)
            R hupc = hup_0, hvpc = hvp_0;
            if constexpr (EXACT && frozen) {
                hupc = R(q.hupc0);
                hvpc = R(q.hvpc0);
            }
            R un, vn;
            if constexpr (EXACT) {
                // zonal momentum (vel_ssh.f90:536-551) with uv_trans (:59-81) and uv_diff2 (:146-159)
                R rhsx = rhs_u;
                R fx_p = (u_0 * d * hu_0 + u0r * d * hu0r) / 2.0 * (u_0 + u0r) / 2.0;
                R fx_m = (u_0 * d * hu_0 + u0l * d * hu0l) / 2.0 * (u_0 + u0l) / 2.0;
                R fy_p = (v_0 * d * hv_0 + v0r * d * hv0r) / 2.0 * (u_p1 + u_0) / 2.0 * luu0;
                R fy_m = (v_m1 * d * hv_m1 + vm1r * d * hvm1r) / 2.0 * (u_m1 + u_0) / 2.0 * luum1;
                R radv = -(fx_p - fx_m + fy_p - fy_m) +
                         (vo_0 * hh_0 * (v0r + v_0) + vo_m1 * hh_m1 * (vm1r + v_m1)) / 4.0;
                R muh_p = (mu_0 + mu_0r + mu_p1 + mu_p1r) / 4.0;
                R muh_m = (mu_0 + mu_0r + mu_m1 + mu_m1r) / 4.0;
                R pp = (d * d * mu_0r * hq0r * stt0r - d * d * mu_0 * hq_0 * stt_0) / d;
                R qq = (d * d * muh_p * hh_0 * sts_0 - d * d * muh_m * hh_m1 * sts_m1) / d;
                R rdfx = pp + qq;
                R bp = hun_0 * d * d / 2.0 / tau;
                R bp0 = hupc * d * d / 2.0 / tau;
                R slx = -G * (s0r - s_0) * d * hu_0;
                R grx = rhsx + slx + rdfx + radv - (rd_0 + rd0r) / 2.0 * up0 * d * d * hu_0 +
                        (f_0 * hh_0 * d * d * (v0r + v_0) + f_m1 * hh_m1 * d * d * (vm1r + v_m1)) / 4.0;
                un = (up0 * bp0 + grx) / bp;

                // meridional momentum (vel_ssh.f90:554-569) with uv_trans (:84-105) and uv_diff2 (:162-175)
                R rhsy = rhs_v;
                R gy_p = (v_0 * d * hv_0 + v_p1 * d * hv_p1) / 2.0 * (v_0 + v_p1) / 2.0;
                R gy_m = (v_0 * d * hv_0 + v_m1 * d * hv_m1) / 2.0 * (v_0 + v_m1) / 2.0;
                R gx_p = (u_0 * d * hu_0 + u_p1 * d * hu_p1) / 2.0 * (v0r + v_0) / 2.0;
                R gx_m = (u0l * d * hu0l + up1l_ * d * hup1l) / 2.0 * (v0l + v_0) / 2.0;
                R rady = -(gx_p - gx_m + gy_p - gy_m) -
                         (vo_0 * hh_0 * (u_p1 + u_0) + vo0l * hh0l * (up1l_ + u0l)) / 4.0;
                R nuh_p = (mu_0 + mu_0r + mu_p1 + mu_p1r) / 4.0;
                R nuh_m = (mu_0 + mu_0l + mu_p1 + mu_p1l) / 4.0;
                R py = (d * d * mu_p1 * hq_p1 * stt_p1 - d * d * mu_0 * hq_0 * stt_0) / d;
                R qy = (d * d * nuh_p * hh_0 * sts_0 - d * d * nuh_m * hh0l * sts0l) / d;
                R rdfy = -py + qy;
                R bq = hvn_0 * d * d / 2.0 / tau;
                R bq0 = hvpc * d * d / 2.0 / tau;
                R sly = -G * (s_p1 - s_0) * d * hv_0;
                R gry = rhsy + sly + rdfy + rady - (rd_0 + rd_p1) / 2.0 * vp0 * d * d * hv_0 -
                        (f_0 * hh_0 * d * d * (u_p1 + u_0) + f_0 * hh0l * d * d * (up1l_ + u0l)) / 4.0;
                vn = (vp0 * bq0 + gry) / bq;
            } else {
#pragma clang fp contract(fast)
                const R id = op.inv_d;
                const R two_tau = R(2.0) * tau;
                // forces per unit area (the reference's terms divided by dxt*dyh)
                R adv = R(-0.25) * ((u_0 * hu_0 + u0r * hu0r) * (u_0 + u0r) -
                                    (u_0 * hu_0 + u0l * hu0l) * (u_0 + u0l) +
                                    luu0 * (v_0 * hv_0 + v0r * hv0r) * (u_p1 + u_0) -
                                    luum1 * (v_m1 * hv_m1 + vm1r * hvm1r) * (u_m1 + u_0));
                R muh_p = R(0.25) * (mu_0 + mu_0r + mu_p1 + mu_p1r);
                R muh_m = R(0.25) * (mu_0 + mu_0r + mu_m1 + mu_m1r);
                R dif = (mu_0r * hq0r * stt0r - mu_0 * hq_0 * stt_0) +
                        (muh_p * hh_0 * sts_0 - muh_m * hh_m1 * sts_m1);
                R F = rhs_u - G * id * (s0r - s_0) * hu_0 +
                      id * (adv + dif) - R(0.5) * (rd_0 + rd0r) * up0 * hu_0 +
                      R(0.25) * (f_0 * hh_0 * (v0r + v_0) + f_m1 * hh_m1 * (vm1r + v_m1));
                un = (up0 * hupc + two_tau * F) / hun_0;

                R advy = R(-0.25) * ((u_0 * hu_0 + u_p1 * hu_p1) * (v0r + v_0) -
                                     (u0l * hu0l + up1l_ * hup1l) * (v0l + v_0) +
                                     (v_0 * hv_0 + v_p1 * hv_p1) * (v_0 + v_p1) -
                                     (v_0 * hv_0 + v_m1 * hv_m1) * (v_0 + v_m1));
                R nuh_p = R(0.25) * (mu_0 + mu_0r + mu_p1 + mu_p1r);
                R nuh_m = R(0.25) * (mu_0 + mu_0l + mu_p1 + mu_p1l);
                R dify = -(mu_p1 * hq_p1 * stt_p1 - mu_0 * hq_0 * stt_0) +
                         (nuh_p * hh_0 * sts_0 - nuh_m * hh0l * sts0l);
                R Fy = rhs_v - G * id * (s_p1 - s_0) * hv_0 +
                       id * (advy + dify) - R(0.5) * (rd_0 + rd_p1) * vp0 * hv_0 -
                       R(0.25) * f_0 * (hh_0 * (u_p1 + u_0) + hh0l * (up1l_ + u0l));
                vn = (vp0 * hvpc + two_tau * Fy) / hvn_0;
            }
            const bool uw = (mb_0 & MB_LCU) && inx && iny;
            const bool vw = (mb_0 & MB_LCV) && inx && iny;
            un = uw ? un : R(0);
            vn = vw ? vn : R(0);

            // time-level shift with the Robert-Asselin filter (vel_ssh.f90:588-612)
            R s_new = s_0, u_new = u_0, v_new = v_0;
            double sp_new = sp_0, up_new = up_0, vp_new = vp_0;
            if (inx1 && n >= 2 && n <= p.ny - 1) {
                const double ts = p.ts, cts = 0.5 * p.ts / p.dn;
                if (mb_0 & MB_LU) {
                    if constexpr (EXACT) {
                        sp_new = s_0 + ts * (sn_0 - 2.0 * s_0 + sp_0) / 2.0 / p.dn;
                    } else {
#pragma clang fp contract(fast)
                        sp_new = double(s_0) + cts * (sn_0 - 2.0 * double(s_0) + sp_0);
                    }
                    s_new = R(sn_0);
                }
                if (mb_0 & MB_LCU) {
                    if constexpr (EXACT) {
                        up_new = u_0 + ts * (un - 2.0 * u_0 + up_0) / 2.0 / p.dn;
                    } else {
#pragma clang fp contract(fast)
                        up_new = double(u_0) + cts * (double(un) - 2.0 * double(u_0) + up_0);
                    }
                    u_new = un;
                }
                if (mb_0 & MB_LCV) {
                    if constexpr (EXACT) {
                        vp_new = v_0 + ts * (vn - 2.0 * v_0 + vp_0) / 2.0 / p.dn;
                    } else {
#pragma clang fp contract(fast)
                        vp_new = double(v_0) + cts * (double(vn) - 2.0 * double(v_0) + vp_0);
                    }
                    v_new = vn;
                }
            }
            o.s = s_new;
            o.u = u_new;
            o.v = v_new;
            o.sp = sp_new;
            o.up = up_new;
            o.vp = vp_new;
            if constexpr (EXACT) {
                // hh_shift of the U/V depths for the next substep (depth.f90:195-203)
                const double ts = p.ts;
                double hu_new = hupc, hv_new = hvpc;
                if (inx1 && n >= 2 && n <= p.ny - 1) {
                    if (mb_0 & MB_LLU) hu_new = hu_0 + ts * (hun_0 - 2.0 * hu_0 + hupc) / 2.0 / p.dn;
                    if (mb_0 & MB_LLV) hv_new = hv_0 + ts * (hvn_0 - 2.0 * hv_0 + hvpc) / 2.0 / p.dn;
                }
                o.hu = hu_new;
                o.hv = hv_new;
            }
        }

        // ---- advance the window one row
        u_m1 = u_0; u_0 = u_p1;
        v_m1 = v_0; v_0 = v_p1;
        s_0 = s_p1; s_p1 = s_p2;
        H_p1 = H_p2;
        up_0 = up_p1; vp_0 = vp_p1; sp_0 = sp_p1; rd_0 = rd_p1;
        mb_m1 = mb_0; mb_0 = mb_p1; mb_p1 = mb_p2;
        hq_0 = hq_p1; hq_p1 = hq_p2;
        hqp_0 = hqp_p1; hqn_0 = hqn_p1;
        hu_0 = hu_p1; hup_0 = hup_p1;
        hv_m1 = hv_0; hv_0 = hv_p1;
        hh_m1 = hh_0;
        sn_0 = sn_p1; hun_0 = hun_p1;
        stt_0 = stt_p1; bf_0 = bf_p1;
        sts_m1 = sts_0; vo_m1 = vo_0;
    }
};

// Loads of one iteration n of a strip (issued one iteration ahead of their
// use: measured on MI355X at 2048 x 1536, two iterations ahead runs 5 %
// slower in FP32 and halves FP64's occupancy).
template <typename R, bool EXACT, bool FROZEN>
struct Loader {
    const FusedParams &p;
    const RowAddr ra;
    const bool colok;
    struct Pre {
        R s, H, u, v, rd, fx, fy;
        double up, vp, sp, hupc, hvpc, f, sx, sy, wf;
        int mb;
    } q;

    __device__ Loader(const FusedParams &pp, int m)
        : p(pp), ra{pp.nx, pp.ny, min(max(m, 1), pp.nx)}, colok(m >= 1 && m <= pp.nx)
    {
    }
    __device__ __forceinline__ bool rowok(int row) const { return colok && row >= 1 && row <= p.ny; }
    __device__ __forceinline__ double tab(const double *t, int row) const
    {
        return t[min(max(row, 1), p.ny) - 1];
    }
    __device__ __forceinline__ void issue(int n)  // the loads iteration n consumes
    {
        const R *S = static_cast<const R *>(p.s), *Hh = static_cast<const R *>(p.H);
        const R *U = static_cast<const R *>(p.u), *V = static_cast<const R *>(p.v);
        const R *RD = static_cast<const R *>(p.rd);
        const R *FX = static_cast<const R *>(p.frx_in), *FY = static_cast<const R *>(p.fry_in);
        const size_t i2 = ra.at(n + 2), i1 = ra.at(n + 1), i0 = ra.at(n);
        q.s = S[i2];
        q.H = Hh[i2];
        q.mb = p.mb[i2];
        q.u = U[i1];
        q.v = V[i1];
        q.up = p.up[i1];
        q.vp = p.vp[i1];
        q.sp = p.sp[i1];
        q.rd = RD[i1];
        // per-cell frozen forcing and carried depths of the output row n
        if constexpr (FROZEN) {
            q.fx = FX[i0];
            q.fy = FY[i0];
        } else {
            q.fx = q.fy = R(0);
        }
        if constexpr (EXACT && FROZEN) {
            q.hupc = p.hup_c[i0];
            q.hvpc = p.hvp_c[i0];
        } else {
            q.hupc = q.hvpc = 0.0;
        }
        q.f = tab(p.row_f, n);
        q.sx = tab(p.row_sx, n);
        q.sy = tab(p.row_sy, n);
        q.wf = tab(p.row_wf, n + 1);
    }
    // the prefetched values of iteration n, zeroed outside the array
    __device__ __forceinline__ void take(int n, StageIn<R> &in) const
    {
        const int k1 = n + 1;
        in.s2 = rowok(n + 2) ? q.s : R(0);
        in.H2 = rowok(n + 2) ? q.H : R(0);
        in.mb2 = rowok(n + 2) ? q.mb : 0;
        in.u1 = rowok(k1) ? q.u : R(0);
        in.v1 = rowok(k1) ? q.v : R(0);
        in.up1 = rowok(k1) ? q.up : 0.0;
        in.vp1 = rowok(k1) ? q.vp : 0.0;
        in.sp1 = rowok(k1) ? q.sp : 0.0;
        in.rd1 = rowok(k1) ? q.rd : R(0);
        in.fx0 = q.fx;
        in.fy0 = q.fy;
        in.hupc0 = q.hupc;
        in.hvpc0 = q.hvpc;
        in.f0 = (n >= 1 && n <= p.ny) ? R(q.f) : R(0);
        in.sx0 = R(q.sx);
        in.sy0 = R(q.sy);
        in.wf1 = (k1 >= 1 && k1 <= p.ny) ? R(q.wf) : R(0);
    }
};

// Store policy (measured, see st_out): level-n outputs non-temporal in the
// FP64 modes, the FP64 filtered level non-temporal in every mode.
template <typename R>
__device__ __forceinline__ void store_state(const FusedParams &p, size_t i, const StageOut<R> &o)
{
    constexpr bool NTS = sizeof(R) == 8;
    st_out<NTS, R>(p.so, i, o.s);
    st_out<NTS, R>(p.uo, i, o.u);
    st_out<NTS, R>(p.vo, i, o.v);
    st_out<true, double>(p.spo, i, o.sp);
    st_out<true, double>(p.upo, i, o.up);
    st_out<true, double>(p.vpo, i, o.vp);
}

// One work item: `chunk` output rows of one 60-column strip of one basin,
// substep p.sub (FORM: p.sub == 1).  Called by every lane of a wave with
// wave-uniform arguments.
template <typename R, bool EXACT, bool FORM>
__device__ __forceinline__ void fused_chunk(const FusedParams &p, int chunk_id, int strip)
{
    const int lane = threadIdx.x;
    const int m = strip * SW + lane - 1;
    const bool colok = (m >= 1 && m <= p.nx);
    const bool writer = colok && lane >= 2 && lane <= SW + 1;
    const int r0 = chunk_id * p.chunk + 1;
    const int r1 = min(r0 + p.chunk, p.ny + 1);

    Stage<R, EXACT, FORM> st(p, m);
    Loader<R, EXACT, !FORM> ld(p, m);
    StageIn<R> in;
    StageOut<R> o;
    ld.issue(r0 - 3);
    for (int n = r0 - 3; n < r1; ++n) {
        ld.take(n, in);
        if (n + 1 < r1) ld.issue(n + 1);
        st.row(n, in, n >= r0, o);
        if (n >= r0 && writer && n <= p.ny) {
            const size_t io = (size_t)(n - 1) * p.nx + (m - 1);
            // the forcing formed by the first substep, for the later ones
            if (FORM && p.nsub > 1) {
                st_out<false, R>(p.frx, io, o.rhs_u);
                st_out<false, R>(p.fry, io, o.rhs_v);
            }
            if constexpr (EXACT) {
                if (p.nsub > 1 && p.sub < p.nsub) {
                    st_out<sizeof(R) == 8, double>(p.hup_c, io, o.hu);
                    st_out<sizeof(R) == 8, double>(p.hvp_c, io, o.hv);
                }
            }
            store_state<R>(p, io, o);
        }
    }
}

// Two substeps, p.sub and p.sub + 1, in one march (AFORM: p.sub == 1):
// stage A reads the arrays; stage B runs two rows behind A on A's outputs
// and on the static rows A loaded two iterations earlier (depth at rest,
// mask bits, Rayleigh coefficient), which never leave the registers
// (re-reading them from cache instead was measured 5 % slower).  A strip
// is 58 output columns (B's output at lane L needs
// A's lanes L-1..L+2, which need the loads of L-2..L+4); A starts two rows
// above the chunk so that B's window sees A's rows from r0-2 on.  The
// results are those of two launches of fused_chunk.
template <typename R, bool EXACT, bool AFORM>
__device__ __forceinline__ void fused_chunk2(const FusedParams &p, int chunk_id, int strip)
{
    const int lane = threadIdx.x;
    const int m = strip * SW2 + lane - 1;
    const bool colok = (m >= 1 && m <= p.nx);
    const bool writer = colok && lane >= 2 && lane <= SW2 + 1;
    const int r0 = chunk_id * p.chunk + 1;
    const int r1 = min(r0 + p.chunk, p.ny + 1);
    const int subB = p.sub + 1;

    Stage<R, EXACT, AFORM> A(p, m);
    Stage<R, EXACT, false> B(p, m);
    Loader<R, EXACT, !AFORM> ld(p, m);
    StageIn<R> ia, ib;
    StageOut<R> oa, ob;
    // delay lines from A's iteration to B's (B's iteration n - 2 runs in
    // A's iteration n): the static rows A loaded, A's forcing / carried
    // depths, and A's outputs of row n - 1 (B's level-n input one row ahead)
    R H_d1 = 0, H_d2 = 0, rd_d1 = 0, rd_d2 = 0;
    int mb_d1 = 0, mb_d2 = 0;
    R fx_d1 = 0, fx_d2 = 0, fy_d1 = 0, fy_d2 = 0;
    double hu_d1 = 0, hu_d2 = 0, hv_d1 = 0, hv_d2 = 0;
    R u_prev = 0, v_prev = 0;
    double up_prev = 0, vp_prev = 0, sp_prev = 0;
    ld.issue(r0 - 5);
    for (int n = r0 - 5; n < r1 + 2; ++n) {
        const int nb = n - 2;
        const bool runB = nb >= r0 - 3;
        ld.take(n, ia);
        if (n + 1 < r1 + 2) ld.issue(n + 1);
        const bool keepA = n >= r0 - 2;
        A.row(n, ia, keepA, oa);
        // A's row n as the 2-launch sequence would have stored it: zero
        // outside the array, as B's loads would read it
        const bool okn = ld.rowok(n) && keepA;
        if (runB) {
            ib.s2 = okn ? oa.s : R(0);
            ib.H2 = H_d2;
            ib.mb2 = mb_d2;
            ib.u1 = u_prev;
            ib.v1 = v_prev;
            ib.up1 = up_prev;
            ib.vp1 = vp_prev;
            ib.sp1 = sp_prev;
            ib.rd1 = rd_d2;
            ib.fx0 = fx_d2;
            ib.fy0 = fy_d2;
            ib.hupc0 = hu_d2;
            ib.hvpc0 = hv_d2;
            ib.f0 = (nb >= 1 && nb <= p.ny) ? R(ld.tab(p.row_f, nb)) : R(0);
            ib.sx0 = R(ld.tab(p.row_sx, nb));
            ib.sy0 = R(ld.tab(p.row_sy, nb));
            ib.wf1 = (nb + 1 >= 1 && nb + 1 <= p.ny) ? R(ld.tab(p.row_wf, nb + 1)) : R(0);
            const bool produce = nb >= r0;
            B.row(nb, ib, produce, ob);
            if (produce && writer && nb <= p.ny) {
                const size_t io = (size_t)(nb - 1) * p.nx + (m - 1);
                if (AFORM && p.nsub > subB) {
                    st_out<false, R>(p.frx, io, ob.rhs_u);
                    st_out<false, R>(p.fry, io, ob.rhs_v);
                }
                if constexpr (EXACT) {
                    if (subB < p.nsub) {
                        st_out<sizeof(R) == 8, double>(p.hup_c, io, ob.hu);
                        st_out<sizeof(R) == 8, double>(p.hvp_c, io, ob.hv);
                    }
                }
                store_state<R>(p, io, ob);
            }
        }
        // shift the delay lines
        H_d2 = H_d1; H_d1 = ia.H2;
        mb_d2 = mb_d1; mb_d1 = ia.mb2;
        rd_d2 = rd_d1; rd_d1 = ia.rd1;
        fx_d2 = fx_d1; fx_d1 = keepA ? oa.rhs_u : R(0);
        fy_d2 = fy_d1; fy_d1 = keepA ? oa.rhs_v : R(0);
        if constexpr (EXACT) {
            hu_d2 = hu_d1; hu_d1 = keepA ? oa.hu : 0.0;
            hv_d2 = hv_d1; hv_d1 = keepA ? oa.hv : 0.0;
        }
        u_prev = okn ? oa.u : R(0);
        v_prev = okn ? oa.v : R(0);
        up_prev = okn ? oa.up : 0.0;
        vp_prev = okn ? oa.vp : 0.0;
        sp_prev = okn ? oa.sp : 0.0;
    }
}

// ------------------------------------------------------------------
// K substeps per launch as a pipeline of K waves (one workgroup of 64 K
// threads per work item): wave w runs substep sub + w on row n - 3 w at
// iteration n, on what wave w - 1 produced in the three iterations before
// (its rows n - 3 w + 2, + 1 and + 0), handed over through a ring of
// PIPE_SLOTS rows in LDS with one workgroup barrier per iteration.  Each
// wave holds one substep's window (registers of one stage), the
// intermediate levels never leave the chip, and wave 0 alone reads the
// arrays; the last wave writes them.  The results are those of K launches
// of fused_chunk bit for bit (the same Stage and Loader code).

constexpr int PIPE_SLOTS = 4;
// Rows per unrolled group of the pipeline's march (the window's registers
// rotate by renaming instead of moves).  Measured on MI355X at 2048 x 1536,
// us per substep with 1 / 2 / 4 rows: FP64 61.7 / 58.5 / 57.1, FP32
// 50.1 / 49.0 / 51.3 (the 64-member FP32 ensemble 820 / 801 / 830 us per
// outer step).
template <typename R>
__host__ __device__ constexpr int pipe_unroll()
{
    return sizeof(R) == 8 ? 4 : 2;
}
// output columns per strip with K stages: stage w's results are valid on
// lanes w + 1 .. 61 - 2 w (a stage's result at lane j needs its inputs at
// j - 1 .. j + 2); the last stage writes lanes K .. 63 - 2 K
__host__ __device__ constexpr int pipe_width(int K) { return WAVE - 3 * K; }
// iterations of a work item beyond its rows: the first stage starts
// 2 K + 1 rows above the chunk, the last ends 3 (K - 1) rows later
__host__ __device__ constexpr int pipe_extra(int K) { return 5 * K - 2; }

// What stage w hands to stage w + 1 for one row (each lane its column).
template <typename R>
struct PipeSlot {
    R s[WAVE], u[WAVE], v[WAVE], fx[WAVE], fy[WAVE], H[WAVE], rd[WAVE];
    double sp[WAVE], up[WAVE], vp[WAVE];
    uint8_t mb[WAVE];
};

template <typename R, int K, bool LOAD, bool FORM>
__device__ __forceinline__ void fused_pipe_stage(const FusedParams &p, int chunk_id, int strip, int w,
                                                 bool save_forcing, PipeSlot<R> (*ring)[PIPE_SLOTS])
{
    constexpr int SWK = pipe_width(K);
    const int lane = threadIdx.x % WAVE;
    const int m = strip * SWK + lane - (K - 1);
    const bool colok = (m >= 1 && m <= p.nx);
    const bool writer = colok && lane >= K && lane <= WAVE - 1 - 2 * K;
    const int r0 = chunk_id * p.chunk + 1;
    const int r1 = min(r0 + p.chunk, p.ny + 1);
    constexpr int U = pipe_unroll<R>();
    const int n0 = r0 - 2 * K - 1;  // first iteration
    // iterations n0..n1, a whole number of unrolled groups of U (the rows
    // past the chunk that this adds are computed, not stored)
    const int n1 = n0 + (r1 + 3 * (K - 1) - n0 + U - 1) / U * U - 1;
    const int valid_from = n0 + 3 + 2 * w;  // first row of valid results
    const bool last = (w == K - 1);

    Stage<R, false, FORM> st(p, m);
    Loader<R, false, !FORM> ld(p, m);
    StageIn<R> in;
    StageOut<R> o;
    // the row's own static fields for the next stage (inputs arrive two rows,
    // and the Rayleigh coefficient one row, ahead)
    R H_d1 = 0, H_d2 = 0, rd_d1 = 0;
    int mb_d1 = 0, mb_d2 = 0;
    if constexpr (LOAD) ld.issue(n0);
    for (int nb = n0; nb <= n1; nb += U)
#pragma unroll
    for (int n = nb; n < nb + U; ++n) {
        const int r = n - 3 * w;
        if constexpr (LOAD) {
            ld.take(r, in);
            if (r + 1 <= n1) ld.issue(r + 1);
        } else {
            const PipeSlot<R> &a = ring[w - 1][(n - 1) & (PIPE_SLOTS - 1)];  // row r + 2
            const PipeSlot<R> &b = ring[w - 1][(n - 2) & (PIPE_SLOTS - 1)];  // row r + 1
            const PipeSlot<R> &c = ring[w - 1][(n - 3) & (PIPE_SLOTS - 1)];  // row r
            in.s2 = a.s[lane];
            in.H2 = a.H[lane];
            in.mb2 = a.mb[lane];
            in.u1 = b.u[lane];
            in.v1 = b.v[lane];
            in.up1 = b.up[lane];
            in.vp1 = b.vp[lane];
            in.sp1 = b.sp[lane];
            in.rd1 = b.rd[lane];
            in.fx0 = c.fx[lane];
            in.fy0 = c.fy[lane];
            in.hupc0 = in.hvpc0 = 0.0;
            in.f0 = (r >= 1 && r <= p.ny) ? R(ld.tab(p.row_f, r)) : R(0);
            in.sx0 = R(ld.tab(p.row_sx, r));
            in.sy0 = R(ld.tab(p.row_sy, r));
            in.wf1 = (r + 1 >= 1 && r + 1 <= p.ny) ? R(ld.tab(p.row_wf, r + 1)) : R(0);
        }
        const bool keep = last ? (r >= r0) : (r >= valid_from);
        st.row(r, in, keep, o);
        if (!last) {
            // row r for stage w + 1, zero outside the array or before the
            // results are valid (what its loads would have read there)
            PipeSlot<R> &q = ring[w][n & (PIPE_SLOTS - 1)];
            const bool ok = ld.rowok(r) && keep;
            q.s[lane] = ok ? o.s : R(0);
            q.u[lane] = ok ? o.u : R(0);
            q.v[lane] = ok ? o.v : R(0);
            q.sp[lane] = ok ? o.sp : 0.0;
            q.up[lane] = ok ? o.up : 0.0;
            q.vp[lane] = ok ? o.vp : 0.0;
            q.fx[lane] = keep ? o.rhs_u : R(0);
            q.fy[lane] = keep ? o.rhs_v : R(0);
            q.H[lane] = H_d2;
            q.rd[lane] = rd_d1;
            q.mb[lane] = (uint8_t)mb_d2;
        } else if (keep && writer && r < r1) {
            const size_t io = (size_t)(r - 1) * p.nx + (m - 1);
            // the forcing of a call that continues after this launch
            if (save_forcing) {
                st_out<false, R>(p.frx, io, o.rhs_u);
                st_out<false, R>(p.fry, io, o.rhs_v);
            }
            store_state<R>(p, io, o);
        }
        H_d2 = H_d1;
        H_d1 = in.H2;
        mb_d2 = mb_d1;
        mb_d1 = in.mb2;
        rd_d1 = in.rd1;
        __syncthreads();
    }
}

// One work item of the K-stage pipeline; every wave of the workgroup
// calls it.  form_mask bit w: stage w is the first substep of its call
// (forms the wind and drag forcing); save_forcing: the last stage's call
// continues after this launch.
template <typename R, int K>
__device__ __forceinline__ void fused_pipe(const FusedParams &p, int chunk_id, int strip, int form_mask,
                                           bool save_forcing, PipeSlot<R> (*ring)[PIPE_SLOTS])
{
    // the stage index as a scalar: each wave takes one branch, the same
    // number of barriers in every branch
    const int w = __builtin_amdgcn_readfirstlane((int)(threadIdx.x / WAVE));
    const bool form = (form_mask >> w) & 1;
    if (w == 0) {
        if (form)
            fused_pipe_stage<R, K, true, true>(p, chunk_id, strip, w, save_forcing, ring);
        else
            fused_pipe_stage<R, K, true, false>(p, chunk_id, strip, w, save_forcing, ring);
    } else {
        if (form)
            fused_pipe_stage<R, K, false, true>(p, chunk_id, strip, w, save_forcing, ring);
        else
            fused_pipe_stage<R, K, false, false>(p, chunk_id, strip, w, save_forcing, ring);
    }
}

}  // namespace ibt
