// The reference's barotropic_dynamics (Function/vel_ssh.f90:379-679) as one
// fused kernel launch per leap-frog substep, behind the reference's own
// boundary: 2*nstep substeps (nstep state-advancing, nstep look-ahead),
// the state after substep nstep returned as the call's result (:638-648),
// the averages ubrtr_i, vbrtr_i, ssh4gradx, ssh4grady accumulated over all
// 2*nstep substeps (:549-550, :567-568), the caller's RHS, water flux,
// viscosity and Rayleigh coefficient per cell, and any metrics.  The
// depths at entry are hh_init(ssh_e, sshp_e) as the reference holds them
// between calls (depth.f90:2-105, vel_ssh.f90:668-675); nothing of them is
// stored: level-n depths are formed from the surface height and the
// filtered U/V depths are carried between substeps as hh_shift filters
// them (depth.f90:174-220).
//
// Arithmetic: the reference's expressions operation by operation, in its
// evaluation order, without fused multiply-add (the library is compiled
// with -ffp-contract=off): the results are the reference's bit for bit.
//
// Work decomposition as in the fused step (bt_fused.hip): one 64-lane wave
// owns a strip of 60 output columns (lanes 0-1 and 62-63 are halo) and
// marches down a chunk of rows with a rolling register window three rows
// deep; x-neighbours come from the adjacent lanes by DPP row shifts.  An
// intermediate that the reference evaluates at a neighbouring point is
// taken from the neighbour's lane (or the previous row) only where the
// reference's expression there is the same, operand for operand, up to the
// order of an addition of two terms (fx_m(m) = fx_p(m-1), fy_m(n) =
// fy_p(n-1)); everything else is evaluated as written.
//
// Eligible calls: full_free_surface = 1, no periodicity, and land on the
// two halo cells of every side (the reference's loop bounds then coincide
// with its masks).  Other calls take the stage-by-stage path.  With ksw4 >
// 0 the biharmonic term uv_diff4 (vel_ssh.f90:187-317) of each substep is
// formed before its launch by the reference-order kernels of the
// stage-by-stage path from the substep's entering state (its two terms per
// component, lit_diff4_terms) and added by the kernel to its uv_diff2 result
// in the reference's order.
// Metrics: per cell (any grid), or per row when every metric and the
// Coriolis parameter are uniform along x (Cartesian and longitude-latitude
// grids): then they are read from row tables instead of 8 + 1 arrays.
#include <cmath>
#include <cstring>
#include <mutex>
#include <vector>

#include "bt_dropin.h"
#include "bt_fused_body.h"
#include "bt_literal.h"

namespace ibt {

struct DropinParams {
    int nx, ny, strips, chunk, nchunks, total, per_xcd;
    const uint8_t *mb;
    const double *H, *rd, *mu, *wf, *rx, *ry;
    // dx, dy, dxt, dyt, dxh, dyh, dxb, dyb and rlh_s: per cell, or (ROWMET)
    // per row, row n at [n - 1]
    const double *met[9];
    const double *u, *v, *s, *up, *vp, *sp;  // state entering the substep
    double *uo, *vo, *so, *upo, *vpo, *spo;  // state after it (uo null: not written)
    double *cu, *cv, *cs, *cup, *cvp, *csp;  // copy of the entering state (cu null: none)
    double *hupc, *hvpc;                     // filtered U/V depths carried in place (substeps > 1 read)
    bool carry_out;                          // write the filtered U/V depths for the next substep
    double *acc[4];                          // ubrtr_i, vbrtr_i, ssh4gradx, ssh4grady, in place
    const double *d4[4];                     // ksw4 > 0: uv_diff4's terms px, qx, py, qy of the substep
    double g, rho0, ts, tau, ffs, dn, d4n;
};

// What iteration n consumes (zeroed outside the array): row n+2 of the
// surface height, depth at rest, masks, dx, dy; row n+1 of the level-n and
// filtered velocities and height, Rayleigh coefficient, viscosity, water
// flux, dxt, dyh, dxh, dyt; row n of the Coriolis parameter, the caller's
// RHS, dxb, dyb, the carried depths and the accumulators.
struct DIn {
    double s2, H2, dx2, dy2;
    int mb2;
    double u1, v1, up1, vp1, sp1, rd1, mu1, wf1, dxt1, dyh1, dxh1, dyt1;
    double f0, rx0, ry0, dxb0, dyb0, hupc0, hvpc0, a0, a1, a2, a3;
    double d4[4];  // uv_diff4's terms of row n (D4)
};

// What one substep produces for row n.
struct DOut {
    double s, u, v, sp, up, vp;  // state after the Robert-Asselin shift
    double hu, hv;               // hh_shift-filtered U/V depths
    double a0, a1, a2, a3;       // accumulators
};

// One substep marching down the rows of one strip.  Suffixes: m1 = row
// n-1, 0 = row n, p1 = row n+1, p2 = row n+2.  FIRST: the call's first
// substep (filtered depths from hh_init, accumulators start at zero).  D4:
// the biharmonic viscosity is on (ksw4 > 0); its terms come precomputed
// per cell and are added to the uv_diff2 result in the reference's order.
template <bool FIRST, bool D4>
struct DStage {
    const DropinParams &p;
    const bool inx, inx1;

    double u_0 = 0, u_p1 = 0, v_0 = 0, v_p1 = 0, s_0 = 0, s_p1 = 0, s_p2 = 0, H_p1 = 0, H_p2 = 0;
    double up_0 = 0, up_p1 = 0, vp_0 = 0, vp_p1 = 0, sp_0 = 0, sp_p1 = 0, rd_0 = 0, rd_p1 = 0;
    double mu_m1 = 0, mu_0 = 0, mu_p1 = 0;
    int mb_m1 = 0, mb_0 = 0, mb_p1 = 0, mb_p2 = 0;
    double dx_p1 = 0, dx_p2 = 0, dy_p1 = 0, dy_p2 = 0;
    double dxt_0 = 0, dxt_p1 = 0, dyh_0 = 0, dyh_p1 = 0, dxh_0 = 0, dxh_p1 = 0, dyt_0 = 0, dyt_p1 = 0;
    double dxb_m1 = 0;
    double hq_0 = 0, hq_p1 = 0, hq_p2 = 0;  // level-n T depth
    double w_0 = 0, w_p1 = 0, w_p2 = 0;     // hq*dx*dy*lu (depth.f90 area weights)
    double wp_0 = 0, wp_p1 = 0;             // the same for the filtered level (hh_init)
    double wn_0 = 0, wn_p1 = 0;             // the same for the new level (hh_update)
    double hu_0 = 0, hu_p1 = 0, hup_0 = 0, hup_p1 = 0, hv_0 = 0, hv_p1 = 0, hun_0 = 0, hun_p1 = 0;
    double hh_m1 = 0, hh_0 = 0;
    double sn_0 = 0, sn_p1 = 0;
    double P_0 = 0, Px_0 = 0;      // dy^2*mu*hq*str_t and dx^2*mu*hq*str_t on T points
    double sts_m1 = 0;             // str_s of row n-1
    double Cc_m1 = 0, Fc_m1 = 0;   // row n-1 vorticity and Coriolis terms of the zonal equation
    double A_0 = 0;                // u*dyh*hu on U points
    double Bv_0 = 0;               // v*dxh*hv on V points
    double fv_0 = 0;               // v*hv*dxh on V points (continuity)
    double b_0 = 0, c_0 = 0, h_0 = 0;  // vp/dxh (V), up/dxt and u*dxt (U)
    double fyU_m1 = 0, fyV_m1 = 0;     // fy_p of row n-1 (zonal and meridional equations)

    __device__ DStage(const DropinParams &pp, int m)
        : p(pp), inx(m >= 3 && m <= pp.nx - 2), inx1(m >= 2 && m <= pp.nx - 1)
    {
    }

    // hh_init / hh_update's average onto a U or V point of two T area
    // weights, and onto an H point of four (Function/depth.f90:29-66):
    // (w_a + w_b) / slu / metric_1 / metric_2, slu the number of wet T cells
    __device__ static double avg(double wsum, double slu, double m1, double m2) { return wsum / slu / m1 / m2; }

    __device__ __forceinline__ void row(int n, const DIn &q, bool produce, DOut &o)
    {
        const double G = p.g, tau = p.tau;
        const int k1 = n + 1;
        s_p2 = q.s2;
        H_p2 = q.H2;
        mb_p2 = q.mb2;
        dx_p2 = q.dx2;
        dy_p2 = q.dy2;
        u_p1 = q.u1;
        v_p1 = q.v1;
        up_p1 = q.up1;
        vp_p1 = q.vp1;
        sp_p1 = q.sp1;
        rd_p1 = q.rd1;
        mu_p1 = q.mu1;
        dxt_p1 = q.dxt1;
        dyh_p1 = q.dyh1;
        dxh_p1 = q.dxh1;
        dyt_p1 = q.dyt1;
        const double dxb_0 = q.dxb0, dyb_0 = q.dyb0;

        const double lu_0 = bitf<double>(mb_0, MB_LU), lu_p1 = bitf<double>(mb_p1, MB_LU);
        const double lu_p2 = bitf<double>(mb_p2, MB_LU);
        const double lu_0r = xr(lu_0), lu_p1r = xr(lu_p1);
        const bool iny1 = (k1 >= 3 && k1 <= p.ny - 2);

        // ---- depths at level n (hh_init at the first substep, hh_update's
        // results of the previous substep afterwards: both are these
        // averages of hq = H + ssh, depth.f90:21-66, 122-150)
        hq_p2 = H_p2 + s_p2;
        w_p2 = hq_p2 * dx_p2 * dy_p2 * lu_p2;
        {
            const double w1r = xr(w_p1);
            hu_p1 = (mb_p1 & MB_LLU) ? avg(w_p1 + w1r, lu_p1 + lu_p1r, dxt_p1, dyh_p1) : 0.0;
            hv_p1 = (mb_p1 & MB_LLV) ? avg(w_p1 + w_p2, lu_p1 + lu_p2, dxh_p1, dyt_p1) : 0.0;
            const double w0r = xr(w_0);
            hh_0 = (mb_0 & MB_LUH) ? avg(w_0 + w0r + w_p1 + w1r, lu_0 + lu_0r + lu_p1 + lu_p1r, dxb_0, dyb_0)
                                   : 0.0;
        }
        double hvp_0 = 0.0;
        if constexpr (FIRST) {
            // filtered-level depths of hh_init (depth.f90:22, :32-34, :41-43)
            const double hqp_p1 = H_p1 + sp_p1;
            wp_p1 = hqp_p1 * dx_p1 * dy_p1 * lu_p1;
            hup_p1 = (mb_p1 & MB_LLU) ? avg(wp_p1 + xr(wp_p1), lu_p1 + lu_p1r, dxt_p1, dyh_p1) : 0.0;
            hvp_0 = (mb_0 & MB_LLV) ? avg(wp_0 + wp_p1, lu_0 + lu_p1, dxh_0, dyt_0) : 0.0;
        }

        // ---- T point of row n+1: continuity (vel_ssh.f90:487-499)
        const double fu_p1 = u_p1 * hu_p1 * dyh_p1;
        const double fv_p1 = v_p1 * hv_p1 * dxh_p1;
        {
            const bool tw = (mb_p1 & MB_LU) && inx && iny1;
            const double sn = sp_p1 + 2.0 * tau *
                                          (q.wf1 / p.rho0 * p.ffs -
                                           (fu_p1 - xl(fu_p1) + fv_p1 - fv_0) / (dx_p1 * dy_p1));
            sn_p1 = tw ? sn : 0.0;
        }
        // new-level depths (hh_update, depth.f90:122-150)
        const double hqn_p1 = H_p1 + sn_p1;
        wn_p1 = hqn_p1 * dx_p1 * dy_p1 * lu_p1;
        hun_p1 = (mb_p1 & MB_LLU) ? avg(wn_p1 + xr(wn_p1), lu_p1 + lu_p1r, dxt_p1, dyh_p1) : 0.0;
        const double hvn_0 = (mb_0 & MB_LLV) ? avg(wn_0 + wn_p1, lu_0 + lu_p1, dxh_0, dyt_0) : 0.0;

        // tension of the filtered level on T points of row n+1 (mixing.f90:129-132)
        // and the two products uv_diff2 takes of it (vel_ssh.f90:146-148, 162-164)
        double P_p1, Px_p1;
        {
            const bool tw = (mb_p1 & MB_LU) && inx && iny1;
            const double a_p1 = up_p1 / dyh_p1;
            const double b_p1 = vp_p1 / dxh_p1;
            const double st = dy_p1 / dx_p1 * (a_p1 - xl(a_p1)) - dx_p1 / dy_p1 * (b_p1 - b_0);
            const double stt_p1 = tw ? st : 0.0;
            P_p1 = dy_p1 * dy_p1 * mu_p1 * hq_p1 * stt_p1;
            Px_p1 = dx_p1 * dx_p1 * mu_p1 * hq_p1 * stt_p1;
            b_0 = b_p1;  // consumed above as row n; row n+1 from here on
        }
        // flux-form advection factors on U and V points of row n+1 (vel_ssh.f90:63-70, 88-95)
        const double A_p1 = u_p1 * dyh_p1 * hu_p1;
        const double Bv_p1 = v_p1 * dxh_p1 * hv_p1;

        // ---- H point of row n: shear of the filtered level (mixing.f90:141-144)
        // and the metric vorticity (vel_ssh.f90:35-37)
        double sts_0, vo_0;
        {
            const bool hw = (mb_0 & MB_LUU) && inx && (n >= 3 && n <= p.ny - 2);
            const double c_p1 = up_p1 / dxt_p1;
            const double e_0 = vp_0 / dyt_0;
            const double ss = dxb_0 / dyb_0 * (c_p1 - c_0) + dyb_0 / dxb_0 * (xr(e_0) - e_0);
            const double h_p1 = u_p1 * dxt_p1;
            const double g_0 = v_0 * dyt_0;
            const double vo = (xr(g_0) - g_0) - (h_p1 - h_0) - ((xr(v_0) - v_0) * dyb_0 - (u_p1 - u_0) * dxb_0);
            sts_0 = hw ? ss : 0.0;
            vo_0 = hw ? vo : 0.0;
            c_0 = c_p1;
            h_0 = h_p1;
        }
        const double v0r = xr(v_0);
        // fy_p of the zonal equation on U(m, n) (vel_ssh.f90:69-70), every row:
        // row n's is fy_m of row n+1
        const double fyU_0 = (Bv_0 + xr(Bv_0)) / 2.0 * (u_p1 + u_0) / 2.0 * bitf<double>(mb_0, MB_LUU);
        // fy_p of the meridional equation on V(m, n) (:88-89)
        const double fyV_0 = (Bv_0 + Bv_p1) / 2.0 * (v_0 + v_p1) / 2.0;
        // vorticity and Coriolis terms on H(m, n) (:76-77, vel_ssh.f90:545-546)
        const double vh = vo_0 * hh_0;
        const double Cc_0 = vh * (v0r + v_0);
        const double fhdd = q.f0 * hh_0 * dxb_0 * dyb_0;
        const double Fc_0 = fhdd * (v0r + v_0);

        if (produce) {
            const bool iny = (n >= 3 && n <= p.ny - 2);
            const double hupc = FIRST ? hup_0 : q.hupc0;
            const double hvpc = FIRST ? hvp_0 : q.hvpc0;
            const double muh_p = (mu_0 + xr(mu_0) + mu_p1 + xr(mu_p1)) / 4.0;

            // zonal momentum (vel_ssh.f90:534-551) with uv_trans (:59-81) and uv_diff2 (:146-159)
            const double fx_p = (A_0 + xr(A_0)) / 2.0 * (u_0 + xr(u_0)) / 2.0;
            const double fx_m = xl(fx_p);
            const double radv = -(fx_p - fx_m + fyU_0 - fyU_m1) + (Cc_0 + Cc_m1) / 4.0;
            const double muh_m = (mu_0 + xr(mu_0) + mu_m1 + xr(mu_m1)) / 4.0;
            double rdfx = (xr(P_0) - P_0) / dyh_0 +
                          (dxb_0 * dxb_0 * muh_p * hh_0 * sts_0 - dxb_m1 * dxb_m1 * muh_m * hh_m1 * sts_m1) / dxt_0;
            // uv_diff4 adds to uv_diff2's result (vel_ssh.f90:292-295)
            if constexpr (D4) rdfx = rdfx + q.d4[0] + q.d4[1];
            const double bp = hun_0 * dxt_0 * dyh_0 / 2.0 / tau;
            const double bp0 = hupc * dxt_0 * dyh_0 / 2.0 / tau;
            const double slx = -G * (xr(s_0) - s_0) * dyh_0 * hu_0;
            const double slxn = -G * (xr(sn_0) - sn_0) * dyh_0 * hun_0;
            const double grx = q.rx0 + slx + rdfx + radv - (rd_0 + xr(rd_0)) / 2.0 * up_0 * dxt_0 * dyh_0 * hu_0 +
                               (Fc_0 + Fc_m1) / 4.0;
            double un = (up_0 * bp0 + grx) / bp;

            // meridional momentum (vel_ssh.f90:554-569) with uv_trans (:84-105) and uv_diff2 (:162-175)
            const double fxV_p = (A_0 + A_p1) / 2.0 * (v0r + v_0) / 2.0;
            const double fxV_m = xl(fxV_p);
            const double D = vh * (u_p1 + u_0);
            const double rady = -(fxV_p - fxV_m + fyV_0 - fyV_m1) - (D + xl(D)) / 4.0;
            const double muhv_m = (mu_0 + xl(mu_0) + mu_p1 + xl(mu_p1)) / 4.0;
            const double dybl = xl(dyb_0), hhl = xl(hh_0), stsl = xl(sts_0);
            double rdfy = -(Px_p1 - Px_0) / dxh_0 +
                          (dyb_0 * dyb_0 * muh_p * hh_0 * sts_0 - dybl * dybl * muhv_m * hhl * stsl) / dyt_0;
            // (vel_ssh.f90:306-309)
            if constexpr (D4) rdfy = rdfy - q.d4[2] + q.d4[3];
            const double bq = hvn_0 * dyt_0 * dxh_0 / 2.0 / tau;
            const double bq0 = hvpc * dyt_0 * dxh_0 / 2.0 / tau;
            const double sly = -G * (s_p1 - s_0) * dxh_0 * hv_0;
            const double slyn = -G * (sn_p1 - sn_0) * dxh_0 * hvn_0;
            const double Gc = fhdd * (u_p1 + u_0);
            const double gry = q.ry0 + sly + rdfy + rady - (rd_0 + rd_p1) / 2.0 * vp_0 * dxh_0 * dyt_0 * hv_0 -
                               (Gc + xl(Gc)) / 4.0;
            double vn = (vp_0 * bq0 + gry) / bq;

            const bool uw = (mb_0 & MB_LCU) && inx && iny;
            const bool vw = (mb_0 & MB_LCV) && inx && iny;
            un = uw ? un : 0.0;
            vn = vw ? vn : 0.0;

            // averages over the call's substeps (vel_ssh.f90:549-550, 567-568)
            const double a0 = FIRST ? 0.0 : q.a0, a1 = FIRST ? 0.0 : q.a1;
            const double a2 = FIRST ? 0.0 : q.a2, a3 = FIRST ? 0.0 : q.a3;
            o.a0 = uw ? a0 + (u_0 * hu_0 + un * hun_0) / p.d4n : a0;
            o.a2 = uw ? a2 + (slx + slxn) / p.d4n : a2;
            o.a1 = vw ? a1 + (v_0 * hv_0 + vn * hvn_0) / p.d4n : a1;
            o.a3 = vw ? a3 + (sly + slyn) / p.d4n : a3;

            // time-level shift with the Robert-Asselin filter (vel_ssh.f90:588-612)
            // and hh_shift of the U/V depths (depth.f90:195-203)
            double s_new = s_0, u_new = u_0, v_new = v_0;
            double sp_new = sp_0, up_new = up_0, vp_new = vp_0;
            double hu_new = hupc, hv_new = hvpc;
            if (inx1 && n >= 2 && n <= p.ny - 1) {
                const double ts = p.ts;
                if (mb_0 & MB_LU) {
                    sp_new = s_0 + ts * (sn_0 - 2.0 * s_0 + sp_0) / 2.0 / p.dn;
                    s_new = sn_0;
                }
                if (mb_0 & MB_LCU) {
                    up_new = u_0 + ts * (un - 2.0 * u_0 + up_0) / 2.0 / p.dn;
                    u_new = un;
                }
                if (mb_0 & MB_LCV) {
                    vp_new = v_0 + ts * (vn - 2.0 * v_0 + vp_0) / 2.0 / p.dn;
                    v_new = vn;
                }
                if (mb_0 & MB_LLU) hu_new = hu_0 + ts * (hun_0 - 2.0 * hu_0 + hupc) / 2.0 / p.dn;
                if (mb_0 & MB_LLV) hv_new = hv_0 + ts * (hvn_0 - 2.0 * hv_0 + hvpc) / 2.0 / p.dn;
            }
            o.s = s_new;
            o.u = u_new;
            o.v = v_new;
            o.sp = sp_new;
            o.up = up_new;
            o.vp = vp_new;
            o.hu = hu_new;
            o.hv = hv_new;
        }

        // ---- advance the window one row
        u_0 = u_p1;
        v_0 = v_p1;
        s_0 = s_p1;
        s_p1 = s_p2;
        H_p1 = H_p2;
        up_0 = up_p1;
        vp_0 = vp_p1;
        sp_0 = sp_p1;
        rd_0 = rd_p1;
        mu_m1 = mu_0;
        mu_0 = mu_p1;
        mb_m1 = mb_0;
        mb_0 = mb_p1;
        mb_p1 = mb_p2;
        dx_p1 = dx_p2;
        dy_p1 = dy_p2;
        dxt_0 = dxt_p1;
        dyh_0 = dyh_p1;
        dxh_0 = dxh_p1;
        dyt_0 = dyt_p1;
        dxb_m1 = dxb_0;
        hq_0 = hq_p1;
        hq_p1 = hq_p2;
        w_0 = w_p1;
        w_p1 = w_p2;
        wp_0 = wp_p1;
        wn_0 = wn_p1;
        hu_0 = hu_p1;
        hup_0 = hup_p1;
        hv_0 = hv_p1;
        hun_0 = hun_p1;
        hh_m1 = hh_0;
        sn_0 = sn_p1;
        P_0 = P_p1;
        Px_0 = Px_p1;
        sts_m1 = sts_0;
        Cc_m1 = Cc_0;
        Fc_m1 = Fc_0;
        A_0 = A_p1;
        Bv_0 = Bv_p1;
        fv_0 = fv_p1;
        fyU_m1 = fyU_0;
        fyV_m1 = fyV_0;
    }
};

// Loads of one iteration, issued one iteration ahead of their use.
template <bool FIRST, bool ROWMET, bool D4>
struct DLoader {
    const DropinParams &p;
    const RowAddr ra;
    const bool colok;
    struct Pre {
        double s, H, dx, dy, u, v, up, vp, sp, rd, mu, wf, dxt, dyh, dxh, dyt, f, rx, ry, dxb, dyb, hupc, hvpc;
        double a[4];
        double d4[D4 ? 4 : 1];
        int mb;
    } q;

    __device__ DLoader(const DropinParams &pp, int m)
        : p(pp), ra{pp.nx, pp.ny, min(max(m, 1), pp.nx)}, colok(m >= 1 && m <= pp.nx)
    {
    }
    __device__ __forceinline__ bool rowok(int row) const { return colok && row >= 1 && row <= p.ny; }
    // metric k (0..7: dx dy dxt dyt dxh dyh dxb dyb; 8: rlh_s) at (this column, row)
    __device__ __forceinline__ double met(int k, int row, size_t i) const
    {
        if constexpr (ROWMET)
            return p.met[k][min(max(row, 1), p.ny) - 1];
        else
            return p.met[k][i];
    }
    __device__ __forceinline__ void issue(int n)
    {
        const size_t i2 = ra.at(n + 2), i1 = ra.at(n + 1), i0 = ra.at(n);
        q.s = p.s[i2];
        q.H = p.H[i2];
        q.mb = p.mb[i2];
        q.dx = met(0, n + 2, i2);
        q.dy = met(1, n + 2, i2);
        q.u = p.u[i1];
        q.v = p.v[i1];
        q.up = p.up[i1];
        q.vp = p.vp[i1];
        q.sp = p.sp[i1];
        q.rd = p.rd[i1];
        q.mu = p.mu[i1];
        q.wf = p.wf[i1];
        q.dxt = met(2, n + 1, i1);
        q.dyt = met(3, n + 1, i1);
        q.dxh = met(4, n + 1, i1);
        q.dyh = met(5, n + 1, i1);
        q.dxb = met(6, n, i0);
        q.dyb = met(7, n, i0);
        q.f = met(8, n, i0);
        q.rx = p.rx[i0];
        q.ry = p.ry[i0];
        if constexpr (D4)
            for (int k = 0; k < 4; ++k) q.d4[k] = p.d4[k][i0];
        if constexpr (!FIRST) {
            q.hupc = p.hupc[i0];
            q.hvpc = p.hvpc[i0];
            for (int k = 0; k < 4; ++k) q.a[k] = p.acc[k][i0];
        }
    }
    __device__ __forceinline__ void take(int n, DIn &in) const
    {
        const bool r2 = rowok(n + 2), r1 = rowok(n + 1), r0 = rowok(n);
        in.s2 = r2 ? q.s : 0.0;
        in.H2 = r2 ? q.H : 0.0;
        in.mb2 = r2 ? q.mb : 0;
        in.dx2 = r2 ? q.dx : 0.0;
        in.dy2 = r2 ? q.dy : 0.0;
        in.u1 = r1 ? q.u : 0.0;
        in.v1 = r1 ? q.v : 0.0;
        in.up1 = r1 ? q.up : 0.0;
        in.vp1 = r1 ? q.vp : 0.0;
        in.sp1 = r1 ? q.sp : 0.0;
        in.rd1 = r1 ? q.rd : 0.0;
        in.mu1 = r1 ? q.mu : 0.0;
        in.wf1 = r1 ? q.wf : 0.0;
        in.dxt1 = r1 ? q.dxt : 0.0;
        in.dyh1 = r1 ? q.dyh : 0.0;
        in.dxh1 = r1 ? q.dxh : 0.0;
        in.dyt1 = r1 ? q.dyt : 0.0;
        in.f0 = r0 ? q.f : 0.0;
        in.rx0 = r0 ? q.rx : 0.0;
        in.ry0 = r0 ? q.ry : 0.0;
        in.dxb0 = r0 ? q.dxb : 0.0;
        in.dyb0 = r0 ? q.dyb : 0.0;
        if constexpr (D4)
            for (int k = 0; k < 4; ++k) in.d4[k] = q.d4[k];
        if constexpr (!FIRST) {
            in.hupc0 = q.hupc;
            in.hvpc0 = q.hvpc;
            in.a0 = q.a[0];
            in.a1 = q.a[1];
            in.a2 = q.a[2];
            in.a3 = q.a[3];
        } else {
            in.hupc0 = in.hvpc0 = in.a0 = in.a1 = in.a2 = in.a3 = 0.0;
        }
    }
};

// Two waves per SIMD (the later substeps would take 256 VGPRs + 4 AGPRs,
// one wave; bounded they spill 12 bytes per lane): at 2048 x 1536 on
// MI355X a call of 4 launches takes 0.782 ms instead of 0.867.  With the
// biharmonic terms (D4) one wave, whose registers hold what two waves
// would spill.
template <bool FIRST, bool ROWMET, bool D4>
__global__ void __launch_bounds__(WAVE, D4 ? 1 : 2) k_dropin(DropinParams p)
{
    const int w = blockIdx.x;
    const int t = (w % NXCD) * p.per_xcd + (w / NXCD);
    if (t >= p.total) return;
    const int chunk_id = t / p.strips, strip = t - chunk_id * p.strips;
    const int lane = threadIdx.x;
    const int m = strip * SW + lane - 1;
    const bool colok = (m >= 1 && m <= p.nx);
    const bool writer = colok && lane >= 2 && lane <= SW + 1;
    const int r0 = chunk_id * p.chunk + 1;
    const int r1 = min(r0 + p.chunk, p.ny + 1);

    DStage<FIRST, D4> st(p, m);
    DLoader<FIRST, ROWMET, D4> ld(p, m);
    DIn in;
    DOut o;
    ld.issue(r0 - 3);
    for (int n = r0 - 3; n < r1; ++n) {
        ld.take(n, in);
        if (n + 1 < r1) ld.issue(n + 1);
        const bool produce = n >= r0;
        if (produce && p.cu && writer) {
            // the state entering this substep (the call's result when this is
            // substep nstep + 1, vel_ssh.f90:638-648), row n of the window
            const size_t io = (size_t)(n - 1) * p.nx + (m - 1);
            st_out<true, double>(p.cu, io, st.u_0);
            st_out<true, double>(p.cv, io, st.v_0);
            st_out<true, double>(p.cs, io, st.s_0);
            st_out<true, double>(p.cup, io, st.up_0);
            st_out<true, double>(p.cvp, io, st.vp_0);
            st_out<true, double>(p.csp, io, st.sp_0);
        }
        st.row(n, in, produce, o);
        if (produce && writer) {
            const size_t io = (size_t)(n - 1) * p.nx + (m - 1);
            if (p.uo) {
                st_out<true, double>(p.so, io, o.s);
                st_out<true, double>(p.uo, io, o.u);
                st_out<true, double>(p.vo, io, o.v);
                st_out<true, double>(p.spo, io, o.sp);
                st_out<true, double>(p.upo, io, o.up);
                st_out<true, double>(p.vpo, io, o.vp);
            }
            if (p.carry_out) {
                p.hupc[io] = o.hu;
                p.hvpc[io] = o.hv;
            }
            p.acc[0][io] = o.a0;
            p.acc[1][io] = o.a1;
            p.acc[2][io] = o.a2;
            p.acc[3][io] = o.a3;
        }
    }
}

// ------------------------------------------------------------------ host

struct DropinAlloc {
    double *S[2][6] = {};  // ping-pong state of the substeps: u, up, v, vp, ssh, sshp
    double *hupc = nullptr, *hvpc = nullptr;
    double *rowmet = nullptr;  // 9 x ny: the metrics and rlh_s per row (when uniform along x)
    uint8_t *mb = nullptr;
    int eligible = -1;  // grid part of the eligibility: -1 not checked, 0 no, 1 yes
    std::string why;
    bool rowmet_ok = false;
    int chunk = 0;
    long long slots = 0;
    double *b4[6] = {};  // fx, fy, px, qx, py, qy of the biharmonic pre-pass (ksw4 > 0)
};

namespace {

// doubles per plane of the substeps' buffers, rounded up to 256 bytes
size_t scratch_plane(size_t n) { return (n + 31) / 32 * 32; }

#define DHC(expr)                                                                              \
    do {                                                                                       \
        hipError_t e_ = (expr);                                                                \
        if (e_ != hipSuccess) {                                                                \
            set_error(std::string(#expr) + ": " + hipGetErrorString(e_));                      \
            return INMSOM_BT_EDEVICE;                                                          \
        }                                                                                      \
    } while (0)

__global__ void k_dropin_masks(size_t n, const float *lu, const float *luu, const float *luh, const float *lcu,
                               const float *lcv, const float *llu, const float *llv, uint8_t *mb)
{
    const size_t i = (size_t)blockIdx.x * blockDim.x + threadIdx.x;
    if (i >= n) return;
    uint8_t b = 0;
    if (lu[i] > 0.5f) b |= MB_LU;
    if (lcu[i] > 0.5f) b |= MB_LCU;
    if (lcv[i] > 0.5f) b |= MB_LCV;
    if (luu[i] > 0.5f) b |= MB_LUU;
    if (luh[i] > 0.5f) b |= MB_LUH;
    if (llu[i] > 0.5f) b |= MB_LLU;
    if (llv[i] > 0.5f) b |= MB_LLV;
    mb[i] = b;
}

// The grid part of the eligibility and the row tables, checked once per
// context (the static grid does not change).
int check_grid(inmsom_bt_ctx *c, DropinAlloc *a)
{
    if (a->eligible >= 0) return INMSOM_BT_OK;
    a->eligible = 0;
    if (c->g.ffs != 1) {
        a->why = "full_free_surface != 1";
        return INMSOM_BT_OK;
    }
    if (c->g.px || c->g.py) {
        a->why = "periodic grid";
        return INMSOM_BT_OK;
    }
    const int nx = c->nx, ny = c->ny;
    if (nx < 5 || ny < 5) {
        a->why = "grid smaller than 5 x 5";
        return INMSOM_BT_OK;
    }
    const size_t n = c->n;
    std::vector<float> lu(n);
    DHC(hipMemcpy(lu.data(), c->mask[0], n * 4, hipMemcpyDeviceToHost));
    for (int r = 1; r <= ny; ++r)
        for (int m = 1; m <= nx; ++m) {
            const bool halo = m <= 2 || m >= nx - 1 || r <= 2 || r >= ny - 1;
            if (halo && lu[(size_t)(r - 1) * nx + (m - 1)] > 0.5f) {
                a->why = "ocean on the two halo cells of a side (the loop bounds then differ from the masks)";
                return INMSOM_BT_OK;
            }
        }
    // metrics and rlh_s uniform along x on columns 2..nx-1: row tables.  With
    // land on the halo the outputs never depend on columns 1 and nx (the
    // reference's own grid set-up leaves its metrics zero there and rlh_s
    // unscaled, Service/basin_parameters.f90:250-300).
    std::vector<double> f(n), tab((size_t)9 * ny);
    bool rows = true;
    for (int k = 0; k < 9 && rows; ++k) {
        DHC(hipMemcpy(f.data(), k < 8 ? c->metric[k] : c->rlh_s, n * 8, hipMemcpyDeviceToHost));
        for (int r = 0; r < ny && rows; ++r) {
            const double *row = &f[(size_t)r * nx];
            for (int m = 2; m < nx - 1; ++m)
                if (std::memcmp(&row[m], &row[1], sizeof(double)) != 0) {
                    rows = false;
                    break;
                }
            tab[(size_t)k * ny + r] = row[1];
        }
    }
    a->rowmet_ok = rows;
    if (rows) {
        DHC(hipMalloc(&a->rowmet, tab.size() * 8));
        DHC(hipMemcpy(a->rowmet, tab.data(), tab.size() * 8, hipMemcpyHostToDevice));
    }
    DHC(hipMalloc(&a->mb, n));
    hipLaunchKernelGGL(k_dropin_masks, dim3((unsigned)((n + 255) / 256)), dim3(256), 0, c->stream, n, c->mask[0],
                       c->mask[1], c->mask[2], c->mask[3], c->mask[4], c->mask[5], c->mask[6], a->mb);
    DHC(hipGetLastError());
    // launch shape: rows per wave filling whole rounds of the device's wave slots
    int dev = 0, cus = 0, per_cu = 0;
    DHC(hipGetDevice(&dev));
    DHC(hipDeviceGetAttribute(&cus, hipDeviceAttributeMultiprocessorCount, dev));
    const void *kf = rows ? (const void *)k_dropin<false, true, false> : (const void *)k_dropin<false, false, false>;
    DHC(hipOccupancyMaxActiveBlocksPerMultiprocessor(&per_cu, kf, WAVE, 0));
    a->slots = (long long)cus * (per_cu > 0 ? per_cu : 1);
    const int strips = (nx + SW - 1) / SW;
    a->chunk = fit_chunk(ny, strips, a->slots, 3, 2);
    DHC(hipStreamSynchronize(c->stream));
    a->eligible = 1;
    return INMSOM_BT_OK;
}

// the substeps' two state sets and carried depths (14 planes of n doubles)
// in library-owned memory, allocated at the first call that needs them
int ensure_buffers(inmsom_bt_ctx *c, DropinAlloc *a)
{
    const size_t n = c->n;
    for (int b = 0; b < 2; ++b)
        for (int k = 0; k < 6; ++k)
            if (!a->S[b][k]) DHC(hipMalloc(&a->S[b][k], n * 8));
    if (!a->hvpc) DHC(hipMalloc(&a->hvpc, n * 8));
    if (!a->hupc) DHC(hipMalloc(&a->hupc, n * 8));
    return INMSOM_BT_OK;
}

// the work planes of the biharmonic pre-pass (lit_diff4_terms), zeroed once
int ensure_diff4(inmsom_bt_ctx *c, DropinAlloc *a)
{
    for (double *&q : a->b4)
        if (!q) {
            DHC(hipMalloc(&q, c->n * 8));
            DHC(hipMemsetAsync(q, 0, c->n * 8, c->stream));
        }
    return INMSOM_BT_OK;
}

template <bool FIRST>
void launch_dropin(inmsom_bt_ctx *c, const DropinParams &p, bool rowmet)
{
    dim3 grid((unsigned)(p.per_xcd * NXCD)), blk(WAVE);
    const bool d4 = p.d4[0] != nullptr;
    const void *k = rowmet ? (d4 ? (const void *)k_dropin<FIRST, true, true> : (const void *)k_dropin<FIRST, true, false>)
                           : (d4 ? (const void *)k_dropin<FIRST, false, true> : (const void *)k_dropin<FIRST, false, false>);
    void *args[] = {const_cast<DropinParams *>(&p)};
    (void)hipLaunchKernel(k, grid, blk, args, 0, c->stream);
}

}  // namespace

static DropinAlloc *dalloc(inmsom_bt_ctx *c)
{
    if (!c->da) c->da = new DropinAlloc();
    return c->da;
}

bool dropin_eligible(inmsom_bt_ctx *c, int ksw4, std::string *why)
{
    DropinAlloc *a = dalloc(c);
    if (check_grid(c, a) != INMSOM_BT_OK) {
        if (why) *why = std::string("device error while checking the grid: ") + inmsom_bt_last_error();
        return false;
    }
    if (!a->eligible) {
        if (why) *why = a->why;
        return false;
    }
    (void)ksw4;  // the biharmonic term is covered by a pre-pass per substep
    return true;
}

int dropin_setup(inmsom_bt_ctx *c) { return check_grid(c, dalloc(c)); }

size_t dropin_scratch_doubles(const inmsom_bt_ctx *c) { return 14 * scratch_plane(c->n); }

int dropin_call(inmsom_bt_ctx *c, double tau, int nstep, double *const st[6], double *const acc[4],
                const double *rhsx, const double *rhsy, const double *wflux, const double *mu, const double *rdis,
                double *scratch, int ksw4, const double *mu4)
{
    DropinAlloc *a = dalloc(c);
    int rc = check_grid(c, a);
    if (rc) return rc;
    if (!a->eligible) {
        set_error("the fused barotropic_dynamics does not cover this grid: " + a->why);
        return INMSOM_BT_ESTATE;
    }
    // the substeps' state sets and carried depths: 14 planes
    double *S[2][6], *hupc, *hvpc;
    if (scratch) {
        const size_t plane = scratch_plane(c->n);
        for (int b = 0; b < 2; ++b)
            for (int k = 0; k < 6; ++k) S[b][k] = scratch + (size_t)(6 * b + k) * plane;
        hupc = scratch + 12 * plane;
        hvpc = scratch + 13 * plane;
    } else {
        if ((rc = ensure_buffers(c, a))) return rc;
        for (int b = 0; b < 2; ++b)
            for (int k = 0; k < 6; ++k) S[b][k] = a->S[b][k];
        hupc = a->hupc;
        hvpc = a->hvpc;
    }
    DropinParams p{};
    p.nx = c->nx;
    p.ny = c->ny;
    p.strips = (c->nx + SW - 1) / SW;
    p.chunk = a->chunk;
    p.nchunks = (c->ny + a->chunk - 1) / a->chunk;
    p.total = p.strips * p.nchunks;
    p.per_xcd = (p.total + NXCD - 1) / NXCD;
    p.mb = a->mb;
    p.H = c->hr;
    p.rd = rdis;
    p.mu = mu;
    p.wf = wflux;
    p.rx = rhsx;
    p.ry = rhsy;
    for (int k = 0; k < 9; ++k)
        p.met[k] = a->rowmet_ok ? a->rowmet + (size_t)k * c->ny : (k < 8 ? c->metric[k] : c->rlh_s);
    p.hupc = hupc;
    p.hvpc = hvpc;
    for (int k = 0; k < 4; ++k) p.acc[k] = acc[k];
    p.g = c->g.g;
    p.rho0 = c->g.rho0;
    p.ts = c->g.ts;
    p.tau = tau / (double)nstep;  // tau_inner, vel_ssh.f90:465
    p.ffs = (double)c->g.ffs;
    p.dn = (double)nstep;
    p.d4n = (double)(4 * nstep);
    if (ksw4 > 0) {
        if (!mu4) {
            set_error("mu4 is required when ksw4 > 0");
            return INMSOM_BT_EINVAL;
        }
        if ((rc = ensure_diff4(c, a))) return rc;
    }
    const int last = 2 * nstep;
    for (int k = 1; k <= last; ++k) {
        // substep k reads the caller's state (k = 1) or the buffer substep
        // k - 1 wrote, and writes S[k % 2] (nothing after the last substep)
        const double *const *in = k == 1 ? (const double *const *)st : (const double *const *)S[(k - 1) % 2];
        p.u = in[0];
        p.up = in[1];
        p.v = in[2];
        p.vp = in[3];
        p.s = in[4];
        p.sp = in[5];
        double *const *out = S[k % 2];
        const bool wr = k < last;
        p.uo = wr ? out[0] : nullptr;
        p.upo = wr ? out[1] : nullptr;
        p.vo = wr ? out[2] : nullptr;
        p.vpo = wr ? out[3] : nullptr;
        p.so = wr ? out[4] : nullptr;
        p.spo = wr ? out[5] : nullptr;
        // substep nstep + 1 enters with the state after substep nstep: the
        // call's result (vel_ssh.f90:638-648)
        // uv_diff4 of this substep from its entering state (level-n surface,
        // filtered velocities), added by the kernel to its uv_diff2 result
        if (ksw4 > 0) {
            lit_diff4_terms(c, in[1], in[3], in[4], mu4, a->b4, a->rowmet_ok ? a->rowmet : nullptr, a->mb);
            for (int q = 0; q < 4; ++q) p.d4[q] = a->b4[2 + q];
        }
        const bool snap = k == nstep + 1;
        p.cu = snap ? st[0] : nullptr;
        p.cup = snap ? st[1] : nullptr;
        p.cv = snap ? st[2] : nullptr;
        p.cvp = snap ? st[3] : nullptr;
        p.cs = snap ? st[4] : nullptr;
        p.csp = snap ? st[5] : nullptr;
        p.carry_out = k < last;
        if (k == 1)
            launch_dropin<true>(c, p, a->rowmet_ok);
        else
            launch_dropin<false>(c, p, a->rowmet_ok);
    }
    DHC(hipGetLastError());
    return INMSOM_BT_OK;
}

// Scratch of the entry point that allocates nothing
// (inmsom_bt_barotropic_dynamics_noalloc): the substeps' two state sets and
// carried depths, 14 planes of nx*ny doubles, in a fixed array every GPU the
// library runs on holds until the process ends.  Calls on one GPU share it:
// calls that overlap on one GPU race.  Nothing in it needs to start at
// zero (every plane is written before it is read within a call).
constexpr size_t NOALLOC_SCRATCH_BYTES = (size_t)384 << 20;
static __device__ __align__(256) double scratch[NOALLOC_SCRATCH_BYTES / sizeof(double)];

namespace {
constexpr int MAX_DEVICES = 64;
std::mutex g_scratch_mu;
double *g_scratch[MAX_DEVICES] = {};
}  // namespace

int noalloc_scratch_setup(int device)
{
    if (device < 0 || device >= MAX_DEVICES) {
        set_error("device index beyond the scratch table");
        return INMSOM_BT_EINVAL;
    }
    std::lock_guard<std::mutex> lock(g_scratch_mu);
    if (g_scratch[device]) return INMSOM_BT_OK;
    void *ptr = nullptr;  // the array's address on the current device (each GPU has its own copy)
    const hipError_t e = hipGetSymbolAddress(&ptr, HIP_SYMBOL(scratch));
    if (e != hipSuccess || !ptr) {
        set_error(std::string("hipGetSymbolAddress(scratch): ") + hipGetErrorString(e));
        return INMSOM_BT_EDEVICE;
    }
    g_scratch[device] = static_cast<double *>(ptr);
    return INMSOM_BT_OK;
}

double *noalloc_scratch(int device, size_t *doubles)
{
    *doubles = NOALLOC_SCRATCH_BYTES / sizeof(double);
    if (device < 0 || device >= MAX_DEVICES) return nullptr;
    std::lock_guard<std::mutex> lock(g_scratch_mu);
    return g_scratch[device];
}

void dropin_release(inmsom_bt_ctx *c)
{
    DropinAlloc *a = c->da;
    if (!a) return;
    for (auto &b : a->S)
        for (double *q : b)
            if (q) (void)hipFree(q);
    for (void *q : {(void *)a->hupc, (void *)a->hvpc, (void *)a->rowmet, (void *)a->mb})
        if (q) (void)hipFree(q);
    for (double *q : a->b4)
        if (q) (void)hipFree(q);
    delete a;
    c->da = nullptr;
}

void dropin_shape(const inmsom_bt_ctx *c, int *rows_per_wave, bool *row_metrics)
{
    *rows_per_wave = c->da ? c->da->chunk : 0;
    *row_metrics = c->da ? c->da->rowmet_ok : false;
}

}  // namespace ibt
