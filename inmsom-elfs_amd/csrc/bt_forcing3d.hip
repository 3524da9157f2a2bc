// The forcing of the barotropic step from the 3-D model (Control/
// dyncall.f90:37-130 with ksw_vert = ksw_lat = ksw_lat4 = 0) in one kernel:
// stress components of the level n-1 velocities (:38-39), uv_trans and
// uv_diff2 at every level and on the 2-D velocities (:65-81), the caller's
// pressure-gradient terms (:92-98), depth_ave with dz (:100-104), the
// surface and bottom stresses (:53-59, the bottom level k = nz of the 3-D
// velocities) and the assembly of RHSx2d / RHSy2d (:107-126).
//
// The reference materialises stress_t/stress_s, r_vort and four 3-D RHS
// arrays per level and re-reads them; here a block of F3_NW waves owns a
// 64-column tile of F3_NW * F3_R rows (62 output columns and all rows but
// the first and last: a one-point halo), keeps each owned point's metrics
// and depths in registers for the whole call and marches the levels: per
// level it reads uu, vv, uup, vvp, amuv (and pgx, pgy) once, exchanges x
// neighbours by DPP lane shifts and y neighbours through the thread's own
// rows or a small LDS exchange between waves, and adds the level's terms
// times dz to depth sums held in registers.  Every value is formed with the reference's expression and
// operation order (this file is compiled with -ffp-contract=off), so the
// result is the reference's bit for bit.  On periodic boundaries the
// wrapped halo values of the stress, vorticity and bottom-friction arrays
// (Service/oc_algorithms.f90:50-95, :118-135) come from a small table a
// pre-pass fills from their source points.
#include <string>

#include "bt_fused_body.h"
#include "bt_internal.h"

namespace ibt {

namespace {

constexpr int F3_R = 1;             // rows per thread
constexpr int F3_NW = 12;           // waves per block
constexpr int F3_ROWS = F3_R * F3_NW;  // tile rows including the two halo rows
constexpr int F3_COLS = 62;         // output columns per block (lanes 1..62)

struct F3Params {
    DevGrid g;
    int nz, with_pg;
    const double *dz;                                         // nz values
    const double *uu, *vv, *uup, *vvp, *amuv, *pgx, *pgy;     // (nz, ny, nx)
    const double *uu2d, *vv2d, *uup2d, *vvp2d, *amuv2d;
    const double *hhq, *hhu, *hhv, *hhh, *taux, *tauy, *slpr;
    const double *rhsx_in, *rhsy_in;
    double *rhsx, *rhsy;                                      // may alias rhs*_in
    const uint8_t *mb;
    double cb_nl, ebottom;
    // periodic halo table: (st, ss, vort) per wrapped point and level
    // (level nz: the 2-D velocities; level nz + 1: bf in slot 0)
    double *halo;
    int nslots;
    int nbx, nby;  // tiles along x and y
};

__device__ __forceinline__ bool interior(const DevGrid &g, int m, int n)
{
    return m >= 3 && m <= g.nx - 2 && n >= 3 && n <= g.ny - 2;
}

// Values of the stress, vorticity and bottom-friction arrays at (m, n)
// read from global memory: used for the periodic halo points, whose values
// are those of the wrapped source point.  Same expressions as the owned-
// point path below and as Function/mixing.f90:115-168, vel_ssh.f90:30-42,
// fluxes.f90:231-243.
__device__ double st_at(const DevGrid &g, const uint8_t *mb, const double *u, const double *v, int m, int n)
{
    if (!interior(g, m, n)) return 0.0;
    const size_t i = IDX(g, m, n), il = IDX(g, m - 1, n), is = IDX(g, m, n - 1);
    if (!(mb[i] & MB_LU)) return 0.0;
    return g.dy[i] / g.dx[i] * (u[i] / g.dyh[i] - u[il] / g.dyh[il]) -
           g.dx[i] / g.dy[i] * (v[i] / g.dxh[i] - v[is] / g.dxh[is]);
}

__device__ double ss_at(const DevGrid &g, const uint8_t *mb, const double *u, const double *v, int m, int n)
{
    if (!interior(g, m, n)) return 0.0;
    const size_t i = IDX(g, m, n), iN = IDX(g, m, n + 1), iE = IDX(g, m + 1, n);
    if (!(mb[i] & MB_LUU)) return 0.0;
    return g.dxb[i] / g.dyb[i] * (u[iN] / g.dxt[iN] - u[i] / g.dxt[i]) +
           g.dyb[i] / g.dxb[i] * (v[iE] / g.dyt[iE] - v[i] / g.dyt[i]);
}

__device__ double vort_at(const DevGrid &g, const uint8_t *mb, const double *u, const double *v, int m, int n)
{
    if (!interior(g, m, n)) return 0.0;
    const size_t i = IDX(g, m, n), iN = IDX(g, m, n + 1), iE = IDX(g, m + 1, n);
    if (!(mb[i] & MB_LUU)) return 0.0;
    return (v[iE] * g.dyt[iE] - v[i] * g.dyt[i]) - (u[iN] * g.dxt[iN] - u[i] * g.dxt[i]) -
           ((v[iE] - v[i]) * g.dyb[i] - (u[iN] - u[i]) * g.dxb[i]);
}

__device__ double bf_at(const F3Params &p, const double *uu, const double *vv, int m, int n)
{
    const DevGrid &g = p.g;
    if (!interior(g, m, n)) return 0.0;
    const size_t i = IDX(g, m, n), il = IDX(g, m - 1, n), is = IDX(g, m, n - 1);
    if (!(p.mb[i] & MB_LU)) return 0.0;
    const double ut = (uu[i] * g.dxt[i] * g.dyh[i] * p.hhu[i] + uu[il] * g.dxt[il] * g.dyh[il] * p.hhu[il]) /
                      2.0 / g.dx[i] / g.dy[i] / p.hhq[i];
    const double vt = (vv[i] * g.dxh[i] * g.dyt[i] * p.hhv[i] + vv[is] * g.dxh[is] * g.dyt[is] * p.hhv[is]) /
                      2.0 / g.dx[i] / g.dy[i] / p.hhq[i];
    return p.cb_nl * sqrt(ut * ut + vt * vt + p.ebottom);
}

// Row of the periodic halo table of (m, n), -1 when (m, n) is not a
// wrapped point: columns 2 and nx - 1 with periodicity in x, rows 2 and
// ny - 1 with periodicity in y.
__device__ __forceinline__ int halo_slot(const DevGrid &g, int m, int n)
{
    if (g.px && (m == 2 || m == g.nx - 1)) return (m == 2 ? 0 : g.ny) + n - 1;
    if (g.py && (n == 2 || n == g.ny - 1)) return 2 * g.ny + (n == 2 ? 0 : g.nx) + m - 1;
    return -1;
}

// The periodic wrap (Service/oc_algorithms.f90:50-95 then :118-135) as a
// map from a halo point to the point whose value it receives.
__device__ __forceinline__ void wrap_src(const DevGrid &g, int m, int n, int &mm, int &nn)
{
    mm = (g.px && m == 2) ? g.nx - 2 : (g.px && m == g.nx - 1) ? 3 : m;
    nn = (g.py && n == 2) ? g.ny - 2 : (g.py && n == g.ny - 1) ? 3 : n;
}

// One thread per (wrapped point, level): the stress, vorticity (levels
// 0 .. nz) and bottom-friction (level nz + 1) values the wrap gives it.
__global__ void k_wrap_halo(F3Params p)
{
    const DevGrid &g = p.g;
    const int slot = blockIdx.x * blockDim.x + threadIdx.x, lvl = blockIdx.y;
    if (slot >= p.nslots) return;
    int m, n;
    if (slot < 2 * g.ny) {
        if (!g.px) return;
        m = slot < g.ny ? 2 : g.nx - 1;
        n = slot % g.ny + 1;
    } else {
        if (!g.py) return;
        const int r = slot - 2 * g.ny;
        n = r < g.nx ? 2 : g.ny - 1;
        m = r % g.nx + 1;
    }
    if (halo_slot(g, m, n) != slot) return;  // a corner listed under x
    int mm, nn;
    wrap_src(g, m, n, mm, nn);
    double *h = p.halo + ((size_t)lvl * p.nslots + slot) * 3;
    const size_t pl = (size_t)g.nx * g.ny;
    if (lvl == p.nz + 1) {
        const size_t bot = (size_t)(p.nz - 1) * pl;
        h[0] = bf_at(p, p.uu + bot, p.vv + bot, mm, nn);
        return;
    }
    const bool l2 = lvl == p.nz;
    const size_t off = l2 ? 0 : (size_t)lvl * pl;
    const double *U = l2 ? p.uu2d : p.uu + off, *V = l2 ? p.vv2d : p.vv + off;
    const double *UP = l2 ? p.uup2d : p.uup + off, *VP = l2 ? p.vvp2d : p.vvp + off;
    h[0] = st_at(g, p.mb, UP, VP, mm, nn);
    h[1] = ss_at(g, p.mb, UP, VP, mm, nn);
    h[2] = vort_at(g, p.mb, U, V, mm, nn);
}

// Level-independent values of one owned point held in registers; the
// others (KS_*) live in LDS for the whole call.
struct Pt {
    double dxh, dyh, dxt, dyt, hu, hv, hh, hq;
};
enum { KS_DYDX, KS_DXDY, KS_BXY, KS_BYX, KS_DY2, KS_DX2, KS_DXB, KS_DYB, KS_DXB2, KS_DYB2, KS_N };

// What a thread holds for its R owned points (rows n[0] .. n[R-1] of
// column m) for the whole call.
template <int R>
struct Own {
    int lane, w, m, t;  // t: thread index in the block
    int n[R], slot[R];
    unsigned idx[R];  // cell index m - 1 + (n - 1) nx (the array holds fewer than 2^29 cells)
    bool ok[R], out[R], in[R];
    uint8_t mbit[R];
    Pt s[R];
    double dxb2_b, hh_b;  // of the point below row 0
};

template <int R, int NW>
struct Lds {
    double e0[NW][9][64];  // before the stress step
    double e1[NW][5][64];  // before the last step of the level
    double ks[KS_N][NW * R * 64];
};

// One level's inputs at the owned points.
template <int R>
struct LevelIn {
    double u[R], v[R], up[R], vp[R], mu[R];
};

// a[i] as a scalar base plus a 32-bit byte offset (one address form for
// every load of a level)
__device__ __forceinline__ double ld(const double *a, unsigned i)
{
    return *(const double *)((const char *)a + i * 8u);
}

template <int R>
__device__ __forceinline__ void load_level(const Own<R> &o, const double *U, const double *V, const double *UP,
                                           const double *VP, const double *MU, LevelIn<R> &in)
{
    // points outside the array (o.idx = 0) feed only halo points: their
    // values are never used and need no mask
#pragma unroll
    for (int j = 0; j < R; ++j) {
        const unsigned i = o.idx[j];
        in.u[j] = ld(U, i);
        in.v[j] = ld(V, i);
        in.up[j] = ld(UP, i);
        in.vp[j] = ld(VP, i);
        in.mu[j] = ld(MU, i);
    }
}

// uv_trans and uv_diff2 of one level (Function/vel_ssh.f90:30-179 with the
// stress components of mixing.f90:115-168) at the owned points: radv/rady
// the transport terms, rdfx/rdfy the viscous terms, zero off lcu/lcv or
// outside the interior.  `in`: level-n velocities u, v, level n-1 up, vp
// and the viscosity mu; `lvl` selects the level's rows of the periodic
// halo table.
//
// The reference's backward fluxes are its forward fluxes of the
// neighbouring point with the operands of each sum swapped, and IEEE
// addition is commutative: fx_m(m, n) = fx_p(m-1, n), fy_m(m, n) =
// fy_p(m, n-1), and the second vorticity term of the U (V) face is the
// first one of the point below (to the left).  They are taken from the
// neighbour, bit for bit the values the reference forms.
//
// Two barriers; LDS e0 and e1 are rewritten by the next call only after
// every wave has passed the next call's first barrier or this call's
// second one respectively.
template <int R, int NW>
__device__ __forceinline__ void level_terms(const F3Params &p, const Own<R> &o, Lds<R, NW> &L, int lvl,
                                            const LevelIn<R> &in, double (&radv)[R], double (&rdfx)[R],
                                            double (&rady)[R], double (&rdfy)[R])
{
    const int lane = o.lane, w = o.w;
    // neighbouring waves; the first and last wave hold halo rows, whose
    // values from outside the tile are never used: they read their own
    const int wd = w > 0 ? w - 1 : 0, wu = w < NW - 1 ? w + 1 : NW - 1;
    const double *u = in.u, *v = in.v, *mu = in.mu;
    double a[R], b[R], c[R], d[R], ev[R], fu[R], A[R], B[R];
#pragma unroll
    for (int j = 0; j < R; ++j) {
        const Pt &s = o.s[j];
        a[j] = in.up[j] / s.dyh;
        b[j] = in.vp[j] / s.dxh;
        c[j] = in.up[j] / s.dxt;
        d[j] = in.vp[j] / s.dyt;
        ev[j] = v[j] * s.dyt;
        fu[j] = u[j] * s.dxt;
        A[j] = u[j] * s.dyh * s.hu;  // the transport products u dyh hu, v dxh hv
        B[j] = v[j] * s.dxh * s.hv;
    }
    // first row: what the wave above needs at n + 1; last row: what the
    // wave below needs at n - 1
    L.e0[w][0][lane] = c[0];
    L.e0[w][1][lane] = fu[0];
    L.e0[w][2][lane] = u[0];
    L.e0[w][3][lane] = A[0];
    L.e0[w][4][lane] = B[0];
    L.e0[w][5][lane] = v[0];
    L.e0[w][6][lane] = mu[0];
    L.e0[w][7][lane] = b[R - 1];
    L.e0[w][8][lane] = mu[R - 1];
    __syncthreads();
    // per point after this step: the U-face sums fx_p - fx_m + fy_p (XU),
    // fy_p (FYU), the vorticity term (T1), pp (PU), the first q product
    // (QU) and muh_m (MU_M); the V-face XV, FYV, TV (both vorticity terms),
    // qq (QV); E and ss
    double XU[R], FYU[R], T1[R], PU[R], QU[R], MUM[R], XV[R], FYV[R], TV[R], QV[R], E[R], ss[R];
#pragma unroll
    for (int j = 0; j < R; ++j) {
        const Pt &s = o.s[j];
        const int t = o.t * R + j;
        const double b_s = j > 0 ? b[j - 1] : L.e0[wd][7][lane];
        const double mu_s = j > 0 ? mu[j - 1] : L.e0[wd][8][lane];
        const double c_n = j < R - 1 ? c[j + 1] : L.e0[wu][0][lane];
        const double f_n = j < R - 1 ? fu[j + 1] : L.e0[wu][1][lane];
        const double u_n = j < R - 1 ? u[j + 1] : L.e0[wu][2][lane];
        const double A_n = j < R - 1 ? A[j + 1] : L.e0[wu][3][lane];
        const double B_n = j < R - 1 ? B[j + 1] : L.e0[wu][4][lane];
        const double v_n = j < R - 1 ? v[j + 1] : L.e0[wu][5][lane];
        const double mu_n = j < R - 1 ? mu[j + 1] : L.e0[wu][6][lane];
        const double a_w = xl(a[j]), d_e = xr(d[j]), e_e = xr(ev[j]), v_e = xr(v[j]);
        const double A_e = xr(A[j]), u_e = xr(u[j]), B_e = xr(B[j]);
        const double mu_e = xr(mu[j]), mu_ne = xr(mu_n), mu_se = xr(mu_s), mu_w = xl(mu[j]), mu_nw = xl(mu_n);
        const bool inj = o.in[j];
        const uint8_t mb = o.mbit[j];
        double st = (inj && (mb & MB_LU))
                        ? L.ks[KS_DYDX][t] * (a[j] - a_w) - L.ks[KS_DXDY][t] * (b[j] - b_s)
                        : 0.0;
        ss[j] = (inj && (mb & MB_LUU)) ? L.ks[KS_BXY][t] * (c_n - c[j]) + L.ks[KS_BYX][t] * (d_e - d[j]) : 0.0;
        double vo = (inj && (mb & MB_LUU)) ? (e_e - ev[j]) - (f_n - fu[j]) -
                                                 ((v_e - v[j]) * L.ks[KS_DYB][t] - (u_n - u[j]) * L.ks[KS_DXB][t])
                                           : 0.0;
        if (o.slot[j] >= 0) {
            const double *h = p.halo + ((size_t)lvl * p.nslots + o.slot[j]) * 3;
            st = h[0];
            ss[j] = h[1];
            vo = h[2];
        }
        const double Cv = vo * s.hh;
        const double D = L.ks[KS_DY2][t] * mu[j] * s.hq * st;
        E[j] = L.ks[KS_DX2][t] * mu[j] * s.hq * st;
        const double luu = (mb & MB_LUU) ? 1.0 : 0.0;
        const double fxU = (A[j] + A_e) / 2.0 * (u[j] + u_e) / 2.0;
        FYU[j] = (B[j] + B_e) / 2.0 * (u_n + u[j]) / 2.0 * luu;
        XU[j] = fxU - xl(fxU) + FYU[j];
        T1[j] = Cv * (v_e + v[j]);
        const double fxV = (A[j] + A_n) / 2.0 * (v_e + v[j]) / 2.0;
        FYV[j] = (B[j] + B_n) / 2.0 * (v[j] + v_n) / 2.0;
        XV[j] = fxV - xl(fxV) + FYV[j];
        const double T2 = Cv * (u_n + u[j]);
        TV[j] = T2 + xl(T2);
        const double muh_p = (mu[j] + mu_e + mu_n + mu_ne) / 4.0;
        MUM[j] = (mu[j] + mu_e + mu_s + mu_se) / 4.0;
        PU[j] = (xr(D) - D) / s.dyh;
        QU[j] = L.ks[KS_DXB2][t] * muh_p * s.hh * ss[j];
        const double muh_mv = (mu[j] + mu_w + mu_n + mu_nw) / 4.0;
        const double dyb2 = L.ks[KS_DYB2][t];
        QV[j] = (dyb2 * muh_p * s.hh * ss[j] - xl(dyb2) * muh_mv * xl(s.hh) * xl(ss[j])) / s.dyt;
    }
    L.e1[w][0][lane] = E[0];
    L.e1[w][1][lane] = FYU[R - 1];
    L.e1[w][2][lane] = FYV[R - 1];
    L.e1[w][3][lane] = T1[R - 1];
    L.e1[w][4][lane] = ss[R - 1];
    __syncthreads();
#pragma unroll
    for (int j = 0; j < R; ++j) {
        const Pt &s = o.s[j];
        const double E_n = j < R - 1 ? E[j + 1] : L.e1[wu][0][lane];
        const double fyU_s = j > 0 ? FYU[j - 1] : L.e1[wd][1][lane];
        const double fyV_s = j > 0 ? FYV[j - 1] : L.e1[wd][2][lane];
        const double T1_s = j > 0 ? T1[j - 1] : L.e1[wd][3][lane];
        const double ss_s = j > 0 ? ss[j - 1] : L.e1[wd][4][lane];
        const double dxb2_s = j > 0 ? L.ks[KS_DXB2][o.t * R + j - 1] : o.dxb2_b;
        const double hh_s = j > 0 ? o.s[j - 1].hh : o.hh_b;
        const bool cu = o.in[j] && (o.mbit[j] & MB_LCU), cv = o.in[j] && (o.mbit[j] & MB_LCV);
        const double trU = -(XU[j] - fyU_s) + (T1[j] + T1_s) / 4.0;
        const double qU = (QU[j] - dxb2_s * MUM[j] * hh_s * ss_s) / s.dxt;
        radv[j] = cu ? trU : 0.0;
        rdfx[j] = cu ? PU[j] + qU : 0.0;
        const double trV = -(XV[j] - fyV_s) - TV[j] / 4.0;
        const double pV = (E_n - E[j]) / s.dxh;
        rady[j] = cv ? trV : 0.0;
        rdfy[j] = cv ? -pV + QV[j] : 0.0;
    }
}

// Tile of a block.  Blocks are dealt to the eight XCDs in turn (block b to
// XCD b % 8); each XCD takes a contiguous run of the tiles listed band by
// band (bands TILE_BAND tiles wide, rows bottom to top within a band), so
// the blocks one XCD runs at once cover a compact patch and the halo rows
// and columns they share are read from HBM once into that XCD's L2.
constexpr int TILE_BAND = 4;
__device__ __forceinline__ void tile_of_block(int nbx, int nby, int b, int &bx, int &by)
{
    const int T = nbx * nby, x = b % NXCD, q = b / NXCD;
    const int per = T / NXCD, extra = T % NXCD;  // XCD x runs per + (x < extra) blocks
    const int pos = x * per + (x < extra ? x : extra) + q;
    const int band = pos / (TILE_BAND * nby);
    const int bw = min(TILE_BAND, nbx - band * TILE_BAND);
    const int r = pos - band * TILE_BAND * nby;
    by = r / bw;
    bx = band * TILE_BAND + r % bw;
}

template <int R, int NW>
__global__ __launch_bounds__(64 * NW) void k_forcing(F3Params p)
{
    const DevGrid &g = p.g;
    const int nx = g.nx, ny = g.ny;
    const size_t pl = (size_t)nx * ny;
    __shared__ Lds<R, NW> L;
    Own<R> o;
    o.t = threadIdx.x;
    o.lane = threadIdx.x & 63;
    o.w = threadIdx.x >> 6;
    const int lane = o.lane, w = o.w;
    int bx, by;
    tile_of_block(p.nbx, p.nby, (int)blockIdx.x, bx, by);
    // tile: columns 1 + 62 bx .. 64 + 62 bx, rows 1 + (R NW - 2) by .. ;
    // owned rows of this thread: n0 .. n0 + R - 1
    const int m = 1 + bx * F3_COLS + lane;
    const int n0 = 1 + by * (R * NW - 2) + w * R;
    o.m = m;
#pragma unroll
    for (int j = 0; j < R; ++j) {
        const int n = n0 + j, e = w * R + j;  // e: tile row
        const int t = o.t * R + j;
        o.n[j] = n;
        o.ok[j] = m <= nx && n <= ny;
        o.out[j] = o.ok[j] && lane >= 1 && lane <= 62 && e >= 1 && e <= R * NW - 2 && m >= 2 && m <= nx - 1 &&
                   n >= 2 && n <= ny - 1;
        o.in[j] = interior(g, m, n);
        o.slot[j] = o.ok[j] ? halo_slot(g, m, n) : -1;
        o.idx[j] = o.ok[j] ? (unsigned)IDX(g, m, n) : 0u;
        const unsigned i = o.idx[j];
        Pt &s = o.s[j];
        {
            const double dx = g.dx[i], dy = g.dy[i], dxb = g.dxb[i], dyb = g.dyb[i];
            L.ks[KS_DYDX][t] = dy / dx;
            L.ks[KS_DXDY][t] = dx / dy;
            L.ks[KS_BXY][t] = dxb / dyb;
            L.ks[KS_BYX][t] = dyb / dxb;
            L.ks[KS_DY2][t] = dy * dy;
            L.ks[KS_DX2][t] = dx * dx;
            L.ks[KS_DXB][t] = dxb;
            L.ks[KS_DYB][t] = dyb;
            L.ks[KS_DXB2][t] = dxb * dxb;
            L.ks[KS_DYB2][t] = dyb * dyb;
            s.dxh = g.dxh[i];
            s.dyh = g.dyh[i];
            s.dxt = g.dxt[i];
            s.dyt = g.dyt[i];
            s.hu = p.hhu[i];
            s.hv = p.hhv[i];
            s.hh = p.hhh[i];
            s.hq = p.hhq[i];
            o.mbit[j] = o.ok[j] ? p.mb[i] : 0;
        }
    }
    // dxb^2 and hh of the point below row 0 (n0 - 1)
    L.e1[w][0][lane] = L.ks[KS_DXB2][o.t * R + R - 1];
    L.e1[w][1][lane] = o.s[R - 1].hh;
    __syncthreads();
    o.dxb2_b = w > 0 ? L.e1[w - 1][0][lane] : 1.0;
    o.hh_b = w > 0 ? L.e1[w - 1][1][lane] : 1.0;

    // depth sums (depth_ave, Function/depth.f90:223-254) of RHS3d = uv_diff2
    // + pressure gradients and of the transport terms, in level order
    double ax[R], atx[R], ay[R], aty[R];
#pragma unroll
    for (int j = 0; j < R; ++j) ax[j] = atx[j] = ay[j] = aty[j] = 0.0;
    // one level: its terms at the owned points added to the depth sums
    auto level = [&](int k, const LevelIn<R> &cur) {
        const size_t off = (size_t)k * pl;
        double pgx[R], pgy[R];
        if (p.with_pg) {
#pragma unroll
            for (int j = 0; j < R; ++j) {
                pgx[j] = ld(p.pgx + off, o.idx[j]);
                pgy[j] = ld(p.pgy + off, o.idx[j]);
            }
        }
        double radv[R], rdfx[R], rady[R], rdfy[R];
        level_terms<R, NW>(p, o, L, k, cur, radv, rdfx, rady, rdfy);
        const double dzk = p.dz[k];
#pragma unroll
        for (int j = 0; j < R; ++j) {
            double rx = rdfx[j], ry = rdfy[j];
            if (p.with_pg) {
                rx = rx + pgx[j];
                ry = ry + pgy[j];
            }
            ax[j] = ax[j] + rx * dzk;
            ay[j] = ay[j] + ry * dzk;
            atx[j] = atx[j] + radv[j] * dzk;
            aty[j] = aty[j] + rady[j] * dzk;
        }
    };
    // loads of level k (k = nz: the 2-D velocities) into `in`
    auto fetch = [&](int k, LevelIn<R> &in) {
        if (k < p.nz) {
            const size_t off = (size_t)k * pl;
            load_level(o, p.uu + off, p.vv + off, p.uup + off, p.vvp + off, p.amuv + off, in);
        } else {
            load_level(o, p.uu2d, p.vv2d, p.uup2d, p.vvp2d, p.amuv2d, in);
        }
    };
    // level k + 1 is loaded while level k is formed
    LevelIn<R> nxt;
    fetch(0, nxt);
    for (int k = 0; k < p.nz; ++k) {
        const LevelIn<R> cur = nxt;
        fetch(k + 1, nxt);
        level(k, cur);
    }
    // the 2-D velocities (RHSx2d_tran_disp, RHSx2d_diff_disp)
    double txd[R], dfx[R], tyd[R], dfy[R];
    level_terms<R, NW>(p, o, L, p.nz, nxt, txd, dfx, tyd, dfy);

    // surface and bottom stresses (Function/fluxes.f90:29-37, :231-272) on
    // the bottom level, then the assembly (Control/dyncall.f90:107-126)
    const size_t bot = (size_t)(p.nz - 1) * pl;
    const double *UB = p.uu + bot, *VB = p.vv + bot;
    double G[R], H[R], wx[R], wy[R], sl[R], dxr[R], dyr[R];
#pragma unroll
    for (int j = 0; j < R; ++j) {
        const size_t i = o.idx[j];
        const double dx = g.dx[i], dy = g.dy[i];
        dxr[j] = dx;
        dyr[j] = dy;
        G[j] = UB[i] * o.s[j].dxt * o.s[j].dyh * o.s[j].hu;
        H[j] = VB[i] * o.s[j].dxh * o.s[j].dyt * o.s[j].hv;
        wx[j] = p.taux[i] * dx * dy;
        wy[j] = p.tauy[i] * dx * dy;
        sl[j] = p.slpr[i];
    }
    const int wd = w > 0 ? w - 1 : 0, wu = w < NW - 1 ? w + 1 : NW - 1;
    __syncthreads();  // every wave is done with e0/e1 of the last level
    L.e0[w][0][lane] = H[R - 1];
    L.e0[w][1][lane] = wy[0];
    L.e0[w][2][lane] = sl[0];
    __syncthreads();
    double bf[R];
#pragma unroll
    for (int j = 0; j < R; ++j) {
        const double H_s = j > 0 ? H[j - 1] : L.e0[wd][0][lane];
        const double G_w = xl(G[j]);
        const double hq = o.s[j].hq;
        const double ut = (G[j] + G_w) / 2.0 / dxr[j] / dyr[j] / hq;
        const double vt = (H[j] + H_s) / 2.0 / dxr[j] / dyr[j] / hq;
        bf[j] = (o.in[j] && (o.mbit[j] & MB_LU)) ? p.cb_nl * sqrt(ut * ut + vt * vt + p.ebottom) : 0.0;
        if (o.slot[j] >= 0) bf[j] = p.halo[((size_t)(p.nz + 1) * p.nslots + o.slot[j]) * 3];
    }
    L.e1[w][0][lane] = bf[0];
    __syncthreads();
#pragma unroll
    for (int j = 0; j < R; ++j) {
        const double bf_n = j < R - 1 ? bf[j + 1] : L.e1[wu][0][lane];
        const double wy_n = j < R - 1 ? wy[j + 1] : L.e0[wu][1][lane];
        const double sl_n = j < R - 1 ? sl[j + 1] : L.e0[wu][2][lane];
        const double bf_e = xr(bf[j]), wx_e = xr(wx[j]), sl_e = xr(sl[j]);
        if (!o.out[j]) continue;
        const Pt &s = o.s[j];
        const size_t i = o.idx[j];
        const uint8_t mbj = o.mbit[j];
        double X = (mbj & MB_LLU) ? ax[j] : p.rhsx_in[i];
        double Y = (mbj & MB_LLV) ? ay[j] : p.rhsy_in[i];
        const double tx2 = (mbj & MB_LLU) ? atx[j] : 0.0;
        const double ty2 = (mbj & MB_LLV) ? aty[j] : 0.0;
        const double rho0 = g.rho0;
        if (o.in[j] && (mbj & MB_LCU)) {
            const double sfx = (wx[j] + wx_e) / rho0 / 2.0 / s.dxt / s.dyh;
            const double bfx = -(bf[j] + bf_e) / 2.0 * p.uup[bot + i];
            X = X - dfx[j] + tx2 - txd[j] + (sfx + bfx) * s.dxt * s.dyh - (sl_e - sl[j]) * s.hu * s.dyh / rho0;
        }
        if (o.in[j] && (mbj & MB_LCV)) {
            const double sfy = (wy[j] + wy_n) / rho0 / 2.0 / s.dxh / s.dyt;
            const double bfy = -(bf[j] + bf_n) / 2.0 * p.vvp[bot + i];
            Y = Y - dfy[j] + ty2 - tyd[j] + (sfy + bfy) * s.dyt * s.dxh - (sl_n - sl[j]) * s.hv * s.dxh / rho0;
        }
        p.rhsx[i] = X;
        p.rhsy[i] = Y;
    }
}

__global__ void k_pack_forcing_masks(size_t n, const float *lu, const float *luu, const float *lcu,
                                     const float *lcv, const float *llu, const float *llv, uint8_t *mb)
{
    const size_t i = (size_t)blockIdx.x * blockDim.x + threadIdx.x;
    if (i >= n) return;
    uint8_t b = 0;
    if (lu[i] > 0.5f) b |= MB_LU;
    if (luu[i] > 0.5f) b |= MB_LUU;
    if (lcu[i] > 0.5f) b |= MB_LCU;
    if (lcv[i] > 0.5f) b |= MB_LCV;
    if (llu[i] > 0.5f) b |= MB_LLU;
    if (llv[i] > 0.5f) b |= MB_LLV;
    mb[i] = b;
}

int ffail(const std::string &m, int code)
{
    set_error(m);
    return code;
}

#define CHF(expr)                                                                                   \
    do {                                                                                            \
        hipError_t e_ = (expr);                                                                     \
        if (e_ != hipSuccess) return ffail(std::string(#expr) + ": " + hipGetErrorString(e_),        \
                                           INMSOM_BT_EDEVICE);                                      \
    } while (0)
#define TRYF(expr)                         \
    do {                                   \
        int r_ = (expr);                   \
        if (r_ != INMSOM_BT_OK) return r_; \
    } while (0)

// staging of the host-array entry: 3-D inputs (nz planes each), 2-D
// planes, dz; the mask byte per cell is kept for every call
enum { UU, VV, UUP, VVP, AMUV, PGX, PGY, N3 };
enum { UU2D, VV2D, UUP2D, VVP2D, AMUV2D, HHQ, HHU, HHV, HHH, TAUX, TAUY, SLPR, RHSX, RHSY, DZ, N2 };

int ensure_masks(inmsom_bt_ctx *c)
{
    if (c->f3_mb) return INMSOM_BT_OK;
    CHF(hipMalloc(&c->f3_mb, c->n));
    hipLaunchKernelGGL(k_pack_forcing_masks, dim3((unsigned)((c->n + 255) / 256)), dim3(256), 0, c->stream, c->n,
                       c->mask[0], c->mask[1], c->mask[3], c->mask[4], c->mask[5], c->mask[6], c->f3_mb);
    CHF(hipGetLastError());
    return INMSOM_BT_OK;
}

int ensure_staging(inmsom_bt_ctx *c, int nz)
{
    const size_t need = ((size_t)N3 * nz + N2) * c->n;
    if (c->f3 && c->f3_n >= need) return INMSOM_BT_OK;
    if (c->f3) CHF(hipFree(c->f3));
    c->f3 = nullptr;
    c->f3_n = 0;
    c->f3_nz = 0;
    CHF(hipMalloc(&c->f3, need * sizeof(double)));
    c->f3_n = need;
    return INMSOM_BT_OK;
}

int ensure_halo(inmsom_bt_ctx *c, int nz)
{
    if (!c->g.px && !c->g.py) return INMSOM_BT_OK;
    const size_t need = (size_t)(nz + 2) * 2 * (c->nx + c->ny) * 3;
    if (c->f3_halo && c->f3_halo_n >= need) return INMSOM_BT_OK;
    if (c->f3_halo) CHF(hipFree(c->f3_halo));
    c->f3_halo = nullptr;
    c->f3_halo_n = 0;
    CHF(hipMalloc(&c->f3_halo, need * sizeof(double)));
    c->f3_halo_n = need;
    return INMSOM_BT_OK;
}

inline double *p3(inmsom_bt_ctx *c, int nz, int k) { return c->f3 + (size_t)k * nz * c->n; }
inline double *p2(inmsom_bt_ctx *c, int nz, int k) { return c->f3 + ((size_t)N3 * nz + k) * c->n; }

F3Params staged_params(inmsom_bt_ctx *c, int nz, bool with_pg)
{
    F3Params p{};
    p.g = c->g;
    p.nz = nz;
    p.with_pg = with_pg;
    p.dz = p2(c, nz, DZ);
    p.uu = p3(c, nz, UU);
    p.vv = p3(c, nz, VV);
    p.uup = p3(c, nz, UUP);
    p.vvp = p3(c, nz, VVP);
    p.amuv = p3(c, nz, AMUV);
    p.pgx = p3(c, nz, PGX);
    p.pgy = p3(c, nz, PGY);
    p.uu2d = p2(c, nz, UU2D);
    p.vv2d = p2(c, nz, VV2D);
    p.uup2d = p2(c, nz, UUP2D);
    p.vvp2d = p2(c, nz, VVP2D);
    p.amuv2d = p2(c, nz, AMUV2D);
    p.hhq = p2(c, nz, HHQ);
    p.hhu = p2(c, nz, HHU);
    p.hhv = p2(c, nz, HHV);
    p.hhh = p2(c, nz, HHH);
    p.taux = p2(c, nz, TAUX);
    p.tauy = p2(c, nz, TAUY);
    p.slpr = p2(c, nz, SLPR);
    p.rhsx_in = p.rhsx = p2(c, nz, RHSX);
    p.rhsy_in = p.rhsy = p2(c, nz, RHSY);
    p.mb = c->f3_mb;
    p.cb_nl = c->cb_nl;
    p.ebottom = c->ebottom;
    p.halo = c->f3_halo;
    p.nslots = 2 * (c->nx + c->ny);
    p.nbx = (c->nx - 2 + F3_COLS - 1) / F3_COLS;
    p.nby = (c->ny - 2 + F3_ROWS - 3) / (F3_ROWS - 2);
    return p;
}

void launch_forcing(inmsom_bt_ctx *c, const F3Params &p)
{
    if (c->g.px || c->g.py)
        hipLaunchKernelGGL(k_wrap_halo, dim3((unsigned)((p.nslots + 255) / 256), (unsigned)(p.nz + 2)), dim3(256), 0,
                           c->stream, p);
    hipLaunchKernelGGL((k_forcing<F3_R, F3_NW>), dim3((unsigned)(p.nbx * p.nby)), dim3(64 * F3_NW), 0, c->stream,
                       p);
}

int check_args(inmsom_bt_ctx *c, int nz)
{
    if (nz < 1 || nz > 1000) return ffail("nz must be in 1..1000", INMSOM_BT_EINVAL);
    if (c->nx < 5 || c->ny < 5) return ffail("the forcing needs nx, ny >= 5", INMSOM_BT_EINVAL);
    if (c->n >= ((size_t)1 << 29)) return ffail("the forcing needs nx * ny < 2^29", INMSOM_BT_EINVAL);
    return INMSOM_BT_OK;
}

}  // namespace

}  // namespace ibt

using namespace ibt;

extern "C" {

int inmsom_bt_forcing_3d(inmsom_bt_ctx *c, int nz, const double *dz, const double *uu, const double *vv,
                         const double *uup, const double *vvp, const double *amuv, const double *uu2d,
                         const double *vv2d, const double *uup2d, const double *vvp2d, const double *amuv2d,
                         const double *hhq, const double *hhu, const double *hhv, const double *hhh,
                         const double *taux, const double *tauy, const double *slpr, const double *pgx,
                         const double *pgy, double *rhsx2d, double *rhsy2d)
{
    if (!c || !dz || !uu || !vv || !uup || !vvp || !amuv || !uu2d || !vv2d || !uup2d || !vvp2d || !amuv2d ||
        !hhq || !hhu || !hhv || !hhh || !taux || !tauy || !slpr || !rhsx2d || !rhsy2d)
        return ffail("null pointer", INMSOM_BT_EINVAL);
    if ((pgx == nullptr) != (pgy == nullptr)) return ffail("pgx and pgy: both or neither", INMSOM_BT_EINVAL);
    TRYF(check_args(c, nz));
    CHF(hipSetDevice(c->device));
    TRYF(ensure_masks(c));
    TRYF(ensure_staging(c, nz));
    TRYF(ensure_halo(c, nz));
    const size_t pl = c->n, n3 = (size_t)nz * pl;
    auto put = [&](double *dst, const double *src, size_t count) {
        return hipMemcpyAsync(dst, src, count * sizeof(double), hipMemcpyHostToDevice, c->stream);
    };
    const double *in3[N3] = {uu, vv, uup, vvp, amuv, pgx, pgy};
    for (int k = 0; k < N3; ++k)
        if (in3[k]) CHF(put(p3(c, nz, k), in3[k], n3));
    const double *in2[DZ] = {uu2d, vv2d, uup2d, vvp2d, amuv2d, hhq, hhu, hhv, hhh, taux, tauy, slpr, rhsx2d, rhsy2d};
    for (int k = 0; k < DZ; ++k) CHF(put(p2(c, nz, k), in2[k], pl));
    CHF(put(p2(c, nz, DZ), dz, nz));
    launch_forcing(c, staged_params(c, nz, pgx != nullptr));
    CHF(hipGetLastError());
    CHF(hipMemcpyAsync(rhsx2d, p2(c, nz, RHSX), pl * sizeof(double), hipMemcpyDeviceToHost, c->stream));
    CHF(hipMemcpyAsync(rhsy2d, p2(c, nz, RHSY), pl * sizeof(double), hipMemcpyDeviceToHost, c->stream));
    CHF(hipStreamSynchronize(c->stream));
    c->f3_nz = nz;
    c->f3_pg = pgx != nullptr;
    return INMSOM_BT_OK;
}

int inmsom_bt_forcing_3d_dev(inmsom_bt_ctx *c, int nz, const double *dz, const double *uu, const double *vv,
                             const double *uup, const double *vvp, const double *amuv, const double *uu2d,
                             const double *vv2d, const double *uup2d, const double *vvp2d, const double *amuv2d,
                             const double *hhq, const double *hhu, const double *hhv, const double *hhh,
                             const double *taux, const double *tauy, const double *slpr, const double *pgx,
                             const double *pgy, double *rhsx2d, double *rhsy2d)
{
    if (!c || !dz || !uu || !vv || !uup || !vvp || !amuv || !uu2d || !vv2d || !uup2d || !vvp2d || !amuv2d ||
        !hhq || !hhu || !hhv || !hhh || !taux || !tauy || !slpr || !rhsx2d || !rhsy2d)
        return ffail("null pointer", INMSOM_BT_EINVAL);
    if ((pgx == nullptr) != (pgy == nullptr)) return ffail("pgx and pgy: both or neither", INMSOM_BT_EINVAL);
    TRYF(check_args(c, nz));
    CHF(hipSetDevice(c->device));
    TRYF(ensure_masks(c));
    TRYF(ensure_halo(c, nz));
    if (!c->f3_dz) CHF(hipMalloc(&c->f3_dz, 1000 * sizeof(double)));
    CHF(hipMemcpyAsync(c->f3_dz, dz, nz * sizeof(double), hipMemcpyHostToDevice, c->stream));
    F3Params p{};
    p.g = c->g;
    p.nz = nz;
    p.with_pg = pgx != nullptr;
    p.dz = c->f3_dz;
    p.uu = uu;
    p.vv = vv;
    p.uup = uup;
    p.vvp = vvp;
    p.amuv = amuv;
    p.pgx = pgx;
    p.pgy = pgy;
    p.uu2d = uu2d;
    p.vv2d = vv2d;
    p.uup2d = uup2d;
    p.vvp2d = vvp2d;
    p.amuv2d = amuv2d;
    p.hhq = hhq;
    p.hhu = hhu;
    p.hhv = hhv;
    p.hhh = hhh;
    p.taux = taux;
    p.tauy = tauy;
    p.slpr = slpr;
    p.rhsx_in = p.rhsx = rhsx2d;
    p.rhsy_in = p.rhsy = rhsy2d;
    p.mb = c->f3_mb;
    p.cb_nl = c->cb_nl;
    p.ebottom = c->ebottom;
    p.halo = c->f3_halo;
    p.nslots = 2 * (c->nx + c->ny);
    p.nbx = (c->nx - 2 + F3_COLS - 1) / F3_COLS;
    p.nby = (c->ny - 2 + F3_ROWS - 3) / (F3_ROWS - 2);
    launch_forcing(c, p);
    CHF(hipGetLastError());
    CHF(hipStreamSynchronize(c->stream));
    return INMSOM_BT_OK;
}

int inmsom_bt_forcing_3d_timed(inmsom_bt_ctx *c, int nz, int nrep, double *ms)
{
    if (!c || !ms) return ffail("null pointer", INMSOM_BT_EINVAL);
    if (nrep < 1) return ffail("nrep must be >= 1", INMSOM_BT_EINVAL);
    if (!c->f3 || c->f3_nz != nz)
        return ffail("no forcing fields staged for nz levels: call inmsom_bt_forcing_3d first", INMSOM_BT_ESTATE);
    CHF(hipSetDevice(c->device));
    const F3Params p = staged_params(c, nz, c->f3_pg);
    hipEvent_t e0, e1;
    CHF(hipEventCreate(&e0));
    CHF(hipEventCreate(&e1));
    CHF(hipEventRecord(e0, c->stream));
    for (int r = 0; r < nrep; ++r) launch_forcing(c, p);
    CHF(hipGetLastError());
    CHF(hipEventRecord(e1, c->stream));
    CHF(hipEventSynchronize(e1));
    float t = 0.0f;
    CHF(hipEventElapsedTime(&t, e0, e1));
    CHF(hipEventDestroy(e0));
    CHF(hipEventDestroy(e1));
    *ms = t;
    return INMSOM_BT_OK;
}

}  // extern "C"
