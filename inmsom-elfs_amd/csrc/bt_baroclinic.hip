// baroclinic_dynamics (Function/vel_ssh.f90:794-1027): for every open U
// (V) face the implicit vertical viscosity / vertical advection column of
// nz levels with Coriolis from the other component, Rayleigh drag and the
// surface and bottom stresses, solved by the sweep of factor8
// (Service/oc_algorithms.f90:494-524); then the Robert-Asselin step of the
// filtered level and uu = un, vv = vn.
//
// k_bc_solve: one thread per interior (m, n), both columns, the sweep's
// coefficients held in registers (nz fixed at compile time, or a generic
// instance with private arrays); consecutive threads take consecutive m,
// so every level's loads are coalesced across a wave.  un, vn go to a
// buffer because the next column's Coriolis term reads the old uu, vv of
// its neighbours.  k_bc_update: the reference's update loop over the
// interior and its one-point ring, un taken at the point the periodic wrap
// copies it from.  (Applying the filter of uup, vvp inside the solve, which
// saves the update a read and a write of them, was 1.5 ms slower at
// 2048 x 1536 x 20: the old levels held for it cost the sweep its load
// batching.)  Reference expressions and operation order throughout
// (-ffp-contract=off): bit-identical to the reference.
#include <string>

#include "bt_internal.h"

namespace ibt {

namespace {

constexpr int BC_NZMAX = 128;  // levels of the generic instance

struct BcParams {
    DevGrid g;
    int nz;
    double tau;
    int tiles_x, tiles_y;  // 64 x 4 tiles of the solve over the interior
    const double *dz, *hzt;                    // nz, nz + 1 (device)
    double *uu, *uup, *vv, *vvp;               // (nz, ny, nx), updated
    const double *ww, *nu;                     // (nz + 1, ny, nx)
    const double *rhsx, *rhsy, *rhsxt, *rhsyt; // (nz, ny, nx)
    const double *pgrx, *pgry, *rdis, *txs, *tys, *txb, *tyb;
    const double *hhu, *hhup, *hhun, *hhv, *hhvp, *hhvn, *hhh;
    double *un, *vn;                           // (nz, ny, nx) work
};

#define I(mm, nn) IDX(g, (mm), (nn))

// U column at (m, n): vel_ssh.f90:845-923
template <int NZ>
__device__ void solve_u(const BcParams &p, int m, int n)
{
    const DevGrid &g = p.g;
    const int nz = NZ > 0 ? NZ : p.nz;
    const size_t pl = (size_t)g.nx * g.ny, i = I(m, n), ie = I(m + 1, n), is = I(m, n - 1), ise = I(m + 1, n - 1);
    const double area0 = g.dxt[i], area1 = g.dyh[i];
    const double bp = p.hhun[i] * area0 * area1 / p.tau / 2.0;
    const double bp0 = p.hhup[i] * area0 * area1 / p.tau / 2.0;
    const double f0 = g.rlh_s[i] * p.hhh[i] * g.dxb[i] * g.dyb[i];
    const double f1 = g.rlh_s[is] * p.hhh[is] * g.dxb[is] * g.dyb[is];
    const double rd = (p.rdis[i] + p.rdis[ie]) / 2.0;
    // ww * dx * dy is (ww * dx) * dy: the metrics are not pre-multiplied
    const double dx0 = g.dx[i], dy0 = g.dy[i], dx1 = g.dx[ie], dy1 = g.dy[ie];
    const double hun = p.hhun[i], hu = p.hhu[i];
    constexpr int NA = NZ > 0 ? NZ : BC_NZMAX;
    constexpr int UR = NZ > 0 ? NZ : 1;  // full unrolling for a fixed nz
    double x[NA], y[NA];
    double dp = 0.0, fz_p = 0.0, uk = p.uu[i];
#pragma unroll UR
    for (int k = 1; k <= NA; ++k) {
        if (k > nz) break;
        const size_t o = (size_t)(k - 1) * pl;
        const double dm = (k == 1) ? 0.0 : dp;
        const double fz_m = (k == 1) ? 0.0 : fz_p;
        const double dzk = p.dz[k - 1];
        if (k < nz) {
            const size_t o1 = o + pl;
            dp = (p.nu[o1 + i] + p.nu[o1 + ie]) / 2.0 / hun / p.hzt[k] * area0 * area1;
            const double uk1 = p.uu[o1 + i];
            fz_p = (p.ww[o1 + i] * dx0 * dy0 + p.ww[o1 + ie] * dx1 * dy1) / 2.0 * (uk + uk1) / 2.0;
            uk = uk1;
        } else {
            dp = 0.0;
            fz_p = 0.0;
        }
        const double c = (k < nz) ? -(dp / dzk) : 0.0;
        const double a = (k == 1) ? 0.0 : -(dm / dzk);
        const double b = (k == 1) ? bp + dp / dzk : (k == nz) ? bp + dm / dzk : bp + (dp + dm) / dzk;
        const double up = p.uup[o + i];
        double e = bp0 * up + p.rhsx[o + i] + p.rhsxt[o + i] + p.pgrx[i] - (fz_p - fz_m) / dzk +
                   (f0 * (p.vv[o + ie] + p.vv[o + i]) + f1 * (p.vv[o + ise] + p.vv[o + is])) / 4.0 -
                   rd * up * area0 * area1 * hu;
        if (k == 1) e = e + p.txs[i] / dzk * area0 * area1;
        if (k == nz) e = e + p.txb[i] / dzk * area0 * area1;
        // factor8 forward sweep
        if (k == 1) {
            x[0] = -(c / b);
            y[0] = e / b;
        } else {
            const double scr = 1.0 / (a * x[k - 2] + b);
            x[k - 1] = -(c * scr);
            y[k - 1] = (e - a * y[k - 2]) * scr;
        }
    }
    // backward sweep; the filtered level of this point is read by no other
    // column, so its Robert-Asselin step (vel_ssh.f90:1006-1007) is done here
    double r = y[nz - 1];
    p.un[(size_t)(nz - 1) * pl + i] = r;
#pragma unroll UR
    for (int k = NA - 1; k >= 1; --k) {
        if (k > nz - 1) continue;
        r = x[k - 1] * r + y[k - 1];
        p.un[(size_t)(k - 1) * pl + i] = r;
    }
}

// V column at (m, n): vel_ssh.f90:925-998
template <int NZ>
__device__ void solve_v(const BcParams &p, int m, int n)
{
    const DevGrid &g = p.g;
    const int nz = NZ > 0 ? NZ : p.nz;
    const size_t pl = (size_t)g.nx * g.ny, i = I(m, n), in_ = I(m, n + 1), iw = I(m - 1, n), inw = I(m - 1, n + 1);
    const double area0 = g.dxh[i], area1 = g.dyt[i];
    const double bp = p.hhvn[i] * area0 * area1 / p.tau / 2.0;
    const double bp0 = p.hhvp[i] * area0 * area1 / p.tau / 2.0;
    const double f0 = g.rlh_s[i] * p.hhh[i] * g.dxb[i] * g.dyb[i];
    const double f1 = g.rlh_s[iw] * p.hhh[iw] * g.dxb[iw] * g.dyb[iw];
    const double rd = (p.rdis[i] + p.rdis[in_]) / 2.0;
    const double dx0 = g.dx[i], dy0 = g.dy[i], dx1 = g.dx[in_], dy1 = g.dy[in_];
    const double hvn = p.hhvn[i], hv = p.hhv[i];
    constexpr int NA = NZ > 0 ? NZ : BC_NZMAX;
    constexpr int UR = NZ > 0 ? NZ : 1;  // full unrolling for a fixed nz
    double x[NA], y[NA];
    double dp = 0.0, fz_p = 0.0, vk = p.vv[i];
#pragma unroll UR
    for (int k = 1; k <= NA; ++k) {
        if (k > nz) break;
        const size_t o = (size_t)(k - 1) * pl;
        const double dm = (k == 1) ? 0.0 : dp;
        const double fz_m = (k == 1) ? 0.0 : fz_p;
        const double dzk = p.dz[k - 1];
        if (k < nz) {
            const size_t o1 = o + pl;
            dp = (p.nu[o1 + i] + p.nu[o1 + in_]) / 2.0 / hvn / p.hzt[k] * area0 * area1;
            const double vk1 = p.vv[o1 + i];
            fz_p = (p.ww[o1 + i] * dx0 * dy0 + p.ww[o1 + in_] * dx1 * dy1) / 2.0 * (vk + vk1) / 2.0;
            vk = vk1;
        } else {
            dp = 0.0;
            fz_p = 0.0;
        }
        const double c = (k < nz) ? -(dp / dzk) : 0.0;
        const double a = (k == 1) ? 0.0 : -(dm / dzk);
        const double b = (k == 1) ? bp + dp / dzk : (k == nz) ? bp + dm / dzk : bp + (dp + dm) / dzk;
        const double vp = p.vvp[o + i];
        double e = bp0 * vp + p.rhsy[o + i] + p.rhsyt[o + i] + p.pgry[i] - (fz_p - fz_m) / dzk -
                   (f0 * (p.uu[o + in_] + p.uu[o + i]) + f1 * (p.uu[o + inw] + p.uu[o + iw])) / 4.0 -
                   rd * vp * area0 * area1 * hv;
        if (k == 1) e = e + p.tys[i] / dzk * area0 * area1;
        if (k == nz) e = e + p.tyb[i] / dzk * area0 * area1;
        if (k == 1) {
            x[0] = -(c / b);
            y[0] = e / b;
        } else {
            const double scr = 1.0 / (a * x[k - 2] + b);
            x[k - 1] = -(c * scr);
            y[k - 1] = (e - a * y[k - 2]) * scr;
        }
    }
    double r = y[nz - 1];
    p.vn[(size_t)(nz - 1) * pl + i] = r;
#pragma unroll UR
    for (int k = NA - 1; k >= 1; --k) {
        if (k > nz - 1) continue;
        r = x[k - 1] * r + y[k - 1];
        p.vn[(size_t)(k - 1) * pl + i] = r;
    }
}

template <int NZ>
__global__ __launch_bounds__(256) void k_bc_solve(BcParams p)
{
    const DevGrid &g = p.g;
    int bx, by;
    xcd_tile(p.tiles_x, p.tiles_y, (int)blockIdx.x, bx, by);
    const int m = bx * 64 + (int)(threadIdx.x & 63) + 3;
    const int n = by * 4 + (int)(threadIdx.x >> 6) + 3;
    if (m > g.nx - 2 || n > g.ny - 2) return;
    const size_t i = I(m, n);
    if (g.lcu[i] > 0.5f) solve_u<NZ>(p, m, n);
    if (g.lcv[i] > 0.5f) solve_v<NZ>(p, m, n);
}

// vel_ssh.f90:1000-1023 over m = 2..nx-1, n = 2..ny-1; un, vn outside the
// interior are zero, or on periodic boundaries the value the wrap of
// Service/oc_algorithms.f90:50-95, :118-135 copies in
__global__ __launch_bounds__(256) void k_bc_update(BcParams p, double ts, int tiles_x, int tiles_y)
{
    const DevGrid &g = p.g;
    int bx, by;
    xcd_tile(tiles_x, tiles_y, (int)blockIdx.x, bx, by);
    const int m = bx * 64 + (int)(threadIdx.x & 63) + 2;
    const int n = by * 4 + (int)(threadIdx.x >> 6) + 2;
    if (m > g.nx - 1 || n > g.ny - 1) return;
    const size_t pl = (size_t)g.nx * g.ny, i = I(m, n);
    const int mm = (g.px && m == 2) ? g.nx - 2 : (g.px && m == g.nx - 1) ? 3 : m;
    const int nn = (g.py && n == 2) ? g.ny - 2 : (g.py && n == g.ny - 1) ? 3 : n;
    const bool src_in = mm >= 3 && mm <= g.nx - 2 && nn >= 3 && nn <= g.ny - 2;
    const size_t is = I(mm, nn);
    if (g.lcu[i] > 0.5f) {
        const bool ok = src_in && g.lcu[is] > 0.5f;
        for (int k = 0; k < p.nz; ++k) {
            const size_t o = (size_t)k * pl;
            const double un = ok ? p.un[o + is] : 0.0;
            const double u = p.uu[o + i];
            p.uup[o + i] = u + ts * (un - 2.0 * u + p.uup[o + i]) / 2.0;
            p.uu[o + i] = un;
        }
    }
    if (g.lcv[i] > 0.5f) {
        const bool ok = src_in && g.lcv[is] > 0.5f;
        for (int k = 0; k < p.nz; ++k) {
            const size_t o = (size_t)k * pl;
            const double vn = ok ? p.vn[o + is] : 0.0;
            const double v = p.vv[o + i];
            p.vvp[o + i] = v + ts * (vn - 2.0 * v + p.vvp[o + i]) / 2.0;
            p.vv[o + i] = vn;
        }
    }
}

#undef I

int bfail(const std::string &m, int code)
{
    set_error(m);
    return code;
}

#define CHB(expr)                                                                                   \
    do {                                                                                            \
        hipError_t e_ = (expr);                                                                     \
        if (e_ != hipSuccess) return bfail(std::string(#expr) + ": " + hipGetErrorString(e_),        \
                                           INMSOM_BT_EDEVICE);                                      \
    } while (0)
#define TRYB(expr)                         \
    do {                                   \
        int r_ = (expr);                   \
        if (r_ != INMSOM_BT_OK) return r_; \
    } while (0)

// staging: (nz)-level fields, (nz+1)-level fields, 2-D planes, vertical grid
enum { S_UU, S_UUP, S_VV, S_VVP, S_RX, S_RY, S_RXT, S_RYT, S_UN, S_VN, S_N3 };
enum { S_WW, S_NU, S_NW };
enum { S_PGX, S_PGY, S_RDIS, S_TXS, S_TYS, S_TXB, S_TYB, S_HU, S_HUP, S_HUN, S_HV, S_HVP, S_HVN, S_HH, S_N2 };

size_t stage_doubles(const inmsom_bt_ctx *c, int nz)
{
    return ((size_t)S_N3 * nz + (size_t)S_NW * (nz + 1) + S_N2) * c->n + 2 * (size_t)(nz + 1);
}

int ensure_stage(inmsom_bt_ctx *c, int nz)
{
    const size_t need = stage_doubles(c, nz);
    if (c->bc && c->bc_n >= need) return INMSOM_BT_OK;
    if (c->bc) CHB(hipFree(c->bc));
    c->bc = nullptr;
    c->bc_n = 0;
    c->bc_nz = 0;
    CHB(hipMalloc(&c->bc, need * sizeof(double)));
    c->bc_n = need;
    return INMSOM_BT_OK;
}

double *s3(inmsom_bt_ctx *c, int nz, int k) { return c->bc + (size_t)k * nz * c->n; }
double *sw(inmsom_bt_ctx *c, int nz, int k) { return c->bc + ((size_t)S_N3 * nz + (size_t)k * (nz + 1)) * c->n; }
double *s2(inmsom_bt_ctx *c, int nz, int k)
{
    return c->bc + ((size_t)S_N3 * nz + (size_t)S_NW * (nz + 1) + k) * c->n;
}
double *svert(inmsom_bt_ctx *c, int nz) { return s2(c, nz, S_N2); }  // dz (nz + 1 slots), hzt (nz + 1)

BcParams staged(inmsom_bt_ctx *c, int nz, double tau)
{
    BcParams p{};
    p.g = c->g;
    p.nz = nz;
    p.tau = tau;
    p.dz = svert(c, nz);
    p.hzt = svert(c, nz) + nz + 1;
    p.uu = s3(c, nz, S_UU);
    p.uup = s3(c, nz, S_UUP);
    p.vv = s3(c, nz, S_VV);
    p.vvp = s3(c, nz, S_VVP);
    p.rhsx = s3(c, nz, S_RX);
    p.rhsy = s3(c, nz, S_RY);
    p.rhsxt = s3(c, nz, S_RXT);
    p.rhsyt = s3(c, nz, S_RYT);
    p.un = s3(c, nz, S_UN);
    p.vn = s3(c, nz, S_VN);
    p.ww = sw(c, nz, S_WW);
    p.nu = sw(c, nz, S_NU);
    const double **two[S_N2] = {&p.pgrx, &p.pgry, &p.rdis, &p.txs, &p.tys, &p.txb, &p.tyb,
                                &p.hhu,  &p.hhup, &p.hhun, &p.hhv, &p.hhvp, &p.hhvn, &p.hhh};
    for (int k = 0; k < S_N2; ++k) *two[k] = s2(c, nz, k);
    return p;
}

void launch(inmsom_bt_ctx *c, const BcParams &p)
{
    BcParams q = p;
    q.tiles_x = (c->nx - 4 + 63) / 64;
    q.tiles_y = (c->ny - 4 + 3) / 4;
    const dim3 gs((unsigned)(q.tiles_x * q.tiles_y));
    if (q.nz == 20)
        hipLaunchKernelGGL(k_bc_solve<20>, gs, dim3(256), 0, c->stream, q);
    else
        hipLaunchKernelGGL(k_bc_solve<0>, gs, dim3(256), 0, c->stream, q);
    // the update's 64 x 4 tiles cover m = 2..nx-1, n = 2..ny-1
    const int ux = (c->nx - 2 + 63) / 64, uy = (c->ny - 2 + 3) / 4;
    hipLaunchKernelGGL(k_bc_update, dim3((unsigned)(ux * uy)), dim3(256), 0, c->stream, q, c->g.ts, ux, uy);
}

int check(inmsom_bt_ctx *c, int nz, double tau)
{
    if (!c) return bfail("null pointer", INMSOM_BT_EINVAL);
    if (nz < 2 || nz > BC_NZMAX) return bfail("nz must be in 2..128", INMSOM_BT_EINVAL);
    if (!(tau > 0.0)) return bfail("tau must be positive", INMSOM_BT_EINVAL);
    if (c->nx < 5 || c->ny < 5) return bfail("baroclinic_dynamics needs nx, ny >= 5", INMSOM_BT_EINVAL);
    return INMSOM_BT_OK;
}

}  // namespace

}  // namespace ibt

using namespace ibt;

extern "C" {

int inmsom_bt_baroclinic_dynamics(inmsom_bt_ctx *c, double tau, int nz, double *uu, double *uup, double *vv,
                                  double *vvp, const double *ww, const double *nu, const double *rhsx,
                                  const double *rhsy, const double *rhsx_tran, const double *rhsy_tran,
                                  const double *pgrx, const double *pgry, const double *rdis, const double *tx_surf,
                                  const double *ty_surf, const double *tx_bot, const double *ty_bot,
                                  const double *dz, const double *hzt, const double *hhu, const double *hhup,
                                  const double *hhun, const double *hhv, const double *hhvp, const double *hhvn,
                                  const double *hhh)
{
    TRYB(check(c, nz, tau));
    const double *in3[8] = {uu, uup, vv, vvp, rhsx, rhsy, rhsx_tran, rhsy_tran};
    const double *in2[S_N2] = {pgrx, pgry, rdis, tx_surf, ty_surf, tx_bot, ty_bot, hhu, hhup, hhun, hhv, hhvp,
                               hhvn, hhh};
    for (const double *q : in3)
        if (!q) return bfail("null pointer", INMSOM_BT_EINVAL);
    for (const double *q : in2)
        if (!q) return bfail("null pointer", INMSOM_BT_EINVAL);
    if (!ww || !nu || !dz || !hzt) return bfail("null pointer", INMSOM_BT_EINVAL);
    CHB(hipSetDevice(c->device));
    TRYB(ensure_stage(c, nz));
    const size_t pl = c->n;
    auto put = [&](double *dst, const double *src, size_t count) {
        return hipMemcpyAsync(dst, src, count * sizeof(double), hipMemcpyHostToDevice, c->stream);
    };
    for (int k = 0; k < 8; ++k) CHB(put(s3(c, nz, k), in3[k], (size_t)nz * pl));
    CHB(put(sw(c, nz, S_WW), ww, (size_t)(nz + 1) * pl));
    CHB(put(sw(c, nz, S_NU), nu, (size_t)(nz + 1) * pl));
    for (int k = 0; k < S_N2; ++k) CHB(put(s2(c, nz, k), in2[k], pl));
    CHB(put(svert(c, nz), dz, nz));
    CHB(put(svert(c, nz) + nz + 1, hzt, nz + 1));
    const BcParams p = staged(c, nz, tau);
    launch(c, p);
    CHB(hipGetLastError());
    double *out[4] = {uu, uup, vv, vvp};
    for (int k = 0; k < 4; ++k)
        CHB(hipMemcpyAsync(out[k], s3(c, nz, k), (size_t)nz * pl * sizeof(double), hipMemcpyDeviceToHost,
                           c->stream));
    CHB(hipStreamSynchronize(c->stream));
    c->bc_nz = nz;
    return INMSOM_BT_OK;
}

int inmsom_bt_baroclinic_dynamics_dev(inmsom_bt_ctx *c, double tau, int nz, double *uu, double *uup, double *vv,
                                      double *vvp, const double *ww, const double *nu, const double *rhsx,
                                      const double *rhsy, const double *rhsx_tran, const double *rhsy_tran,
                                      const double *pgrx, const double *pgry, const double *rdis,
                                      const double *tx_surf, const double *ty_surf, const double *tx_bot,
                                      const double *ty_bot, const double *dz, const double *hzt, const double *hhu,
                                      const double *hhup, const double *hhun, const double *hhv,
                                      const double *hhvp, const double *hhvn, const double *hhh)
{
    TRYB(check(c, nz, tau));
    const void *all[] = {uu, uup, vv, vvp, ww, nu, rhsx, rhsy, rhsx_tran, rhsy_tran, pgrx, pgry, rdis, tx_surf,
                         ty_surf, tx_bot, ty_bot, dz, hzt, hhu, hhup, hhun, hhv, hhvp, hhvn, hhh};
    for (const void *q : all)
        if (!q) return bfail("null pointer", INMSOM_BT_EINVAL);
    CHB(hipSetDevice(c->device));
    // work buffers un, vn and the vertical grid
    const size_t need = (2 * (size_t)nz * c->n) + 2 * (size_t)(nz + 1);
    if (!c->bc_dev || c->bc_dev_n < need) {
        if (c->bc_dev) CHB(hipFree(c->bc_dev));
        c->bc_dev = nullptr;
        c->bc_dev_n = 0;
        CHB(hipMalloc(&c->bc_dev, need * sizeof(double)));
        c->bc_dev_n = need;
    }
    double *vert = c->bc_dev + 2 * (size_t)nz * c->n;
    CHB(hipMemcpyAsync(vert, dz, nz * sizeof(double), hipMemcpyHostToDevice, c->stream));
    CHB(hipMemcpyAsync(vert + nz + 1, hzt, (nz + 1) * sizeof(double), hipMemcpyHostToDevice, c->stream));
    BcParams p{};
    p.g = c->g;
    p.nz = nz;
    p.tau = tau;
    p.dz = vert;
    p.hzt = vert + nz + 1;
    p.uu = uu;
    p.uup = uup;
    p.vv = vv;
    p.vvp = vvp;
    p.ww = ww;
    p.nu = nu;
    p.rhsx = rhsx;
    p.rhsy = rhsy;
    p.rhsxt = rhsx_tran;
    p.rhsyt = rhsy_tran;
    p.pgrx = pgrx;
    p.pgry = pgry;
    p.rdis = rdis;
    p.txs = tx_surf;
    p.tys = ty_surf;
    p.txb = tx_bot;
    p.tyb = ty_bot;
    p.hhu = hhu;
    p.hhup = hhup;
    p.hhun = hhun;
    p.hhv = hhv;
    p.hhvp = hhvp;
    p.hhvn = hhvn;
    p.hhh = hhh;
    p.un = c->bc_dev;
    p.vn = c->bc_dev + (size_t)nz * c->n;
    launch(c, p);
    CHB(hipGetLastError());
    CHB(hipStreamSynchronize(c->stream));
    return INMSOM_BT_OK;
}

int inmsom_bt_baroclinic_timed(inmsom_bt_ctx *c, double tau, int nz, int nrep, double *ms)
{
    if (!c || !ms) return bfail("null pointer", INMSOM_BT_EINVAL);
    if (nrep < 1) return bfail("nrep must be >= 1", INMSOM_BT_EINVAL);
    if (!c->bc || c->bc_nz != nz)
        return bfail("no fields staged for nz levels: call inmsom_bt_baroclinic_dynamics first", INMSOM_BT_ESTATE);
    TRYB(check(c, nz, tau));
    CHB(hipSetDevice(c->device));
    const BcParams p = staged(c, nz, tau);
    hipEvent_t e0, e1;
    CHB(hipEventCreate(&e0));
    CHB(hipEventCreate(&e1));
    CHB(hipEventRecord(e0, c->stream));
    for (int r = 0; r < nrep; ++r) launch(c, p);
    CHB(hipGetLastError());
    CHB(hipEventRecord(e1, c->stream));
    CHB(hipEventSynchronize(e1));
    float t = 0.0f;
    CHB(hipEventElapsedTime(&t, e0, e1));
    CHB(hipEventDestroy(e0));
    CHB(hipEventDestroy(e1));
    *ms = t;
    return INMSOM_BT_OK;
}

}  // extern "C"
