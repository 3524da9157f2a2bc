// The forcing of the barotropic step from the 3-D model (Control/
// dyncall.f90:37-130 with ksw_vert = ksw_lat = ksw_lat4 = 0): stress
// components of the level n-1 3-D and 2-D velocities, the surface and
// bottom stresses (the bottom level k = nz of the 3-D velocities), uv_trans
// and uv_diff2 at every level and on the 2-D velocities, the caller's 3-D
// pressure-gradient terms, depth_ave with dz, and the assembly of
// RHSx2d / RHSy2d (:109-126).  Reference index ranges, masks and
// operation order (-ffp-contract=off), one level at a time through the
// stage-by-stage path's kernels.
#include <string>

#include "bt_internal.h"
#include "bt_literal.h"

namespace ibt {

namespace {

#define KERNEL_MN                                            \
    const int m = blockIdx.x * blockDim.x + threadIdx.x + 1; \
    const int n = blockIdx.y * blockDim.y + threadIdx.y + 1;
#define I(mm, nn) IDX(g, (mm), (nn))
#define IN_INTERIOR(g) (m >= 3 && m <= (g).nx - 2 && n >= 3 && n <= (g).ny - 2)

__global__ void k_fill3(size_t count, double *__restrict__ a, double v)
{
    const size_t i = (size_t)blockIdx.x * blockDim.x + threadIdx.x;
    if (i < count) a[i] = v;
}

// pressure_gradients adds its terms to the 3-D RHS (dyncall.f90:92-98)
__global__ void k_add3(size_t count, double *__restrict__ a, const double *__restrict__ b)
{
    const size_t i = (size_t)blockIdx.x * blockDim.x + threadIdx.x;
    if (i < count) a[i] = a[i] + b[i];
}

// bottom stress, Function/fluxes.f90:261-270
__global__ void k_bot_stress(DevGrid g, const double *__restrict__ bf, const double *__restrict__ uup,
                             const double *__restrict__ vvp, double *__restrict__ bx, double *__restrict__ by)
{
    KERNEL_MN
    if (!IN_INTERIOR(g)) return;
    const size_t i = I(m, n);
    if (g.lcu[i] > 0.5f) bx[i] = -(bf[i] + bf[I(m + 1, n)]) / 2.0 * uup[i];
    if (g.lcv[i] > 0.5f) by[i] = -(bf[i] + bf[I(m, n + 1)]) / 2.0 * vvp[i];
}

// depth_ave (Function/depth.f90:223-254) with kbcl = 0 over
// bnd_x1+1..bnd_x2-1 x bnd_y1+1..bnd_y2-1, on mask llu (v = 0) or llv (v = 1)
__global__ void k_depth_ave(DevGrid g, int nz, const double *__restrict__ dz, const double *__restrict__ u,
                            int v, double *__restrict__ utr)
{
    KERNEL_MN
    if (m < 2 || m > g.nx - 1 || n < 2 || n > g.ny - 1) return;
    const size_t i = I(m, n), plane = (size_t)g.nx * g.ny;
    if ((v ? g.llv[i] : g.llu[i]) > 0.5f) {
        double f = 0.0;
        for (int k = 0; k < nz; ++k) f = f + u[k * plane + i] * dz[k];
        utr[i] = f;
    }
}

struct AsmArgs {
    const double *dfx, *dfy, *tx2, *ty2, *txd, *tyd, *sfx, *sfy, *bfx, *bfy, *hhu, *hhv, *slpr;
    double *rhsx, *rhsy;
};

// Control/dyncall.f90:107-126
__global__ void k_forcing_assembly(DevGrid g, AsmArgs a)
{
    KERNEL_MN
    if (!IN_INTERIOR(g)) return;
    const size_t i = I(m, n);
    if (g.lcu[i] > 0.5f)
        a.rhsx[i] = a.rhsx[i] - a.dfx[i] + a.tx2[i] - a.txd[i] + (a.sfx[i] + a.bfx[i]) * g.dxt[i] * g.dyh[i] -
                    (a.slpr[I(m + 1, n)] - a.slpr[i]) * a.hhu[i] * g.dyh[i] / g.rho0;
    if (g.lcv[i] > 0.5f)
        a.rhsy[i] = a.rhsy[i] - a.dfy[i] + a.ty2[i] - a.tyd[i] + (a.sfy[i] + a.bfy[i]) * g.dyt[i] * g.dxh[i] -
                    (a.slpr[I(m, n + 1)] - a.slpr[i]) * a.hhv[i] * g.dxh[i] / g.rho0;
}

inline dim3 grid2(const DevGrid &g) { return dim3((g.nx + 63) / 64, (g.ny + 3) / 4); }
const dim3 BLK2(64, 4);

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

// staging: 3-D inputs and work arrays (nz planes each), 2-D planes, dz
enum { UU, VV, UUP, VVP, AMUV, PGX, PGY, ST3, SS3, VORT3, TX3, TY3, RX3, RY3, N3 };
enum { UU2D, VV2D, UUP2D, VVP2D, AMUV2D, HHQ, HHU, HHV, HHH, SLPR, TAUX, TAUY, RHSX, RHSY,
       ST2, SS2, VORT2, TXD, TYD, DFX, DFY, SFX, SFY, BFX, BFY, BF, TX2, TY2, DZ, N2 };

size_t need_doubles(const inmsom_bt_ctx *c, int nz) { return ((size_t)N3 * nz + N2) * c->n; }

int ensure_staging(inmsom_bt_ctx *c, int nz)
{
    const size_t need = need_doubles(c, nz);
    if (c->f3 && c->f3_n >= need) return INMSOM_BT_OK;
    if (c->f3) CHF(hipFree(c->f3));
    c->f3 = nullptr;
    c->f3_n = 0;
    CHF(hipMalloc(&c->f3, need * sizeof(double)));
    c->f3_n = need;
    c->f3_nz = 0;
    return INMSOM_BT_OK;
}

inline double *p3(inmsom_bt_ctx *c, int nz, int k) { return c->f3 + (size_t)k * nz * c->n; }
inline double *p2(inmsom_bt_ctx *c, int nz, int k) { return c->f3 + ((size_t)N3 * nz + k) * c->n; }

// the device work of one call on the staged fields (with_pg: add the
// pressure-gradient terms)
void launch_forcing(inmsom_bt_ctx *c, int nz, bool with_pg)
{
    const DevGrid &g = c->g;
    const size_t pl = c->n, n3 = (size_t)nz * pl;
    auto fill = [&](double *a, size_t count) {
        hipLaunchKernelGGL(k_fill3, dim3((unsigned)((count + 255) / 256)), dim3(256), 0, c->stream, count, a, 0.0);
    };
    // work arrays start at zero (the reference's RHSx3d = 0 and calloc'd scratch)
    for (int k = ST3; k <= RY3; ++k) fill(p3(c, nz, k), n3);
    for (int k = ST2; k <= TY2; ++k) fill(p2(c, nz, k), pl);
    // dyncall.f90:38-39
    for (int k = 0; k < nz; ++k)
        lit_stress(c, p3(c, nz, UUP) + k * pl, p3(c, nz, VVP) + k * pl, p3(c, nz, ST3) + k * pl,
                   p3(c, nz, SS3) + k * pl);
    lit_stress(c, p2(c, nz, UUP2D), p2(c, nz, VVP2D), p2(c, nz, ST2), p2(c, nz, SS2));
    // :53-59
    lit_surface_stress(c, p2(c, nz, TAUX), p2(c, nz, TAUY), p2(c, nz, SFX), p2(c, nz, SFY));
    const size_t bot = (size_t)(nz - 1) * pl;
    lit_bottom_friction(c, p3(c, nz, UU) + bot, p3(c, nz, VV) + bot, p2(c, nz, HHU), p2(c, nz, HHV),
                        p2(c, nz, HHQ), p2(c, nz, BF));
    hipLaunchKernelGGL(k_bot_stress, grid2(g), BLK2, 0, c->stream, g, p2(c, nz, BF), p3(c, nz, UUP) + bot,
                       p3(c, nz, VVP) + bot, p2(c, nz, BFX), p2(c, nz, BFY));
    // :65-81, level by level and on the 2-D velocities
    for (int k = 0; k < nz; ++k) {
        const size_t o = k * pl;
        lit_trans_diff2(c, p3(c, nz, UU) + o, p3(c, nz, VV) + o, p3(c, nz, VORT3) + o, p2(c, nz, HHQ),
                        p2(c, nz, HHU), p2(c, nz, HHV), p2(c, nz, HHH), p3(c, nz, AMUV) + o, p3(c, nz, ST3) + o,
                        p3(c, nz, SS3) + o, p3(c, nz, TX3) + o, p3(c, nz, TY3) + o, p3(c, nz, RX3) + o,
                        p3(c, nz, RY3) + o);
    }
    lit_trans_diff2(c, p2(c, nz, UU2D), p2(c, nz, VV2D), p2(c, nz, VORT2), p2(c, nz, HHQ), p2(c, nz, HHU),
                    p2(c, nz, HHV), p2(c, nz, HHH), p2(c, nz, AMUV2D), p2(c, nz, ST2), p2(c, nz, SS2),
                    p2(c, nz, TXD), p2(c, nz, TYD), p2(c, nz, DFX), p2(c, nz, DFY));
    // :92-98
    if (with_pg) {
        const dim3 gr((unsigned)((n3 + 255) / 256));
        hipLaunchKernelGGL(k_add3, gr, dim3(256), 0, c->stream, n3, p3(c, nz, RX3), p3(c, nz, PGX));
        hipLaunchKernelGGL(k_add3, gr, dim3(256), 0, c->stream, n3, p3(c, nz, RY3), p3(c, nz, PGY));
    }
    // :100-104
    const double *dz = p2(c, nz, DZ);
    hipLaunchKernelGGL(k_depth_ave, grid2(g), BLK2, 0, c->stream, g, nz, dz, p3(c, nz, RX3), 0, p2(c, nz, RHSX));
    hipLaunchKernelGGL(k_depth_ave, grid2(g), BLK2, 0, c->stream, g, nz, dz, p3(c, nz, RY3), 1, p2(c, nz, RHSY));
    hipLaunchKernelGGL(k_depth_ave, grid2(g), BLK2, 0, c->stream, g, nz, dz, p3(c, nz, TX3), 0, p2(c, nz, TX2));
    hipLaunchKernelGGL(k_depth_ave, grid2(g), BLK2, 0, c->stream, g, nz, dz, p3(c, nz, TY3), 1, p2(c, nz, TY2));
    // :107-126
    AsmArgs a{p2(c, nz, DFX), p2(c, nz, DFY), p2(c, nz, TX2), p2(c, nz, TY2), p2(c, nz, TXD), p2(c, nz, TYD),
              p2(c, nz, SFX), p2(c, nz, SFY), p2(c, nz, BFX), p2(c, nz, BFY), p2(c, nz, HHU), p2(c, nz, HHV),
              p2(c, nz, SLPR), p2(c, nz, RHSX), p2(c, nz, RHSY)};
    hipLaunchKernelGGL(k_forcing_assembly, grid2(g), BLK2, 0, c->stream, g, a);
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
    if (nz < 1 || nz > 1000) return ffail("nz must be in 1..1000", INMSOM_BT_EINVAL);
    if ((size_t)nz > c->n) return ffail("nz exceeds the plane size", INMSOM_BT_EINVAL);
    CHF(hipSetDevice(c->device));
    TRYF(ensure_staging(c, nz));
    const size_t pl = c->n, n3 = (size_t)nz * pl;
    auto put = [&](double *dst, const double *src, size_t count) {
        return hipMemcpyAsync(dst, src, count * sizeof(double), hipMemcpyHostToDevice, c->stream);
    };
    const double *in3[7] = {uu, vv, uup, vvp, amuv, pgx, pgy};
    for (int k = 0; k < 7; ++k)
        if (in3[k]) CHF(put(p3(c, nz, k), in3[k], n3));
    const double *in2[14] = {uu2d, vv2d, uup2d, vvp2d, amuv2d, hhq, hhu, hhv, hhh, slpr, taux, tauy, rhsx2d, rhsy2d};
    for (int k = 0; k < 14; ++k) CHF(put(p2(c, nz, k), in2[k], pl));
    CHF(put(p2(c, nz, DZ), dz, nz));
    launch_forcing(c, nz, pgx != nullptr);
    CHF(hipGetLastError());
    CHF(hipMemcpyAsync(rhsx2d, p2(c, nz, RHSX), pl * sizeof(double), hipMemcpyDeviceToHost, c->stream));
    CHF(hipMemcpyAsync(rhsy2d, p2(c, nz, RHSY), pl * sizeof(double), hipMemcpyDeviceToHost, c->stream));
    CHF(hipStreamSynchronize(c->stream));
    c->f3_nz = nz;
    c->f3_pg = pgx != nullptr;
    return INMSOM_BT_OK;
}

int inmsom_bt_forcing_3d_timed(inmsom_bt_ctx *c, int nz, int nrep, double *ms)
{
    if (!c || !ms) return ffail("null pointer", INMSOM_BT_EINVAL);
    if (nrep < 1) return ffail("nrep must be >= 1", INMSOM_BT_EINVAL);
    if (!c->f3 || c->f3_nz != nz)
        return ffail("no forcing fields staged for nz levels: call inmsom_bt_forcing_3d first", INMSOM_BT_ESTATE);
    CHF(hipSetDevice(c->device));
    hipEvent_t e0, e1;
    CHF(hipEventCreate(&e0));
    CHF(hipEventCreate(&e1));
    CHF(hipEventRecord(e0, c->stream));
    for (int r = 0; r < nrep; ++r) launch_forcing(c, nz, c->f3_pg);
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
