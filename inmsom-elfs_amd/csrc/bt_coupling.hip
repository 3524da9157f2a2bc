// The 2-D -> 3-D coupling that follows barotropic_dynamics in the model step
// (Control/dyncall.f90:185-231, without vertical_velocity, which the 3-D
// caller runs between the two entry points): the barotropic part removed
// from the 3-D velocity, the time-mean 2-D velocity and the sea-level
// pressure term of the pressure gradient, ssh_internal
// (Function/vel_ssh.f90:682-751), and the barotropic part added back.
// Reference index ranges, mask tests and operation order; compiled with
// -ffp-contract=off like the other reference-order kernels.
#include <string>

#include "bt_internal.h"
#include "bt_literal.h"

namespace ibt {

namespace {

#define KERNEL_MN                                            \
    const int m = blockIdx.x * blockDim.x + threadIdx.x + 1; \
    const int n = blockIdx.y * blockDim.y + threadIdx.y + 1;
#define I(mm, nn) IDX(g, (mm), (nn))
// nx_start-1 .. nx_end+1 x ny_start-1 .. ny_end+1 on one rank
#define IN_WIDE(g) (m >= 2 && m <= (g).nx - 1 && n >= 2 && n <= (g).ny - 1)

// dyncall.f90:185-199 (sign = -1) and :217-231 (sign = +1), levels 1..nz
__global__ void k_barotropic_part(DevGrid g, int nz, double sign, double *__restrict__ uu,
                                  double *__restrict__ vv, const double *__restrict__ uu2d,
                                  const double *__restrict__ vv2d)
{
    KERNEL_MN
    if (!IN_WIDE(g)) return;
    const size_t i = I(m, n), plane = (size_t)g.nx * g.ny;
    if (g.lcu[i] > 0.5f)
        for (int k = 0; k < nz; ++k)
            uu[k * plane + i] = sign < 0.0 ? uu[k * plane + i] - uu2d[i] : uu[k * plane + i] + uu2d[i];
    if (g.lcv[i] > 0.5f)
        for (int k = 0; k < nz; ++k)
            vv[k * plane + i] = sign < 0.0 ? vv[k * plane + i] - vv2d[i] : vv[k * plane + i] + vv2d[i];
}

// dyncall.f90:201-215
__global__ void k_internal_uv(DevGrid g, const double *__restrict__ ubrtr_i, const double *__restrict__ vbrtr_i,
                              const double *__restrict__ hhu, const double *__restrict__ hhv,
                              const double *__restrict__ slpr, double *__restrict__ uu2d,
                              double *__restrict__ vv2d, double *__restrict__ pgrx, double *__restrict__ pgry)
{
    KERNEL_MN
    if (!IN_WIDE(g)) return;
    const size_t i = I(m, n);
    if (g.lcu[i] > 0.5f) {
        uu2d[i] = ubrtr_i[i] / hhu[i];
        pgrx[i] = pgrx[i] - (slpr[I(m + 1, n)] - slpr[i]) * hhu[i] * g.dyh[i] / g.rho0;
    }
    if (g.lcv[i] > 0.5f) {
        vv2d[i] = vbrtr_i[i] / hhv[i];
        pgry[i] = pgry[i] - (slpr[I(m, n + 1)] - slpr[i]) * hhv[i] * g.dxh[i] / g.rho0;
    }
}

// ssh_internal, vel_ssh.f90:705-720 (ssh_new zero outside the wet interior)
__global__ void k_ssh_internal(DevGrid g, double tau, const double *__restrict__ sshp_i,
                               const double *__restrict__ u, const double *__restrict__ v,
                               const double *__restrict__ wflux, double *__restrict__ ssh_new)
{
    KERNEL_MN
    if (m > g.nx || n > g.ny) return;
    const size_t i = I(m, n);
    double s = 0.0;
    if (m >= 3 && m <= g.nx - 2 && n >= 3 && n <= g.ny - 2 && g.lu[i] > 0.5f)
        s = sshp_i[i] + 2.0 * tau *
                            (wflux[i] / g.rho0 * (double)g.ffs -
                             (u[i] * g.dyh[i] - u[I(m - 1, n)] * g.dyh[I(m - 1, n)] + v[i] * g.dxh[i] -
                              v[I(m, n - 1)] * g.dxh[I(m, n - 1)]) /
                                 (g.dx[i] * g.dy[i]));
    ssh_new[i] = s;
}

// vel_ssh.f90:737-748
__global__ void k_ssh_internal_shift(DevGrid g, const double *__restrict__ ssh_new, double *__restrict__ ssh_i,
                                     double *__restrict__ sshp_i)
{
    KERNEL_MN
    if (!IN_WIDE(g)) return;
    const size_t i = I(m, n);
    if (g.lu[i] > 0.5f) {
        sshp_i[i] = ssh_i[i] + g.ts * (ssh_new[i] - 2.0 * ssh_i[i] + sshp_i[i]) / 2.0;
        ssh_i[i] = ssh_new[i];
    }
}

inline dim3 grid2(const DevGrid &g) { return dim3((g.nx + 63) / 64, (g.ny + 3) / 4); }
const dim3 BLK2(64, 4);

int cfail(const std::string &m, int code)
{
    set_error(m);
    return code;
}

#define CHC(expr)                                                                                   \
    do {                                                                                            \
        hipError_t e_ = (expr);                                                                     \
        if (e_ != hipSuccess) return cfail(std::string(#expr) + ": " + hipGetErrorString(e_),        \
                                           INMSOM_BT_EDEVICE);                                      \
    } while (0)

// device staging of the coupling: 17 planes + uu, vv (nz planes each)
constexpr int CPL_2D = 17;
enum { UBI = 0, VBI, HHU, HHV, SLPR, WFL, UU2D, VV2D, PGRX, PGRY, SSHI, SSHPI, HQN, HUN, HVN, HHN, SSHNEW };

int ensure_staging(inmsom_bt_ctx *c, int nz)
{
    const size_t need = (size_t)(CPL_2D + 2 * nz) * c->n;
    if (c->cpl && c->cpl_n >= need) return INMSOM_BT_OK;
    if (c->cpl) CHC(hipFree(c->cpl));
    c->cpl = nullptr;
    c->cpl_n = 0;
    CHC(hipMalloc(&c->cpl, need * sizeof(double)));
    c->cpl_n = need;
    return INMSOM_BT_OK;
}

inline double *plane(inmsom_bt_ctx *c, int k) { return c->cpl + (size_t)k * c->n; }

int put(inmsom_bt_ctx *c, double *dst, const double *src, size_t count)
{
    CHC(hipMemcpyAsync(dst, src, count * sizeof(double), hipMemcpyHostToDevice, c->stream));
    return INMSOM_BT_OK;
}

int get(inmsom_bt_ctx *c, double *dst, const double *src, size_t count)
{
    CHC(hipMemcpyAsync(dst, src, count * sizeof(double), hipMemcpyDeviceToHost, c->stream));
    return INMSOM_BT_OK;
}

#define TRYC(expr)                         \
    do {                                   \
        int r_ = (expr);                   \
        if (r_ != INMSOM_BT_OK) return r_; \
    } while (0)

}  // namespace

}  // namespace ibt

using namespace ibt;

extern "C" {

int inmsom_bt_remove_barotropic(inmsom_bt_ctx *c, int nz, double *uu, double *vv, const double *uu2d,
                                const double *vv2d)
{
    if (!c || !uu || !vv || !uu2d || !vv2d) return cfail("null pointer", INMSOM_BT_EINVAL);
    if (nz < 1) return cfail("nz must be >= 1", INMSOM_BT_EINVAL);
    CHC(hipSetDevice(c->device));
    TRYC(ensure_staging(c, nz));
    const size_t n = c->n, n3 = (size_t)nz * n;
    double *duu = plane(c, CPL_2D), *dvv = duu + n3;
    TRYC(put(c, duu, uu, n3));
    TRYC(put(c, dvv, vv, n3));
    TRYC(put(c, plane(c, UU2D), uu2d, n));
    TRYC(put(c, plane(c, VV2D), vv2d, n));
    hipLaunchKernelGGL(k_barotropic_part, grid2(c->g), BLK2, 0, c->stream, c->g, nz, -1.0, duu, dvv,
                       plane(c, UU2D), plane(c, VV2D));
    CHC(hipGetLastError());
    TRYC(get(c, uu, duu, n3));
    TRYC(get(c, vv, dvv, n3));
    CHC(hipStreamSynchronize(c->stream));
    return INMSOM_BT_OK;
}

int inmsom_bt_internal_update(inmsom_bt_ctx *c, double tau, int nz, const double *ubrtr_i, const double *vbrtr_i,
                              const double *hhu, const double *hhv, const double *slpr, const double *wflux,
                              double *uu2d, double *vv2d, double *pgrx, double *pgry, double *ssh_i,
                              double *sshp_i, double *const hhn[4], double *uu, double *vv)
{
    if (!c || !ubrtr_i || !vbrtr_i || !hhu || !hhv || !slpr || !wflux || !uu2d || !vv2d || !pgrx || !pgry ||
        !ssh_i || !sshp_i || !hhn || !uu || !vv)
        return cfail("null pointer", INMSOM_BT_EINVAL);
    for (int k = 0; k < 4; ++k)
        if (!hhn[k]) return cfail("null depth pointer", INMSOM_BT_EINVAL);
    if (nz < 1) return cfail("nz must be >= 1", INMSOM_BT_EINVAL);
    if (!(tau > 0.0)) return cfail("tau must be positive", INMSOM_BT_EINVAL);
    CHC(hipSetDevice(c->device));
    TRYC(ensure_staging(c, nz));
    const size_t n = c->n, n3 = (size_t)nz * n;
    const double *in2[8] = {ubrtr_i, vbrtr_i, hhu, hhv, slpr, wflux, nullptr, nullptr};
    for (int k = 0; k < 6; ++k) TRYC(put(c, plane(c, k), in2[k], n));
    double *io2[6] = {uu2d, vv2d, pgrx, pgry, ssh_i, sshp_i};
    for (int k = 0; k < 6; ++k) TRYC(put(c, plane(c, UU2D + k), io2[k], n));
    for (int k = 0; k < 4; ++k) TRYC(put(c, plane(c, HQN + k), hhn[k], n));
    double *duu = plane(c, CPL_2D), *dvv = duu + n3;
    TRYC(put(c, duu, uu, n3));
    TRYC(put(c, dvv, vv, n3));

    const DevGrid &g = c->g;
    hipLaunchKernelGGL(k_internal_uv, grid2(g), BLK2, 0, c->stream, g, plane(c, UBI), plane(c, VBI),
                       plane(c, HHU), plane(c, HHV), plane(c, SLPR), plane(c, UU2D), plane(c, VV2D),
                       plane(c, PGRX), plane(c, PGRY));
    hipLaunchKernelGGL(k_ssh_internal, grid2(g), BLK2, 0, c->stream, g, tau, plane(c, SSHPI), plane(c, UBI),
                       plane(c, VBI), plane(c, WFL), plane(c, SSHNEW));
    lit_cyclize(c, plane(c, SSHNEW));
    if (g.ffs > 0)
        lit_hh_update(c, plane(c, SSHNEW), plane(c, HQN), plane(c, HUN), plane(c, HVN), plane(c, HHN));
    hipLaunchKernelGGL(k_ssh_internal_shift, grid2(g), BLK2, 0, c->stream, g, plane(c, SSHNEW), plane(c, SSHI),
                       plane(c, SSHPI));
    hipLaunchKernelGGL(k_barotropic_part, grid2(g), BLK2, 0, c->stream, g, nz, 1.0, duu, dvv, plane(c, UU2D),
                       plane(c, VV2D));
    CHC(hipGetLastError());
    for (int k = 0; k < 6; ++k) TRYC(get(c, io2[k], plane(c, UU2D + k), n));
    for (int k = 0; k < 4; ++k) TRYC(get(c, hhn[k], plane(c, HQN + k), n));
    TRYC(get(c, uu, duu, n3));
    TRYC(get(c, vv, dvv, n3));
    CHC(hipStreamSynchronize(c->stream));
    return INMSOM_BT_OK;
}

}  // extern "C"
