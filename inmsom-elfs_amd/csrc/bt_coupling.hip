// The 2-D -> 3-D coupling that follows barotropic_dynamics in the model step
// (Control/dyncall.f90:169-231, without vertical_velocity, which the 3-D
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

// (m, n) of a 2-D launch over the whole array; the blocks go to the XCDs
// in contiguous row-major runs (xcd_tile)
#define KERNEL_MN                                                                                   \
    int bx_, by_;                                                                                   \
    xcd_tile((int)gridDim.x, (int)gridDim.y, (int)(blockIdx.y * gridDim.x + blockIdx.x), bx_, by_); \
    const int m = bx_ * (int)blockDim.x + (int)threadIdx.x + 1;                                     \
    const int n = by_ * (int)blockDim.y + (int)threadIdx.y + 1;
#define I(mm, nn) IDX(g, (mm), (nn))
// nx_start-1 .. nx_end+1 x ny_start-1 .. ny_end+1 on one rank
#define IN_WIDE(g) (m >= 2 && m <= (g).nx - 1 && n >= 2 && n <= (g).ny - 1)

// dyncall.f90:169-184 (sign = -1) and :216-231 (sign = +1), levels 1..nz.
// One thread per cell in flat (n, m) order, so that a wave's 64 cells are
// consecutive in memory and its accesses start on a cache line whatever
// nx is (rows of 64-wide blocks would not); the thread walks the levels,
// reading the mask and the 2-D operand once.
__global__ void k_barotropic_part(DevGrid g, int nz, double sign, double *__restrict__ uu,
                                  double *__restrict__ vv, const double *__restrict__ uu2d,
                                  const double *__restrict__ vv2d)
{
    const size_t plane = (size_t)g.nx * g.ny;
    const size_t i = (size_t)blockIdx.x * blockDim.x + threadIdx.x;
    if (i >= plane) return;
    const int m = (int)(i % (size_t)g.nx) + 1, n = (int)(i / (size_t)g.nx) + 1;
    if (!IN_WIDE(g)) return;
    if (g.lcu[i] > 0.5f) {
        const double a = uu2d[i];
        for (int k = 0; k < nz; ++k) uu[k * plane + i] = sign < 0.0 ? uu[k * plane + i] - a : uu[k * plane + i] + a;
    }
    if (g.lcv[i] > 0.5f) {
        const double a = vv2d[i];
        for (int k = 0; k < nz; ++k) vv[k * plane + i] = sign < 0.0 ? vv[k * plane + i] - a : vv[k * plane + i] + a;
    }
}

// dyncall.f90:188-206
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


// The arrays one coupling works on (device addresses).
struct Cpl {
    const double *ubrtr_i, *vbrtr_i, *hhu, *hhv, *slpr, *wflux;
    double *uu2d, *vv2d, *pgrx, *pgry, *ssh_i, *sshp_i, *hhn[4], *ssh_new, *uu, *vv;
};

// the staged planes of the host-array entry points
Cpl staged(inmsom_bt_ctx *c, int nz)
{
    Cpl p;
    p.ubrtr_i = plane(c, UBI);
    p.vbrtr_i = plane(c, VBI);
    p.hhu = plane(c, HHU);
    p.hhv = plane(c, HHV);
    p.slpr = plane(c, SLPR);
    p.wflux = plane(c, WFL);
    p.uu2d = plane(c, UU2D);
    p.vv2d = plane(c, VV2D);
    p.pgrx = plane(c, PGRX);
    p.pgry = plane(c, PGRY);
    p.ssh_i = plane(c, SSHI);
    p.sshp_i = plane(c, SSHPI);
    for (int k = 0; k < 4; ++k) p.hhn[k] = plane(c, HQN + k);
    p.ssh_new = plane(c, SSHNEW);
    p.uu = plane(c, CPL_2D);
    p.vv = p.uu + (size_t)nz * c->n;
    return p;
}

// Control/dyncall.f90:169-184
void launch_part(inmsom_bt_ctx *c, int nz, double sign, const Cpl &p)
{
    hipLaunchKernelGGL(k_barotropic_part, dim3((unsigned)((c->n + 255) / 256)), dim3(256), 0, c->stream, c->g, nz,
                       sign, p.uu, p.vv, p.uu2d, p.vv2d);
}

void launch_remove(inmsom_bt_ctx *c, int nz, const Cpl &p) { launch_part(c, nz, -1.0, p); }

// Control/dyncall.f90:188-231 (vertical_velocity excluded)
void launch_update(inmsom_bt_ctx *c, double tau, int nz, const Cpl &p)
{
    const DevGrid &g = c->g;
    hipLaunchKernelGGL(k_internal_uv, grid2(g), BLK2, 0, c->stream, g, p.ubrtr_i, p.vbrtr_i, p.hhu, p.hhv, p.slpr,
                       p.uu2d, p.vv2d, p.pgrx, p.pgry);
    hipLaunchKernelGGL(k_ssh_internal, grid2(g), BLK2, 0, c->stream, g, tau, p.sshp_i, p.ubrtr_i, p.vbrtr_i,
                       p.wflux, p.ssh_new);
    lit_cyclize(c, p.ssh_new);
    if (g.ffs > 0) lit_hh_update(c, p.ssh_new, p.hhn[0], p.hhn[1], p.hhn[2], p.hhn[3]);
    hipLaunchKernelGGL(k_ssh_internal_shift, grid2(g), BLK2, 0, c->stream, g, p.ssh_new, p.ssh_i, p.sshp_i);
    launch_part(c, nz, 1.0, p);
}

int ensure_new_plane(inmsom_bt_ctx *c)
{
    if (!c->cpl_new) CHC(hipMalloc(&c->cpl_new, c->n * sizeof(double)));
    return INMSOM_BT_OK;
}

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
    launch_remove(c, nz, staged(c, nz));
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
    const double *in2[6] = {ubrtr_i, vbrtr_i, hhu, hhv, slpr, wflux};
    for (int k = 0; k < 6; ++k) TRYC(put(c, plane(c, k), in2[k], n));
    double *io2[6] = {uu2d, vv2d, pgrx, pgry, ssh_i, sshp_i};
    for (int k = 0; k < 6; ++k) TRYC(put(c, plane(c, UU2D + k), io2[k], n));
    for (int k = 0; k < 4; ++k) TRYC(put(c, plane(c, HQN + k), hhn[k], n));
    double *duu = plane(c, CPL_2D), *dvv = duu + n3;
    TRYC(put(c, duu, uu, n3));
    TRYC(put(c, dvv, vv, n3));

    launch_update(c, tau, nz, staged(c, nz));
    CHC(hipGetLastError());
    for (int k = 0; k < 6; ++k) TRYC(get(c, io2[k], plane(c, UU2D + k), n));
    for (int k = 0; k < 4; ++k) TRYC(get(c, hhn[k], plane(c, HQN + k), n));
    TRYC(get(c, uu, duu, n3));
    TRYC(get(c, vv, dvv, n3));
    CHC(hipStreamSynchronize(c->stream));
    return INMSOM_BT_OK;
}

int inmsom_bt_coupling_timed(inmsom_bt_ctx *c, double tau, int nz, int nrep, double *ms)
{
    if (!c || !ms) return cfail("null pointer", INMSOM_BT_EINVAL);
    if (nz < 1 || nrep < 1) return cfail("nz and nrep must be >= 1", INMSOM_BT_EINVAL);
    if (!(tau > 0.0)) return cfail("tau must be positive", INMSOM_BT_EINVAL);
    if (!c->cpl || c->cpl_n < (size_t)(CPL_2D + 2 * nz) * c->n)
        return cfail("no coupling fields staged for nz levels: call inmsom_bt_internal_update first",
                     INMSOM_BT_ESTATE);
    CHC(hipSetDevice(c->device));
    hipEvent_t e0, e1;
    CHC(hipEventCreate(&e0));
    CHC(hipEventCreate(&e1));
    CHC(hipEventRecord(e0, c->stream));
    const Cpl p = staged(c, nz);
    for (int r = 0; r < nrep; ++r) {
        launch_remove(c, nz, p);
        launch_update(c, tau, nz, p);
    }
    CHC(hipGetLastError());
    CHC(hipEventRecord(e1, c->stream));
    CHC(hipEventSynchronize(e1));
    float t = 0.0f;
    CHC(hipEventElapsedTime(&t, e0, e1));
    CHC(hipEventDestroy(e0));
    CHC(hipEventDestroy(e1));
    *ms = t;
    return INMSOM_BT_OK;
}

int inmsom_bt_remove_barotropic_dev(inmsom_bt_ctx *c, int nz, double *uu, double *vv, const double *uu2d,
                                    const double *vv2d)
{
    if (!c || !uu || !vv || !uu2d || !vv2d) return cfail("null pointer", INMSOM_BT_EINVAL);
    if (nz < 1) return cfail("nz must be >= 1", INMSOM_BT_EINVAL);
    CHC(hipSetDevice(c->device));
    Cpl p{};
    p.uu = uu;
    p.vv = vv;
    p.uu2d = const_cast<double *>(uu2d);   // read only by the removal kernel
    p.vv2d = const_cast<double *>(vv2d);
    launch_remove(c, nz, p);
    CHC(hipGetLastError());
    CHC(hipStreamSynchronize(c->stream));
    return INMSOM_BT_OK;
}

int inmsom_bt_internal_update_dev(inmsom_bt_ctx *c, double tau, int nz, const double *ubrtr_i,
                                  const double *vbrtr_i, const double *hhu, const double *hhv, const double *slpr,
                                  const double *wflux, double *uu2d, double *vv2d, double *pgrx, double *pgry,
                                  double *ssh_i, double *sshp_i, double *const hhn[4], double *uu, double *vv)
{
    if (!c || !ubrtr_i || !vbrtr_i || !hhu || !hhv || !slpr || !wflux || !uu2d || !vv2d || !pgrx || !pgry ||
        !ssh_i || !sshp_i || !hhn || !uu || !vv)
        return cfail("null pointer", INMSOM_BT_EINVAL);
    for (int k = 0; k < 4; ++k)
        if (!hhn[k]) return cfail("null depth pointer", INMSOM_BT_EINVAL);
    if (nz < 1) return cfail("nz must be >= 1", INMSOM_BT_EINVAL);
    if (!(tau > 0.0)) return cfail("tau must be positive", INMSOM_BT_EINVAL);
    CHC(hipSetDevice(c->device));
    TRYC(ensure_new_plane(c));
    Cpl p{ubrtr_i, vbrtr_i, hhu, hhv, slpr, wflux, uu2d, vv2d, pgrx, pgry, ssh_i, sshp_i,
          {hhn[0], hhn[1], hhn[2], hhn[3]}, c->cpl_new, uu, vv};
    launch_update(c, tau, nz, p);
    CHC(hipGetLastError());
    CHC(hipStreamSynchronize(c->stream));
    return INMSOM_BT_OK;
}

}  // extern "C"
