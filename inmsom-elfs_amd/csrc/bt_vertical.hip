// vertical_velocity (Function/vel_ssh.f90:1030-1076): w of the 3-D model
// from the continuity equation, integrated from the bottom (w(nz+1) = 0)
// on the lu columns of the interior, then wrapped on periodic boundaries
// (Service/oc_algorithms.f90:50-95, :118-135) for all nz+1 levels.
// Cells the reference does not compute keep the caller's values.
//
// k_vertical: one thread per interior (m, n), the column walked from the
// bottom in registers, in 64 x 4 tiles dealt to the XCDs in contiguous
// runs; consecutive lanes take consecutive m, so each level's loads and
// stores are coalesced.  The wrap reads the computed
// values and runs as two launches after it (x first, then y, as the
// reference orders them, which fixes the corners).
#include <string>

#include "bt_internal.h"

namespace ibt {

namespace {

struct VvParams {
    DevGrid g;
    int nz;
    int tiles_x, tiles_y;       // 64 x 4 tiles over the interior
    const double *dz;           // nz (device)
    const double *u, *v;        // (nz, ny, nx)
    double *w;                  // (nz + 1, ny, nx)
    const double *hu, *hv, *wflux;
};

#define I(mm, nn) IDX(g, (mm), (nn))

__global__ __launch_bounds__(256) void k_vertical(VvParams p)
{
    const DevGrid &g = p.g;
    int bx, by;
    xcd_tile(p.tiles_x, p.tiles_y, (int)blockIdx.x, bx, by);
    const int m = bx * 64 + (int)(threadIdx.x & 63) + 3;
    const int n = by * 4 + (int)(threadIdx.x >> 6) + 3;
    if (m > g.nx - 2 || n > g.ny - 2) return;
    const size_t pl = (size_t)g.nx * g.ny, i = I(m, n), iw = I(m - 1, n), is = I(m, n - 1);
    if (!(g.lu[i] > 0.5f)) return;
    const double ffs = (double)g.ffs;
    const double dxy = g.dx[i] * g.dy[i];
    const double hu0 = p.hu[i], hu1 = p.hu[iw], hv0 = p.hv[i], hv1 = p.hv[is];
    const double yh0 = g.dyh[i], yh1 = g.dyh[iw], xh0 = g.dxh[i], xh1 = g.dxh[is];
    const double src = p.wflux[i] / g.rho0 * ffs;
    double wk = 0.0;
    p.w[(size_t)p.nz * pl + i] = 0.0;
    for (int k = p.nz; k >= 1; --k) {
        const size_t o = (size_t)(k - 1) * pl;
        wk = wk + p.dz[k - 1] / dxy *
                      (p.u[o + i] * hu0 * yh0 - p.u[o + iw] * hu1 * yh1 + p.v[o + i] * hv0 * xh0 -
                       p.v[o + is] * hv1 * xh1 + src);
        p.w[o + i] = wk;
    }
}

// column mm -> 2, 3 -> nx - 1 (x); row nn -> 2, 3 -> ny - 1 (y), all levels
__global__ void k_wrap_x(DevGrid g, int nlev, double *w)
{
    const int n = blockIdx.x * blockDim.x + threadIdx.x + 1;
    const int k = blockIdx.y;
    if (n > g.ny || k >= nlev) return;
    double *f = w + (size_t)k * g.nx * g.ny;
    const double a = f[I(g.nx - 2, n)], b = f[I(3, n)];
    f[I(2, n)] = a;
    f[I(g.nx - 1, n)] = b;
}

__global__ void k_wrap_y(DevGrid g, int nlev, double *w)
{
    const int m = blockIdx.x * blockDim.x + threadIdx.x + 1;
    const int k = blockIdx.y;
    if (m > g.nx || k >= nlev) return;
    double *f = w + (size_t)k * g.nx * g.ny;
    const double a = f[I(m, g.ny - 2)], b = f[I(m, 3)];
    f[I(m, 2)] = a;
    f[I(m, g.ny - 1)] = b;
}

#undef I

int vfail(const std::string &m, int code)
{
    set_error(m);
    return code;
}

#define CHV(expr)                                                                                   \
    do {                                                                                            \
        hipError_t e_ = (expr);                                                                     \
        if (e_ != hipSuccess) return vfail(std::string(#expr) + ": " + hipGetErrorString(e_),        \
                                           INMSOM_BT_EDEVICE);                                      \
    } while (0)

void launch(inmsom_bt_ctx *c, VvParams p)
{
    p.tiles_x = (c->nx - 4 + 63) / 64;
    p.tiles_y = (c->ny - 4 + 3) / 4;
    hipLaunchKernelGGL(k_vertical, dim3((unsigned)(p.tiles_x * p.tiles_y)), dim3(256), 0, c->stream, p);
    if (c->g.px)
        hipLaunchKernelGGL(k_wrap_x, dim3((unsigned)((c->ny + 255) / 256), (unsigned)(p.nz + 1)), dim3(256), 0,
                           c->stream, c->g, p.nz + 1, p.w);
    if (c->g.py)
        hipLaunchKernelGGL(k_wrap_y, dim3((unsigned)((c->nx + 255) / 256), (unsigned)(p.nz + 1)), dim3(256), 0,
                           c->stream, c->g, p.nz + 1, p.w);
}

int check(inmsom_bt_ctx *c, int nz)
{
    if (!c) return vfail("null pointer", INMSOM_BT_EINVAL);
    if (nz < 1 || nz > 1000) return vfail("nz must be in 1..1000", INMSOM_BT_EINVAL);
    if (c->nx < 5 || c->ny < 5) return vfail("vertical_velocity needs nx, ny >= 5", INMSOM_BT_EINVAL);
    return INMSOM_BT_OK;
}

// staging: u, v (nz), w (nz + 1), hu, hv, wflux, dz
size_t stage_need(const inmsom_bt_ctx *c, int nz) { return ((size_t)3 * nz + 4) * c->n + (size_t)nz; }

VvParams staged(inmsom_bt_ctx *c, int nz)
{
    const size_t pl = c->n;
    VvParams p{};
    p.g = c->g;
    p.nz = nz;
    double *b = c->vv;
    p.u = b;
    p.v = b + (size_t)nz * pl;
    p.w = b + (size_t)2 * nz * pl;
    p.hu = b + (size_t)(3 * nz + 1) * pl;
    p.hv = p.hu + pl;
    p.wflux = p.hv + pl;
    p.dz = p.wflux + pl;
    return p;
}

}  // namespace

}  // namespace ibt

using namespace ibt;

extern "C" {

int inmsom_bt_vertical_velocity(inmsom_bt_ctx *c, int nz, const double *dz, const double *u, const double *v,
                                double *w, const double *hu, const double *hv, const double *wflux)
{
    int rc = check(c, nz);
    if (rc) return rc;
    if (!dz || !u || !v || !w || !hu || !hv || !wflux) return vfail("null pointer", INMSOM_BT_EINVAL);
    CHV(hipSetDevice(c->device));
    const size_t need = stage_need(c, nz);
    if (!c->vv || c->vv_n < need) {
        if (c->vv) CHV(hipFree(c->vv));
        c->vv = nullptr;
        c->vv_n = 0;
        c->vv_nz = 0;
        CHV(hipMalloc(&c->vv, need * sizeof(double)));
        c->vv_n = need;
    }
    const VvParams p = staged(c, nz);
    const size_t pl = c->n;
    auto put = [&](const double *dst, const double *src, size_t count) {
        return hipMemcpyAsync(const_cast<double *>(dst), src, count * sizeof(double), hipMemcpyHostToDevice,
                              c->stream);
    };
    CHV(put(p.u, u, (size_t)nz * pl));
    CHV(put(p.v, v, (size_t)nz * pl));
    CHV(put(p.w, w, (size_t)(nz + 1) * pl));
    CHV(put(p.hu, hu, pl));
    CHV(put(p.hv, hv, pl));
    CHV(put(p.wflux, wflux, pl));
    CHV(put(p.dz, dz, nz));
    launch(c, p);
    CHV(hipGetLastError());
    CHV(hipMemcpyAsync(w, p.w, (size_t)(nz + 1) * pl * sizeof(double), hipMemcpyDeviceToHost, c->stream));
    CHV(hipStreamSynchronize(c->stream));
    c->vv_nz = nz;
    return INMSOM_BT_OK;
}

int inmsom_bt_vertical_velocity_dev(inmsom_bt_ctx *c, int nz, const double *dz, const double *u, const double *v,
                                    double *w, const double *hu, const double *hv, const double *wflux)
{
    int rc = check(c, nz);
    if (rc) return rc;
    if (!dz || !u || !v || !w || !hu || !hv || !wflux) return vfail("null pointer", INMSOM_BT_EINVAL);
    CHV(hipSetDevice(c->device));
    if (!c->vv_dz) CHV(hipMalloc(&c->vv_dz, 1000 * sizeof(double)));
    CHV(hipMemcpyAsync(c->vv_dz, dz, nz * sizeof(double), hipMemcpyHostToDevice, c->stream));
    VvParams p{};
    p.g = c->g;
    p.nz = nz;
    p.dz = c->vv_dz;
    p.u = u;
    p.v = v;
    p.w = w;
    p.hu = hu;
    p.hv = hv;
    p.wflux = wflux;
    launch(c, p);
    CHV(hipGetLastError());
    CHV(hipStreamSynchronize(c->stream));
    return INMSOM_BT_OK;
}

int inmsom_bt_vertical_velocity_timed(inmsom_bt_ctx *c, int nz, int nrep, double *ms)
{
    if (!c || !ms) return vfail("null pointer", INMSOM_BT_EINVAL);
    if (nrep < 1) return vfail("nrep must be >= 1", INMSOM_BT_EINVAL);
    if (!c->vv || c->vv_nz != nz)
        return vfail("no fields staged for nz levels: call inmsom_bt_vertical_velocity first", INMSOM_BT_ESTATE);
    CHV(hipSetDevice(c->device));
    const VvParams p = staged(c, nz);
    hipEvent_t e0, e1;
    CHV(hipEventCreate(&e0));
    CHV(hipEventCreate(&e1));
    CHV(hipEventRecord(e0, c->stream));
    for (int r = 0; r < nrep; ++r) launch(c, p);
    CHV(hipGetLastError());
    CHV(hipEventRecord(e1, c->stream));
    CHV(hipEventSynchronize(e1));
    float t = 0.0f;
    CHV(hipEventElapsedTime(&t, e0, e1));
    CHV(hipEventDestroy(e0));
    CHV(hipEventDestroy(e1));
    *ms = t;
    return INMSOM_BT_OK;
}

}  // extern "C"
