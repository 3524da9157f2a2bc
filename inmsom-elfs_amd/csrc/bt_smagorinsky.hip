// Lateral Smagorinsky viscosity of the 3-D model and its depth integral,
// the pair Control/dyncall.f90:40-43 runs when ksw_lat > 0:
// smagorinsky_coeff (Function/mixing.f90:171-219) and then
// depth_ave(amuv, amuv2d, lu, 0) (Function/depth.f90:223-254).  amuv2d is
// the barotropic step's lateral viscosity (mu of barotropic_dynamics), so
// this producer feeds the step directly.
//
// On lu points of the interior (3..nx-2 x 3..ny-2) and every level k
//   mu   = (kappa / pi)^2 dx dy sqrt(str_t^2 + (str_s^2 summed over the
//          four H corners of the T cell) / 4)      kappa = 0.4
//   amuv = visc2 + mu,   amts = dift + 0.1 mu,
// in the reference's operation order (no fused multiply-adds), so the
// result is bit for bit the reference's.  Both fields are then wrapped on
// periodic boundaries (Service/oc_algorithms.f90:50-95, :118-135; x before
// y, which fixes the corners) and amuv2d = sum_k amuv dz (k ascending,
// from 0) on lu points of 2..nx-1 x 2..ny-1.
//
// k_smag: one thread per interior (m, n), the column walked in registers,
// 64 x 4 tiles dealt to the XCDs in contiguous runs; consecutive lanes take
// consecutive m, so each level's loads and stores are coalesced and the
// three neighbouring str_s values come from the cache lines the tile's
// other lanes and waves load.  Interior points are never touched by the
// wraps, so the thread adds up its own column's amuv2d as it goes and
// nothing is read back.  k_smag_ring sums the one-cell ring of the
// interior after the wraps (values copied there, or the caller's).
// Cells the reference does not compute keep the caller's values.
//
// stress_components (Function/mixing.f90:115-168) of the 3-D velocities,
// the input of the pair (Control/dyncall.f90:36): tension on lu points and
// shear on luu points of the interior, every level, in the reference's
// operation order, then both wrapped on periodic boundaries.  k_stress3
// walks each interior column like k_smag; the per-point metric ratios are
// the same at every level and formed once.
#include <string>

#include "bt_internal.h"

namespace ibt {

namespace {

struct SmParams {
    DevGrid g;
    int nz;
    int tiles_x, tiles_y;  // 64 x 4 tiles over the interior
    double visc2, dift;
    const double *dz;      // nz (device)
    const double *str_t, *str_s;  // (nz, ny, nx)
    double *amts, *amuv;   // (nz, ny, nx)
    double *amuv2d;        // (ny, nx)
};

#define I(mm, nn) IDX(g, (mm), (nn))

__global__ __launch_bounds__(256) void k_smag(SmParams p)
{
    const DevGrid &g = p.g;
    int bx, by;
    xcd_tile(p.tiles_x, p.tiles_y, (int)blockIdx.x, bx, by);
    const int m = bx * 64 + (int)(threadIdx.x & 63) + 3;
    const int n = by * 4 + (int)(threadIdx.x >> 6) + 3;
    if (m > g.nx - 2 || n > g.ny - 2) return;
    const size_t pl = (size_t)g.nx * g.ny, i = I(m, n), iw = I(m - 1, n), is = I(m, n - 1),
                 isw = I(m - 1, n - 1);
    if (!(g.lu[i] > 0.5f)) return;
    const double *__restrict__ st = p.str_t;
    const double *__restrict__ ss = p.str_s;
    const double *__restrict__ dz = p.dz;
    double *__restrict__ uv = p.amuv;
    double *__restrict__ ts = p.amts;
    // (kappa / pi)^2: 0.4 / pi squared as x * x (Inc/basinpar.fi:72,89)
    const double kp = 0.4 / 3.141592653589793;
    const double cdxy = kp * kp * g.dx[i] * g.dy[i];
    const double visc2 = p.visc2, dift = p.dift;
    double acc = 0.0;
#pragma unroll 4
    for (int k = 0; k < p.nz; ++k) {
        const size_t o = (size_t)k * pl;
        const double a = ss[o + i], b = ss[o + iw], c = ss[o + is], d = ss[o + isw];
        const double t = st[o + i];
        const double s2 = (a * a + b * b + c * c + d * d) / 4.0;
        const double mu = cdxy * sqrt(t * t + s2);
        const double v = visc2 + mu;
        uv[o + i] = v;
        ts[o + i] = dift + mu * 0.1;
        acc = acc + v * dz[k];
    }
    p.amuv2d[i] = acc;
}

struct StParams {
    DevGrid g;
    int nz;
    int tiles_x, tiles_y;
    const double *u, *v;      // (nz, ny, nx)
    double *str_t, *str_s;    // (nz, ny, nx)
};

__global__ __launch_bounds__(256) void k_stress3(StParams p)
{
    const DevGrid &g = p.g;
    int bx, by;
    xcd_tile(p.tiles_x, p.tiles_y, (int)blockIdx.x, bx, by);
    const int m = bx * 64 + (int)(threadIdx.x & 63) + 3;
    const int n = by * 4 + (int)(threadIdx.x >> 6) + 3;
    if (m > g.nx - 2 || n > g.ny - 2) return;
    const size_t pl = (size_t)g.nx * g.ny, i = I(m, n), iw = I(m - 1, n), is = I(m, n - 1), in = I(m, n + 1),
                 ie = I(m + 1, n);
    const bool tw = g.lu[i] > 0.5f, sw = g.luu[i] > 0.5f;
    if (!tw && !sw) return;
    const double *__restrict__ u = p.u;
    const double *__restrict__ v = p.v;
    double *__restrict__ st = p.str_t;
    double *__restrict__ ss = p.str_s;
    // tension: dy/dx (u/dyh - u_w/dyh_w) - dx/dy (v/dxh - v_s/dxh_s)
    const double ryx = g.dy[i] / g.dx[i], rxy = g.dx[i] / g.dy[i];
    const double dyh0 = g.dyh[i], dyhw = g.dyh[iw], dxh0 = g.dxh[i], dxhs = g.dxh[is];
    // shear: dxb/dyb (u_n/dxt_n - u/dxt) + dyb/dxb (v_e/dyt_e - v/dyt)
    const double rbxy = g.dxb[i] / g.dyb[i], rbyx = g.dyb[i] / g.dxb[i];
    const double dxtn = g.dxt[in], dxt0 = g.dxt[i], dyte = g.dyt[ie], dyt0 = g.dyt[i];
#pragma unroll 2
    for (int k = 0; k < p.nz; ++k) {
        const size_t o = (size_t)k * pl;
        const double u0 = u[o + i], v0 = v[o + i];
        if (tw) st[o + i] = ryx * (u0 / dyh0 - u[o + iw] / dyhw) - rxy * (v0 / dxh0 - v[o + is] / dxhs);
        if (sw) ss[o + i] = rbxy * (u[o + in] / dxtn - u0 / dxt0) + rbyx * (v[o + ie] / dyte - v0 / dyt0);
    }
}

// column mm -> 2, 3 -> nx - 1 (x); row nn -> 2, 3 -> ny - 1 (y); both
// fields, all levels
__global__ void k_smag_wrap_x(DevGrid g, int nz, double *f0, double *f1)
{
    const int n = blockIdx.x * blockDim.x + threadIdx.x + 1;
    const int k = blockIdx.y;
    if (n > g.ny || k >= 2 * nz) return;
    double *f = (k < nz ? f0 : f1) + (size_t)(k % nz) * g.nx * g.ny;
    const double a = f[I(g.nx - 2, n)], b = f[I(3, n)];
    f[I(2, n)] = a;
    f[I(g.nx - 1, n)] = b;
}

__global__ void k_smag_wrap_y(DevGrid g, int nz, double *f0, double *f1)
{
    const int m = blockIdx.x * blockDim.x + threadIdx.x + 1;
    const int k = blockIdx.y;
    if (m > g.nx || k >= 2 * nz) return;
    double *f = (k < nz ? f0 : f1) + (size_t)(k % nz) * g.nx * g.ny;
    const double a = f[I(m, g.ny - 2)], b = f[I(m, 3)];
    f[I(m, 2)] = a;
    f[I(m, g.ny - 1)] = b;
}

// The ring of the interior: rows 2 and ny - 1 (m = 2..nx-1), then columns
// 2 and nx - 1 (n = 3..ny-2); depth sum of amuv on its lu cells.
__global__ void k_smag_ring(SmParams p)
{
    const DevGrid &g = p.g;
    const int t = blockIdx.x * blockDim.x + threadIdx.x;
    const int row = g.nx - 2, col = g.ny - 4;
    int m, n;
    if (t < 2 * row) {
        m = 2 + t % row;
        n = t < row ? 2 : g.ny - 1;
    } else if (t < 2 * row + 2 * col) {
        const int u = t - 2 * row;
        n = 3 + u % col;
        m = u < col ? 2 : g.nx - 1;
    } else {
        return;
    }
    const size_t pl = (size_t)g.nx * g.ny, i = I(m, n);
    if (!(g.lu[i] > 0.5f)) return;
    double acc = 0.0;
    for (int k = 0; k < p.nz; ++k) acc = acc + p.amuv[(size_t)k * pl + i] * p.dz[k];
    p.amuv2d[i] = acc;
}

#undef I

int sfail(const std::string &m, int code)
{
    set_error(m);
    return code;
}

#define CHS(expr)                                                                                   \
    do {                                                                                            \
        hipError_t e_ = (expr);                                                                     \
        if (e_ != hipSuccess) return sfail(std::string(#expr) + ": " + hipGetErrorString(e_),        \
                                           INMSOM_BT_EDEVICE);                                      \
    } while (0)

void launch(inmsom_bt_ctx *c, SmParams p)
{
    p.tiles_x = (c->nx - 4 + 63) / 64;
    p.tiles_y = (c->ny - 4 + 3) / 4;
    hipLaunchKernelGGL(k_smag, dim3((unsigned)(p.tiles_x * p.tiles_y)), dim3(256), 0, c->stream, p);
    if (c->g.px)
        hipLaunchKernelGGL(k_smag_wrap_x, dim3((unsigned)((c->ny + 255) / 256), (unsigned)(2 * p.nz)), dim3(256), 0,
                           c->stream, c->g, p.nz, p.amuv, p.amts);
    if (c->g.py)
        hipLaunchKernelGGL(k_smag_wrap_y, dim3((unsigned)((c->nx + 255) / 256), (unsigned)(2 * p.nz)), dim3(256), 0,
                           c->stream, c->g, p.nz, p.amuv, p.amts);
    const int ring = 2 * (c->nx - 2) + 2 * (c->ny - 4);
    hipLaunchKernelGGL(k_smag_ring, dim3((unsigned)((ring + 255) / 256)), dim3(256), 0, c->stream, p);
}

void launch_stress(inmsom_bt_ctx *c, StParams p)
{
    p.tiles_x = (c->nx - 4 + 63) / 64;
    p.tiles_y = (c->ny - 4 + 3) / 4;
    hipLaunchKernelGGL(k_stress3, dim3((unsigned)(p.tiles_x * p.tiles_y)), dim3(256), 0, c->stream, p);
    if (c->g.px)
        hipLaunchKernelGGL(k_smag_wrap_x, dim3((unsigned)((c->ny + 255) / 256), (unsigned)(2 * p.nz)), dim3(256), 0,
                           c->stream, c->g, p.nz, p.str_t, p.str_s);
    if (c->g.py)
        hipLaunchKernelGGL(k_smag_wrap_y, dim3((unsigned)((c->nx + 255) / 256), (unsigned)(2 * p.nz)), dim3(256), 0,
                           c->stream, c->g, p.nz, p.str_t, p.str_s);
}

// staging of stress_components: u, v, str_t, str_s (nz planes each)
StParams staged_stress(inmsom_bt_ctx *c, int nz)
{
    const size_t vol = (size_t)nz * c->n;
    StParams p{};
    p.g = c->g;
    p.nz = nz;
    p.u = c->sst;
    p.v = c->sst + vol;
    p.str_t = c->sst + 2 * vol;
    p.str_s = c->sst + 3 * vol;
    return p;
}

int check(inmsom_bt_ctx *c, int nz)
{
    if (!c) return sfail("null pointer", INMSOM_BT_EINVAL);
    if (nz < 1 || nz > 1000) return sfail("nz must be in 1..1000", INMSOM_BT_EINVAL);
    if (c->nx < 5 || c->ny < 5) return sfail("smagorinsky needs nx, ny >= 5", INMSOM_BT_EINVAL);
    return INMSOM_BT_OK;
}

// staging: str_t, str_s, amts, amuv (nz), amuv2d, dz
size_t stage_need(const inmsom_bt_ctx *c, int nz) { return ((size_t)4 * nz + 1) * c->n + (size_t)nz; }

SmParams staged(inmsom_bt_ctx *c, int nz, double visc2, double dift)
{
    const size_t pl = c->n;
    SmParams p{};
    p.g = c->g;
    p.nz = nz;
    p.visc2 = visc2;
    p.dift = dift;
    double *b = c->sm;
    p.str_t = b;
    p.str_s = b + (size_t)nz * pl;
    p.amts = b + (size_t)2 * nz * pl;
    p.amuv = b + (size_t)3 * nz * pl;
    p.amuv2d = b + (size_t)4 * nz * pl;
    p.dz = p.amuv2d + pl;
    return p;
}

}  // namespace

}  // namespace ibt

using namespace ibt;

extern "C" {

int inmsom_bt_smagorinsky(inmsom_bt_ctx *c, int nz, double visc2, double dift, const double *str_t,
                          const double *str_s, const double *dz, double *amts, double *amuv, double *amuv2d)
{
    int rc = check(c, nz);
    if (rc) return rc;
    if (!str_t || !str_s || !dz || !amts || !amuv || !amuv2d) return sfail("null pointer", INMSOM_BT_EINVAL);
    CHS(hipSetDevice(c->device));
    const size_t need = stage_need(c, nz);
    if (!c->sm || c->sm_n < need) {
        if (c->sm) CHS(hipFree(c->sm));
        c->sm = nullptr;
        c->sm_n = 0;
        c->sm_nz = 0;
        CHS(hipMalloc(&c->sm, need * sizeof(double)));
        c->sm_n = need;
    }
    const SmParams p = staged(c, nz, visc2, dift);
    const size_t pl = c->n, vol = (size_t)nz * pl;
    auto put = [&](const double *dst, const double *src, size_t count) {
        return hipMemcpyAsync(const_cast<double *>(dst), src, count * sizeof(double), hipMemcpyHostToDevice,
                              c->stream);
    };
    CHS(put(p.str_t, str_t, vol));
    CHS(put(p.str_s, str_s, vol));
    CHS(put(p.amts, amts, vol));
    CHS(put(p.amuv, amuv, vol));
    CHS(put(p.amuv2d, amuv2d, pl));
    CHS(put(p.dz, dz, nz));
    launch(c, p);
    CHS(hipGetLastError());
    CHS(hipMemcpyAsync(amts, p.amts, vol * sizeof(double), hipMemcpyDeviceToHost, c->stream));
    CHS(hipMemcpyAsync(amuv, p.amuv, vol * sizeof(double), hipMemcpyDeviceToHost, c->stream));
    CHS(hipMemcpyAsync(amuv2d, p.amuv2d, pl * sizeof(double), hipMemcpyDeviceToHost, c->stream));
    CHS(hipStreamSynchronize(c->stream));
    c->sm_nz = nz;
    c->sm_visc2 = visc2;
    c->sm_dift = dift;
    return INMSOM_BT_OK;
}

int inmsom_bt_smagorinsky_dev(inmsom_bt_ctx *c, int nz, double visc2, double dift, const double *str_t,
                              const double *str_s, const double *dz, double *amts, double *amuv, double *amuv2d)
{
    int rc = check(c, nz);
    if (rc) return rc;
    if (!str_t || !str_s || !dz || !amts || !amuv || !amuv2d) return sfail("null pointer", INMSOM_BT_EINVAL);
    CHS(hipSetDevice(c->device));
    if (!c->sm_dz) CHS(hipMalloc(&c->sm_dz, 1000 * sizeof(double)));
    CHS(hipMemcpyAsync(c->sm_dz, dz, nz * sizeof(double), hipMemcpyHostToDevice, c->stream));
    SmParams p{};
    p.g = c->g;
    p.nz = nz;
    p.visc2 = visc2;
    p.dift = dift;
    p.dz = c->sm_dz;
    p.str_t = str_t;
    p.str_s = str_s;
    p.amts = amts;
    p.amuv = amuv;
    p.amuv2d = amuv2d;
    launch(c, p);
    CHS(hipGetLastError());
    CHS(hipStreamSynchronize(c->stream));
    return INMSOM_BT_OK;
}

int inmsom_bt_stress_components(inmsom_bt_ctx *c, int nz, const double *u, const double *v, double *str_t,
                                double *str_s)
{
    int rc = check(c, nz);
    if (rc) return rc;
    if (!u || !v || !str_t || !str_s) return sfail("null pointer", INMSOM_BT_EINVAL);
    CHS(hipSetDevice(c->device));
    const size_t vol = (size_t)nz * c->n, need = 4 * vol;
    if (!c->sst || c->sst_n < need) {
        if (c->sst) CHS(hipFree(c->sst));
        c->sst = nullptr;
        c->sst_n = 0;
        c->sst_nz = 0;
        CHS(hipMalloc(&c->sst, need * sizeof(double)));
        c->sst_n = need;
    }
    const StParams p = staged_stress(c, nz);
    const size_t bytes = vol * sizeof(double);
    CHS(hipMemcpyAsync(const_cast<double *>(p.u), u, bytes, hipMemcpyHostToDevice, c->stream));
    CHS(hipMemcpyAsync(const_cast<double *>(p.v), v, bytes, hipMemcpyHostToDevice, c->stream));
    CHS(hipMemcpyAsync(p.str_t, str_t, bytes, hipMemcpyHostToDevice, c->stream));
    CHS(hipMemcpyAsync(p.str_s, str_s, bytes, hipMemcpyHostToDevice, c->stream));
    launch_stress(c, p);
    CHS(hipGetLastError());
    CHS(hipMemcpyAsync(str_t, p.str_t, bytes, hipMemcpyDeviceToHost, c->stream));
    CHS(hipMemcpyAsync(str_s, p.str_s, bytes, hipMemcpyDeviceToHost, c->stream));
    CHS(hipStreamSynchronize(c->stream));
    c->sst_nz = nz;
    return INMSOM_BT_OK;
}

int inmsom_bt_stress_components_dev(inmsom_bt_ctx *c, int nz, const double *u, const double *v, double *str_t,
                                    double *str_s)
{
    int rc = check(c, nz);
    if (rc) return rc;
    if (!u || !v || !str_t || !str_s) return sfail("null pointer", INMSOM_BT_EINVAL);
    CHS(hipSetDevice(c->device));
    StParams p{};
    p.g = c->g;
    p.nz = nz;
    p.u = u;
    p.v = v;
    p.str_t = str_t;
    p.str_s = str_s;
    launch_stress(c, p);
    CHS(hipGetLastError());
    CHS(hipStreamSynchronize(c->stream));
    return INMSOM_BT_OK;
}

int inmsom_bt_stress_components_timed(inmsom_bt_ctx *c, int nz, int nrep, double *ms)
{
    if (!c || !ms) return sfail("null pointer", INMSOM_BT_EINVAL);
    if (nrep < 1) return sfail("nrep must be >= 1", INMSOM_BT_EINVAL);
    if (!c->sst || c->sst_nz != nz)
        return sfail("no fields staged for nz levels: call inmsom_bt_stress_components first", INMSOM_BT_ESTATE);
    CHS(hipSetDevice(c->device));
    const StParams p = staged_stress(c, nz);
    hipEvent_t e0, e1;
    CHS(hipEventCreate(&e0));
    CHS(hipEventCreate(&e1));
    CHS(hipEventRecord(e0, c->stream));
    for (int r = 0; r < nrep; ++r) launch_stress(c, p);
    CHS(hipGetLastError());
    CHS(hipEventRecord(e1, c->stream));
    CHS(hipEventSynchronize(e1));
    float t = 0.0f;
    CHS(hipEventElapsedTime(&t, e0, e1));
    CHS(hipEventDestroy(e0));
    CHS(hipEventDestroy(e1));
    *ms = t;
    return INMSOM_BT_OK;
}

int inmsom_bt_smagorinsky_timed(inmsom_bt_ctx *c, int nz, int nrep, double *ms)
{
    if (!c || !ms) return sfail("null pointer", INMSOM_BT_EINVAL);
    if (nrep < 1) return sfail("nrep must be >= 1", INMSOM_BT_EINVAL);
    if (!c->sm || c->sm_nz != nz)
        return sfail("no fields staged for nz levels: call inmsom_bt_smagorinsky first", INMSOM_BT_ESTATE);
    CHS(hipSetDevice(c->device));
    const SmParams p = staged(c, nz, c->sm_visc2, c->sm_dift);
    hipEvent_t e0, e1;
    CHS(hipEventCreate(&e0));
    CHS(hipEventCreate(&e1));
    CHS(hipEventRecord(e0, c->stream));
    for (int r = 0; r < nrep; ++r) launch(c, p);
    CHS(hipGetLastError());
    CHS(hipEventRecord(e1, c->stream));
    CHS(hipEventSynchronize(e1));
    float t = 0.0f;
    CHS(hipEventElapsedTime(&t, e0, e1));
    CHS(hipEventDestroy(e0));
    CHS(hipEventDestroy(e1));
    *ms = t;
    return INMSOM_BT_OK;
}

}  // extern "C"
