// pressure_gradients (Function/vel_ssh.f90:320-376): the hydrostatic
// pressure of the density integrated down the sigma levels between the two
// T columns of each lcu (x) and lcv (y) interior point, its gradient
// subtracted from the 3-D right-hand sides in place.  Cells off lcu (lcv)
// keep the caller's values; the reference does not exchange halos after
// it, and neither does this.
//
// k_pressure: one thread per interior (m, n) carrying both running
// integrals (x and y) down its column in registers, in 64 x 4 tiles dealt
// to the XCDs in contiguous runs.  Consecutive lanes take consecutive m, so
// each level's loads and stores are coalesced; the east and north densities
// are the neighbours' own loads and come from cache.  The arithmetic is the
// reference's, term by term and in its order, so the results are
// bit-identical.
#include <string>

#include "bt_internal.h"

namespace ibt {

namespace {

struct PgParams {
    DevGrid g;
    int nz;
    int tiles_x, tiles_y;        // 64 x 4 tiles over the interior
    const double *z;             // nz sigma levels (device)
    const double *den;           // (nz, ny, nx)
    const double *hhq, *hhu, *hhv;
    double *rhsx, *rhsy;         // (nz, ny, nx), updated in place
};

#define I(mm, nn) IDX(g, (mm), (nn))

__global__ __launch_bounds__(256) void k_pressure(PgParams p)
{
    const DevGrid &g = p.g;
    int bx, by;
    xcd_tile(p.tiles_x, p.tiles_y, (int)blockIdx.x, bx, by);
    const int m = bx * 64 + (int)(threadIdx.x & 63) + 3;
    const int n = by * 4 + (int)(threadIdx.x >> 6) + 3;
    if (m > g.nx - 2 || n > g.ny - 2) return;
    const size_t pl = (size_t)g.nx * g.ny, i = I(m, n), ie = I(m + 1, n), in = I(m, n + 1);
    const bool cu = g.lcu[i] > 0.5f, cv = g.lcv[i] > 0.5f;
    if (!cu && !cv) return;
    const double *__restrict__ den = p.den;
    const double *__restrict__ z = p.z;
    double *__restrict__ rx = p.rhsx;
    double *__restrict__ ry = p.rhsy;
    const double gr = g.g, rho = g.rho0;
    // column i's depth and the east (x) / north (y) neighbour's
    const double h0 = p.hhq[i], hx = cu ? p.hhq[ie] : 0.0, hy = cv ? p.hhq[in] : 0.0;
    const double fx = cu ? p.hhu[i] : 0.0, mx = cu ? g.dyh[i] : 0.0;
    const double fy = cv ? p.hhv[i] : 0.0, my = cv ? g.dxh[i] : 0.0;
    // level 1
    double d0 = den[i], dx = cu ? den[ie] : 0.0, dy = cv ? den[in] : 0.0;
    double px = 0.0, py = 0.0;  // p_hint of the x and y integrals
    {
        const double z1 = z[0];
        if (cu) {
            px = hx * z1 * dx - h0 * z1 * d0;
            const double lat = hx * z1 * d0 - h0 * z1 * dx;
            rx[i] = rx[i] - (px - lat) * fx * mx * gr / rho / 2.0;
        }
        if (cv) {
            py = hy * z1 * dy - h0 * z1 * d0;
            const double lat = hy * z1 * d0 - h0 * z1 * dy;
            ry[i] = ry[i] - (py - lat) * fy * my * gr / rho / 2.0;
        }
    }
    // levels 2..nz: the previous level's densities stay in registers
#pragma unroll 4
    for (int k = 1; k < p.nz; ++k) {
        const size_t o = (size_t)k * pl;
        const double zk = z[k], zp = z[k - 1];
        const double e0 = den[o + i];
        if (cu) {
            const double e1 = den[o + ie];
            px = px + (hx * zk * dx - hx * zp * e1) - (h0 * zk * d0 - h0 * zp * e0);
            const double lat = hx * zk * e0 - h0 * zk * e1;
            rx[o + i] = rx[o + i] - (px - lat) * fx * mx * gr / rho / 2.0;
            dx = e1;
        }
        if (cv) {
            const double e1 = den[o + in];
            py = py + (hy * zk * dy - hy * zp * e1) - (h0 * zk * d0 - h0 * zp * e0);
            const double lat = hy * zk * e0 - h0 * zk * e1;
            ry[o + i] = ry[o + i] - (py - lat) * fy * my * gr / rho / 2.0;
            dy = e1;
        }
        d0 = e0;
    }
}

#undef I

int pfail(const std::string &m, int code)
{
    set_error(m);
    return code;
}

#define CHP(expr)                                                                                   \
    do {                                                                                            \
        hipError_t e_ = (expr);                                                                     \
        if (e_ != hipSuccess) return pfail(std::string(#expr) + ": " + hipGetErrorString(e_),        \
                                           INMSOM_BT_EDEVICE);                                      \
    } while (0)

void launch(inmsom_bt_ctx *c, PgParams p)
{
    p.tiles_x = (c->nx - 4 + 63) / 64;
    p.tiles_y = (c->ny - 4 + 3) / 4;
    hipLaunchKernelGGL(k_pressure, dim3((unsigned)(p.tiles_x * p.tiles_y)), dim3(256), 0, c->stream, p);
}

int check(inmsom_bt_ctx *c, int nz)
{
    if (!c) return pfail("null pointer", INMSOM_BT_EINVAL);
    if (nz < 1 || nz > 1000) return pfail("nz must be in 1..1000", INMSOM_BT_EINVAL);
    if (c->nx < 5 || c->ny < 5) return pfail("pressure_gradients needs nx, ny >= 5", INMSOM_BT_EINVAL);
    return INMSOM_BT_OK;
}

// staging: den, rhsx, rhsy (nz each), hhq, hhu, hhv, z
size_t stage_need(const inmsom_bt_ctx *c, int nz) { return ((size_t)3 * nz + 3) * c->n + (size_t)nz; }

PgParams staged(inmsom_bt_ctx *c, int nz)
{
    const size_t pl = c->n;
    PgParams p{};
    p.g = c->g;
    p.nz = nz;
    double *b = c->pg;
    p.den = b;
    p.rhsx = b + (size_t)nz * pl;
    p.rhsy = b + (size_t)2 * nz * pl;
    p.hhq = b + (size_t)3 * nz * pl;
    p.hhu = p.hhq + pl;
    p.hhv = p.hhu + pl;
    p.z = p.hhv + pl;
    return p;
}

}  // namespace

}  // namespace ibt

using namespace ibt;

extern "C" {

int inmsom_bt_pressure_gradients(inmsom_bt_ctx *c, int nz, const double *z, const double *den, const double *hhq,
                                 const double *hhu, const double *hhv, double *rhsx, double *rhsy)
{
    int rc = check(c, nz);
    if (rc) return rc;
    if (!z || !den || !hhq || !hhu || !hhv || !rhsx || !rhsy) return pfail("null pointer", INMSOM_BT_EINVAL);
    CHP(hipSetDevice(c->device));
    const size_t need = stage_need(c, nz);
    if (!c->pg || c->pg_n < need) {
        if (c->pg) CHP(hipFree(c->pg));
        c->pg = nullptr;
        c->pg_n = 0;
        c->pg_nz = 0;
        CHP(hipMalloc(&c->pg, need * sizeof(double)));
        c->pg_n = need;
    }
    const PgParams p = staged(c, nz);
    const size_t pl = c->n;
    auto put = [&](const double *dst, const double *src, size_t count) {
        return hipMemcpyAsync(const_cast<double *>(dst), src, count * sizeof(double), hipMemcpyHostToDevice,
                              c->stream);
    };
    CHP(put(p.den, den, (size_t)nz * pl));
    CHP(put(p.rhsx, rhsx, (size_t)nz * pl));
    CHP(put(p.rhsy, rhsy, (size_t)nz * pl));
    CHP(put(p.hhq, hhq, pl));
    CHP(put(p.hhu, hhu, pl));
    CHP(put(p.hhv, hhv, pl));
    CHP(put(p.z, z, nz));
    launch(c, p);
    CHP(hipGetLastError());
    CHP(hipMemcpyAsync(rhsx, p.rhsx, (size_t)nz * pl * sizeof(double), hipMemcpyDeviceToHost, c->stream));
    CHP(hipMemcpyAsync(rhsy, p.rhsy, (size_t)nz * pl * sizeof(double), hipMemcpyDeviceToHost, c->stream));
    CHP(hipStreamSynchronize(c->stream));
    c->pg_nz = nz;
    return INMSOM_BT_OK;
}

int inmsom_bt_pressure_gradients_dev(inmsom_bt_ctx *c, int nz, const double *z, const double *den,
                                     const double *hhq, const double *hhu, const double *hhv, double *rhsx,
                                     double *rhsy)
{
    int rc = check(c, nz);
    if (rc) return rc;
    if (!z || !den || !hhq || !hhu || !hhv || !rhsx || !rhsy) return pfail("null pointer", INMSOM_BT_EINVAL);
    CHP(hipSetDevice(c->device));
    if (!c->pg_z) CHP(hipMalloc(&c->pg_z, 1000 * sizeof(double)));
    CHP(hipMemcpyAsync(c->pg_z, z, nz * sizeof(double), hipMemcpyHostToDevice, c->stream));
    PgParams p{};
    p.g = c->g;
    p.nz = nz;
    p.z = c->pg_z;
    p.den = den;
    p.hhq = hhq;
    p.hhu = hhu;
    p.hhv = hhv;
    p.rhsx = rhsx;
    p.rhsy = rhsy;
    launch(c, p);
    CHP(hipGetLastError());
    CHP(hipStreamSynchronize(c->stream));
    return INMSOM_BT_OK;
}

int inmsom_bt_pressure_gradients_timed(inmsom_bt_ctx *c, int nz, int nrep, double *ms)
{
    if (!c || !ms) return pfail("null pointer", INMSOM_BT_EINVAL);
    if (nrep < 1) return pfail("nrep must be >= 1", INMSOM_BT_EINVAL);
    if (!c->pg || c->pg_nz != nz)
        return pfail("no fields staged for nz levels: call inmsom_bt_pressure_gradients first", INMSOM_BT_ESTATE);
    CHP(hipSetDevice(c->device));
    const PgParams p = staged(c, nz);
    hipEvent_t e0, e1;
    CHP(hipEventCreate(&e0));
    CHP(hipEventCreate(&e1));
    CHP(hipEventRecord(e0, c->stream));
    for (int r = 0; r < nrep; ++r) launch(c, p);
    CHP(hipGetLastError());
    CHP(hipEventRecord(e1, c->stream));
    CHP(hipEventSynchronize(e1));
    float t = 0.0f;
    CHP(hipEventElapsedTime(&t, e0, e1));
    CHP(hipEventDestroy(e0));
    CHP(hipEventDestroy(e1));
    *ms = t;
    return INMSOM_BT_OK;
}

}  // extern "C"
