// The density definition of the 3-D step (Control/dyncall.f90:437-449):
// on lu points of m = 2..nx-1, n = 2..ny-1 and every level, the in-situ
// density deviation den at the pressure FreeFallAcc * RefDen * hhq * z(k)
// and the potential one den_pot at zero pressure, from the potential
// temperature tt and salinity ss, by the reference's polynomial equation of
// state (Service/oc_algorithms.f90:138-161, Brydon et al. 1999).  den feeds
// pressure_gradients.  Cells the reference does not write keep the
// caller's values.
//
// k_density: one thread per (m, n), the column walked level by level;
// consecutive lanes take consecutive m, so every level's loads and stores
// are coalesced; 64 x 4 tiles dealt to the XCDs in contiguous runs.  The
// polynomial is evaluated term by term in the reference's order, so the
// results are bit-identical.
#include <string>

#include "bt_internal.h"

namespace ibt {

namespace {

struct EosParams {
    DevGrid g;
    int nz;
    int tiles_x, tiles_y;        // 64 x 4 tiles over m, n = 2..nx-1, 2..ny-1
    const double *z;             // nz sigma levels (device)
    const double *tt, *ss;       // (nz, ny, nx)
    const double *hhq;
    double *den, *den_pot;       // (nz, ny, nx), written on lu
};

#define I(mm, nn) IDX(g, (mm), (nn))

// Service/oc_algorithms.f90:138-161: deviation from 1020 kg/m^3 at t [C],
// s [psu], p_pa [Pa]; the pure-pressure term is left out, as there
__device__ __forceinline__ double eos(double t, double s, double p_pa)
{
    const double p = p_pa * 1.0e-6;
    const double den_cgs = -2.00920601e-02 + (5.10768e-05 - 3.69119e-06 * p + 6.54837e-09 * p * p) * t +
                           (8.05999e-04 - 9.34012e-07 * p + 1.38777e-09 * p * p) * s +
                           (-7.40849e-06 + 5.33243e-08 * p - 1.01563e-10 * p * p) * t * t +
                           (-3.01036e-06 + 1.75145e-08 * p - 2.34892e-11 * p * p) * t * s +
                           (3.32267e-08 - 3.25887e-10 * p + 4.98612e-13 * p * p) * t * t * t +
                           (3.21931e-08 - 1.65849e-10 * p + 2.17612e-13 * p * p) * t * t * s;
    return den_cgs * 1000.0;
}

__global__ __launch_bounds__(256) void k_density(EosParams p)
{
    const DevGrid &g = p.g;
    int bx, by;
    xcd_tile(p.tiles_x, p.tiles_y, (int)blockIdx.x, bx, by);
    const int m = bx * 64 + (int)(threadIdx.x & 63) + 2;
    const int n = by * 4 + (int)(threadIdx.x >> 6) + 2;
    if (m > g.nx - 1 || n > g.ny - 1) return;
    const size_t pl = (size_t)g.nx * g.ny, i = I(m, n);
    if (!(g.lu[i] > 0.5f)) return;
    const double *__restrict__ tt = p.tt;
    const double *__restrict__ ss = p.ss;
    double *__restrict__ den = p.den;
    double *__restrict__ dpot = p.den_pot;
    const double gh = g.g * g.rho0 * p.hhq[i];
#pragma unroll 8
    for (int k = 0; k < p.nz; ++k) {
        const size_t o = (size_t)k * pl + i;
        const double t = tt[o], s = ss[o];
        den[o] = eos(t, s, gh * p.z[k]);
        dpot[o] = eos(t, s, 0.0);
    }
}

#undef I

int efail(const std::string &m, int code)
{
    set_error(m);
    return code;
}

#define CHE(expr)                                                                                   \
    do {                                                                                            \
        hipError_t e_ = (expr);                                                                     \
        if (e_ != hipSuccess) return efail(std::string(#expr) + ": " + hipGetErrorString(e_),        \
                                           INMSOM_BT_EDEVICE);                                      \
    } while (0)

void launch(inmsom_bt_ctx *c, EosParams p)
{
    p.tiles_x = (c->nx - 2 + 63) / 64;
    p.tiles_y = (c->ny - 2 + 3) / 4;
    hipLaunchKernelGGL(k_density, dim3((unsigned)(p.tiles_x * p.tiles_y)), dim3(256), 0, c->stream, p);
}

int check(inmsom_bt_ctx *c, int nz)
{
    if (!c) return efail("null pointer", INMSOM_BT_EINVAL);
    if (nz < 1 || nz > 1000) return efail("nz must be in 1..1000", INMSOM_BT_EINVAL);
    if (c->nx < 5 || c->ny < 5) return efail("density needs nx, ny >= 5", INMSOM_BT_EINVAL);
    return INMSOM_BT_OK;
}

// staging: tt, ss, den, den_pot (nz each), hhq, z
size_t stage_need(const inmsom_bt_ctx *c, int nz) { return ((size_t)4 * nz + 1) * c->n + (size_t)nz; }

EosParams staged(inmsom_bt_ctx *c, int nz)
{
    const size_t pl = c->n;
    EosParams p{};
    p.g = c->g;
    p.nz = nz;
    double *b = c->eos;
    p.tt = b;
    p.ss = b + (size_t)nz * pl;
    p.den = b + (size_t)2 * nz * pl;
    p.den_pot = b + (size_t)3 * nz * pl;
    p.hhq = b + (size_t)4 * nz * pl;
    p.z = p.hhq + pl;
    return p;
}

}  // namespace

}  // namespace ibt

using namespace ibt;

extern "C" {

int inmsom_bt_density(inmsom_bt_ctx *c, int nz, const double *z, const double *tt, const double *ss,
                      const double *hhq, double *den, double *den_pot)
{
    int rc = check(c, nz);
    if (rc) return rc;
    if (!z || !tt || !ss || !hhq || !den || !den_pot) return efail("null pointer", INMSOM_BT_EINVAL);
    CHE(hipSetDevice(c->device));
    const size_t need = stage_need(c, nz);
    if (!c->eos || c->eos_n < need) {
        if (c->eos) CHE(hipFree(c->eos));
        c->eos = nullptr;
        c->eos_n = 0;
        c->eos_nz = 0;
        CHE(hipMalloc(&c->eos, need * sizeof(double)));
        c->eos_n = need;
    }
    const EosParams p = staged(c, nz);
    const size_t pl = c->n;
    auto put = [&](const double *dst, const double *src, size_t count) {
        return hipMemcpyAsync(const_cast<double *>(dst), src, count * sizeof(double), hipMemcpyHostToDevice,
                              c->stream);
    };
    CHE(put(p.tt, tt, (size_t)nz * pl));
    CHE(put(p.ss, ss, (size_t)nz * pl));
    CHE(put(p.den, den, (size_t)nz * pl));
    CHE(put(p.den_pot, den_pot, (size_t)nz * pl));
    CHE(put(p.hhq, hhq, pl));
    CHE(put(p.z, z, nz));
    launch(c, p);
    CHE(hipGetLastError());
    CHE(hipMemcpyAsync(den, p.den, (size_t)nz * pl * sizeof(double), hipMemcpyDeviceToHost, c->stream));
    CHE(hipMemcpyAsync(den_pot, p.den_pot, (size_t)nz * pl * sizeof(double), hipMemcpyDeviceToHost, c->stream));
    CHE(hipStreamSynchronize(c->stream));
    c->eos_nz = nz;
    return INMSOM_BT_OK;
}

int inmsom_bt_density_dev(inmsom_bt_ctx *c, int nz, const double *z, const double *tt, const double *ss,
                          const double *hhq, double *den, double *den_pot)
{
    int rc = check(c, nz);
    if (rc) return rc;
    if (!z || !tt || !ss || !hhq || !den || !den_pot) return efail("null pointer", INMSOM_BT_EINVAL);
    CHE(hipSetDevice(c->device));
    if (!c->eos_z) CHE(hipMalloc(&c->eos_z, 1000 * sizeof(double)));
    CHE(hipMemcpyAsync(c->eos_z, z, nz * sizeof(double), hipMemcpyHostToDevice, c->stream));
    EosParams p{};
    p.g = c->g;
    p.nz = nz;
    p.z = c->eos_z;
    p.tt = tt;
    p.ss = ss;
    p.hhq = hhq;
    p.den = den;
    p.den_pot = den_pot;
    launch(c, p);
    CHE(hipGetLastError());
    CHE(hipStreamSynchronize(c->stream));
    return INMSOM_BT_OK;
}

int inmsom_bt_density_timed(inmsom_bt_ctx *c, int nz, int nrep, double *ms)
{
    if (!c || !ms) return efail("null pointer", INMSOM_BT_EINVAL);
    if (nrep < 1) return efail("nrep must be >= 1", INMSOM_BT_EINVAL);
    if (!c->eos || c->eos_nz != nz)
        return efail("no fields staged for nz levels: call inmsom_bt_density first", INMSOM_BT_ESTATE);
    CHE(hipSetDevice(c->device));
    const EosParams p = staged(c, nz);
    hipEvent_t e0, e1;
    CHE(hipEventCreate(&e0));
    CHE(hipEventCreate(&e1));
    CHE(hipEventRecord(e0, c->stream));
    for (int r = 0; r < nrep; ++r) launch(c, p);
    CHE(hipGetLastError());
    CHE(hipEventRecord(e1, c->stream));
    CHE(hipEventSynchronize(e1));
    float t = 0.0f;
    CHE(hipEventElapsedTime(&t, e0, e1));
    CHE(hipEventDestroy(e0));
    CHE(hipEventDestroy(e1));
    *ms = t;
    return INMSOM_BT_OK;
}

}  // extern "C"
