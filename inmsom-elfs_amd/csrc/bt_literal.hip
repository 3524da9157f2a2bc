// Stage-by-stage device implementation of INMSOM's barotropic step.
//
// Each kernel is one sweep of the reference with the reference's index
// range, mask test and operation order, so that the result is the
// reference's bit for bit.  This file is compiled with
// -ffp-contract=off: a fused multiply-add would round differently.
// These kernels serve the general drop-in (any metrics, nstep, ksw4,
// periodicity); the fused single-kernel step for the common case lives
// in bt_fused.hip.
#include "bt_internal.h"
#include "bt_literal.h"

namespace ibt {

// (m, n) of a 2-D launch over the whole array; the blocks go to the XCDs
// in contiguous row-major runs (xcd_tile)
#define KERNEL_MN                                                                                   \
    int bx_, by_;                                                                                   \
    xcd_tile((int)gridDim.x, (int)gridDim.y, (int)(blockIdx.y * gridDim.x + blockIdx.x), bx_, by_); \
    const int m = bx_ * (int)blockDim.x + (int)threadIdx.x + 1;                                     \
    const int n = by_ * (int)blockDim.y + (int)threadIdx.y + 1;

#define IN_INTERIOR(g) (m >= 3 && m <= (g).nx - 2 && n >= 3 && n <= (g).ny - 2)
#define I(mm, nn) IDX(g, (mm), (nn))

// ---------------------------------------------------------------- helpers
// a plane read at (a, b)
struct Plane {
    const DevGrid *g;
    const double *f;
    __device__ double operator()(int a, int b) const { return f[IDX(*g, a, b)]; }
};

__global__ void k_fill(double *a, size_t n, double v)
{
    size_t i = (size_t)blockIdx.x * blockDim.x + threadIdx.x;
    if (i < n) a[i] = v;
}

__global__ void k_copy(double *__restrict__ dst, const double *__restrict__ src, size_t n)
{
    size_t i = (size_t)blockIdx.x * blockDim.x + threadIdx.x;
    if (i < n) dst[i] = src[i];
}

// Service/oc_algorithms.f90:50-95 as a decomposed run performs it:
// column mm -> mmm-1, mmm -> mm+1 over all rows.
__global__ void k_cyclize_x(DevGrid g, double *f)
{
    int n = blockIdx.x * blockDim.x + threadIdx.x + 1;
    if (n > g.ny) return;
    double a = f[I(g.nx - 2, n)], b = f[I(3, n)];
    f[I(2, n)] = a;
    f[I(g.nx - 1, n)] = b;
}

// Service/oc_algorithms.f90:118-135.
__global__ void k_cyclize_y(DevGrid g, double *f)
{
    int m = blockIdx.x * blockDim.x + threadIdx.x + 1;
    if (m > g.nx) return;
    double a = f[I(m, g.ny - 2)], b = f[I(m, 3)];
    f[I(m, 2)] = a;
    f[I(m, g.ny - 1)] = b;
}

// Metrics and masks of the grid at (a, b), for the stencil helpers below.
// CellView reads the grid's per-cell planes; RowView (for grids whose
// metrics are uniform along x on columns 2..nx-1) reads a table of nine
// rows-long metric vectors (dx dy dxt dyt dxh dyh dxb dyb rlh_s) and packed
// mask bits (MB_*) for the same values.
struct CellView {
    const DevGrid *g;
    __device__ size_t at(int a, int b) const { return IDX(*g, a, b); }
    __device__ double dx(int a, int b) const { return g->dx[at(a, b)]; }
    __device__ double dy(int a, int b) const { return g->dy[at(a, b)]; }
    __device__ double dxt(int a, int b) const { return g->dxt[at(a, b)]; }
    __device__ double dyt(int a, int b) const { return g->dyt[at(a, b)]; }
    __device__ double dxh(int a, int b) const { return g->dxh[at(a, b)]; }
    __device__ double dyh(int a, int b) const { return g->dyh[at(a, b)]; }
    __device__ double dxb(int a, int b) const { return g->dxb[at(a, b)]; }
    __device__ double dyb(int a, int b) const { return g->dyb[at(a, b)]; }
    __device__ float lu(int a, int b) const { return g->lu[at(a, b)]; }
    __device__ bool is(const float *mask, int a, int b) const { return mask[at(a, b)] > 0.5f; }
    __device__ bool is_lu(int a, int b) const { return is(g->lu, a, b); }
    __device__ bool is_luu(int a, int b) const { return is(g->luu, a, b); }
    __device__ bool is_luh(int a, int b) const { return is(g->luh, a, b); }
    __device__ bool is_llu(int a, int b) const { return is(g->llu, a, b); }
    __device__ bool is_llv(int a, int b) const { return is(g->llv, a, b); }
    __device__ bool is_lcu(int a, int b) const { return is(g->lcu, a, b); }
    __device__ bool is_lcv(int a, int b) const { return is(g->lcv, a, b); }
};

struct RowView {
    const DevGrid *g;
    const double *__restrict__ rm;
    const uint8_t *__restrict__ mb;
    __device__ size_t at(int a, int b) const { return IDX(*g, a, b); }
    __device__ double met(int k, int b) const { return rm[(size_t)k * (size_t)g->ny + (size_t)(b - 1)]; }
    __device__ double dx(int, int b) const { return met(0, b); }
    __device__ double dy(int, int b) const { return met(1, b); }
    __device__ double dxt(int, int b) const { return met(2, b); }
    __device__ double dyt(int, int b) const { return met(3, b); }
    __device__ double dxh(int, int b) const { return met(4, b); }
    __device__ double dyh(int, int b) const { return met(5, b); }
    __device__ double dxb(int, int b) const { return met(6, b); }
    __device__ double dyb(int, int b) const { return met(7, b); }
    __device__ float lu(int a, int b) const { return (mb[at(a, b)] & MB_LU) ? 1.0f : 0.0f; }
    __device__ bool is_lu(int a, int b) const { return mb[at(a, b)] & MB_LU; }
    __device__ bool is_luu(int a, int b) const { return mb[at(a, b)] & MB_LUU; }
    __device__ bool is_luh(int a, int b) const { return mb[at(a, b)] & MB_LUH; }
    __device__ bool is_llu(int a, int b) const { return mb[at(a, b)] & MB_LLU; }
    __device__ bool is_llv(int a, int b) const { return mb[at(a, b)] & MB_LLV; }
    __device__ bool is_lcu(int a, int b) const { return mb[at(a, b)] & MB_LCU; }
    __device__ bool is_lcv(int a, int b) const { return mb[at(a, b)] & MB_LCV; }
};

// Depth on U, V and H points from T-point depths hq(m, n) given by Q,
// Function/depth.f90:29-66.
template <class V, class Q>
__device__ inline double avg_u_w(const V &w, Q q, int m, int n)
{
    double slu = (double)(w.lu(m, n) + w.lu(m + 1, n));
    return (q(m, n) * w.dx(m, n) * w.dy(m, n) * (double)w.lu(m, n) +
            q(m + 1, n) * w.dx(m + 1, n) * w.dy(m + 1, n) * (double)w.lu(m + 1, n)) /
           slu / w.dxt(m, n) / w.dyh(m, n);
}

template <class V, class Q>
__device__ inline double avg_v_w(const V &w, Q q, int m, int n)
{
    double slu = (double)(w.lu(m, n) + w.lu(m, n + 1));
    return (q(m, n) * w.dx(m, n) * w.dy(m, n) * (double)w.lu(m, n) +
            q(m, n + 1) * w.dx(m, n + 1) * w.dy(m, n + 1) * (double)w.lu(m, n + 1)) /
           slu / w.dxh(m, n) / w.dyt(m, n);
}

template <class V, class Q>
__device__ inline double avg_h_w(const V &w, Q q, int m, int n)
{
    double slu = (double)(w.lu(m, n) + w.lu(m + 1, n) + w.lu(m, n + 1) + w.lu(m + 1, n + 1));
    return (q(m, n) * w.dx(m, n) * w.dy(m, n) * (double)w.lu(m, n) +
            q(m + 1, n) * w.dx(m + 1, n) * w.dy(m + 1, n) * (double)w.lu(m + 1, n) +
            q(m, n + 1) * w.dx(m, n + 1) * w.dy(m, n + 1) * (double)w.lu(m, n + 1) +
            q(m + 1, n + 1) * w.dx(m + 1, n + 1) * w.dy(m + 1, n + 1) * (double)w.lu(m + 1, n + 1)) /
           slu / w.dxb(m, n) / w.dyb(m, n);
}

template <class Q>
__device__ inline double avg_u(const DevGrid &g, Q q, int m, int n)
{
    return avg_u_w(CellView{&g}, q, m, n);
}

template <class Q>
__device__ inline double avg_v(const DevGrid &g, Q q, int m, int n)
{
    return avg_v_w(CellView{&g}, q, m, n);
}

template <class Q>
__device__ inline double avg_h(const DevGrid &g, Q q, int m, int n)
{
    return avg_h_w(CellView{&g}, q, m, n);
}

// ------------------------------------------------------------ depth.f90
// hh_init (Function/depth.f90:2-105): T depths over the whole array,
// U/V/H depths over nx_start-1..nx_end x ny_start-1..ny_end.
__global__ void k_hh_init(DevGrid g, const double *__restrict__ sh, const double *__restrict__ shp,
                          double *const *__restrict__ d)
{
    KERNEL_MN
    if (m > g.nx || n > g.ny) return;
    const double ffs = (double)g.ffs;
    const size_t i = I(m, n);
    d[0][i] = g.hr[i] + sh[i] * ffs;
    d[1][i] = g.hr[i] + shp[i] * ffs;
    d[2][i] = g.hr[i];
    if (m < 2 || m > g.nx - 2 || n < 2 || n > g.ny - 2) return;
    auto q = [&](int a, int b) { size_t j = I(a, b); return g.hr[j] + sh[j] * ffs; };
    auto qp = [&](int a, int b) { size_t j = I(a, b); return g.hr[j] + shp[j] * ffs; };
    auto qn = [&](int a, int b) { return g.hr[I(a, b)]; };
    if (g.llu[i] > 0.5f) {
        d[3][i] = avg_u(g, q, m, n);
        d[4][i] = avg_u(g, qp, m, n);
        d[5][i] = avg_u(g, qn, m, n);
    }
    if (g.llv[i] > 0.5f) {
        d[6][i] = avg_v(g, q, m, n);
        d[7][i] = avg_v(g, qp, m, n);
        d[8][i] = avg_v(g, qn, m, n);
    }
    if (g.luh[i] > 0.5f) {
        d[9][i] = avg_h(g, q, m, n);
        d[10][i] = avg_h(g, qp, m, n);
        d[11][i] = avg_h(g, qn, m, n);
    }
}

// hh_update (Function/depth.f90:108-170)
__global__ void k_hh_update(DevGrid g, const double *__restrict__ sh, double *__restrict__ hqn,
                            double *__restrict__ hun, double *__restrict__ hvn,
                            double *__restrict__ hhn)
{
    KERNEL_MN
    if (m > g.nx || n > g.ny) return;
    const size_t i = I(m, n);
    hqn[i] = g.hr[i] + sh[i];
    if (m < 2 || m > g.nx - 2 || n < 2 || n > g.ny - 2) return;
    auto q = [&](int a, int b) { size_t j = I(a, b); return g.hr[j] + sh[j]; };
    if (g.llu[i] > 0.5f) hun[i] = avg_u(g, q, m, n);
    if (g.llv[i] > 0.5f) hvn[i] = avg_v(g, q, m, n);
    if (g.luh[i] > 0.5f) hhn[i] = avg_h(g, q, m, n);
}

// time-level shift of ssh/u/v with the Robert-Asselin filter
// (Function/vel_ssh.f90:588-612) and of the depths (depth.f90:174-220),
// over nx_start-1..nx_end+1 x ny_start-1..ny_end+1.
__global__ void k_shift(DevGrid g, double dn, int with_depths, double *__restrict__ ssh,
                        double *__restrict__ sshp, const double *__restrict__ sshn,
                        double *__restrict__ u, double *__restrict__ up, const double *__restrict__ un,
                        double *__restrict__ v, double *__restrict__ vp, const double *__restrict__ vn,
                        double *const *__restrict__ d)
{
    KERNEL_MN
    if (m < 2 || m > g.nx - 1 || n < 2 || n > g.ny - 1) return;
    const size_t i = I(m, n);
    const double ts = g.ts;
    if (g.lu[i] > 0.5f) {
        sshp[i] = ssh[i] + ts * (sshn[i] - 2.0 * ssh[i] + sshp[i]) / 2.0 / dn;
        ssh[i] = sshn[i];
    }
    if (g.lcu[i] > 0.5f) {
        up[i] = u[i] + ts * (un[i] - 2.0 * u[i] + up[i]) / 2.0 / dn;
        u[i] = un[i];
    }
    if (g.lcv[i] > 0.5f) {
        vp[i] = v[i] + ts * (vn[i] - 2.0 * v[i] + vp[i]) / 2.0 / dn;
        v[i] = vn[i];
    }
    if (!with_depths) return;
    auto shift = [&](int a) {
        double *h = d[a], *hp = d[a + 1], *hn = d[a + 2];
        hp[i] = h[i] + ts * (hn[i] - 2.0 * h[i] + hp[i]) / 2.0 / dn;
        h[i] = hn[i];
    };
    if (g.llu[i] > 0.5f) shift(3);
    if (g.llv[i] > 0.5f) shift(6);
    if (g.lu[i] > 0.5f) shift(0);
    if (g.luh[i] > 0.5f) shift(9);
}

// ------------------------------------------------------------ vel_ssh.f90
// continuity (Function/vel_ssh.f90:487-499)
__global__ void k_continuity(DevGrid g, double tau_inner, const double *__restrict__ wflux,
                             const double *__restrict__ u, const double *__restrict__ v,
                             const double *__restrict__ hhu, const double *__restrict__ hhv,
                             const double *__restrict__ sshp, double *__restrict__ sshn)
{
    KERNEL_MN
    if (!IN_INTERIOR(g)) return;
    const size_t i = I(m, n);
    if (g.lu[i] > 0.5f)
        sshn[i] = sshp[i] +
                  2.0 * tau_inner *
                      (wflux[i] / g.rho0 * (double)g.ffs -
                       (u[i] * hhu[i] * g.dyh[i] - u[I(m - 1, n)] * hhu[I(m - 1, n)] * g.dyh[I(m - 1, n)] +
                        v[i] * hhv[i] * g.dxh[i] - v[I(m, n - 1)] * hhv[I(m, n - 1)] * g.dxh[I(m, n - 1)]) /
                           (g.dx[i] * g.dy[i]));
}

// stress_components (Function/mixing.f90:115-168) at (m, n) of the
// velocity pair given by accessors, for a point of the interior: tension on
// lu points, shear on luu points.
template <class V, class A, class B>
__device__ inline double tension_at(const V &w, A u, B v, int m, int n)
{
    return w.dy(m, n) / w.dx(m, n) * (u(m, n) / w.dyh(m, n) - u(m - 1, n) / w.dyh(m - 1, n)) -
           w.dx(m, n) / w.dy(m, n) * (v(m, n) / w.dxh(m, n) - v(m, n - 1) / w.dxh(m, n - 1));
}

template <class V, class A, class B>
__device__ inline double shear_at(const V &w, A u, B v, int m, int n)
{
    return w.dxb(m, n) / w.dyb(m, n) * (u(m, n + 1) / w.dxt(m, n + 1) - u(m, n) / w.dxt(m, n)) +
           w.dyb(m, n) / w.dxb(m, n) * (v(m + 1, n) / w.dyt(m + 1, n) - v(m, n) / w.dyt(m, n));
}

__global__ void k_stress(DevGrid g, const double *__restrict__ u, const double *__restrict__ v,
                         double *__restrict__ str_t, double *__restrict__ str_s)
{
    KERNEL_MN
    if (!IN_INTERIOR(g)) return;
    const size_t i = I(m, n);
    if (g.lu[i] > 0.5f) str_t[i] = tension_at(CellView{&g}, Plane{&g, u}, Plane{&g, v}, m, n);
    if (g.luu[i] > 0.5f) str_s[i] = shear_at(CellView{&g}, Plane{&g, u}, Plane{&g, v}, m, n);
}

// metric vorticity of uv_trans (Function/vel_ssh.f90:30-42)
__global__ void k_vort(DevGrid g, const double *__restrict__ u, const double *__restrict__ v,
                       double *__restrict__ vort)
{
    KERNEL_MN
    if (!IN_INTERIOR(g)) return;
    const size_t i = I(m, n);
    if (g.luu[i] > 0.5f)
        vort[i] = (v[I(m + 1, n)] * g.dyt[I(m + 1, n)] - v[i] * g.dyt[i]) -
                  (u[I(m, n + 1)] * g.dxt[I(m, n + 1)] - u[i] * g.dxt[i]) -
                  ((v[I(m + 1, n)] - v[i]) * g.dyb[i] - (u[I(m, n + 1)] - u[i]) * g.dxb[i]);
}

// Tension part p and shear part q of the viscous stencil,
// Function/vel_ssh.f90:150-156 (u) and :166-172 (v); the stress, depths
// at T and H points come from accessors (a, b) -> value.
template <class V, class ST, class SS, class HQ, class HH>
__device__ inline void visc_u_at(const V &w, const double *mu, ST str_t, SS str_s, HQ hq, HH hh, int m, int n,
                                 double &p, double &q)
{
    auto M = [&](int a, int b) { return mu[w.at(a, b)]; };
    double muh_p = (M(m, n) + M(m + 1, n) + M(m, n + 1) + M(m + 1, n + 1)) / 4.0;
    double muh_m = (M(m, n) + M(m + 1, n) + M(m, n - 1) + M(m + 1, n - 1)) / 4.0;
    double dy1 = w.dy(m + 1, n), dy0 = w.dy(m, n);
    double b1 = w.dxb(m, n), b0 = w.dxb(m, n - 1);
    p = (dy1 * dy1 * M(m + 1, n) * hq(m + 1, n) * str_t(m + 1, n) - dy0 * dy0 * M(m, n) * hq(m, n) * str_t(m, n)) /
        w.dyh(m, n);
    q = (b1 * b1 * muh_p * hh(m, n) * str_s(m, n) - b0 * b0 * muh_m * hh(m, n - 1) * str_s(m, n - 1)) / w.dxt(m, n);
}

template <class V, class ST, class SS, class HQ, class HH>
__device__ inline void visc_v_at(const V &w, const double *mu, ST str_t, SS str_s, HQ hq, HH hh, int m, int n,
                                 double &p, double &q)
{
    auto M = [&](int a, int b) { return mu[w.at(a, b)]; };
    double muh_p = (M(m, n) + M(m + 1, n) + M(m, n + 1) + M(m + 1, n + 1)) / 4.0;
    double muh_m = (M(m, n) + M(m - 1, n) + M(m, n + 1) + M(m - 1, n + 1)) / 4.0;
    double dx1 = w.dx(m, n + 1), dx0 = w.dx(m, n);
    double b1 = w.dyb(m, n), b0 = w.dyb(m - 1, n);
    p = (dx1 * dx1 * M(m, n + 1) * hq(m, n + 1) * str_t(m, n + 1) - dx0 * dx0 * M(m, n) * hq(m, n) * str_t(m, n)) /
        w.dxh(m, n);
    q = (b1 * b1 * muh_p * hh(m, n) * str_s(m, n) - b0 * b0 * muh_m * hh(m - 1, n) * str_s(m - 1, n)) / w.dyt(m, n);
}

__device__ inline void visc_u(const DevGrid &g, const double *mu, const double *str_t,
                              const double *str_s, const double *hq, const double *hh, int m, int n,
                              double &p, double &q)
{
    visc_u_at(CellView{&g}, mu, Plane{&g, str_t}, Plane{&g, str_s}, Plane{&g, hq}, Plane{&g, hh}, m, n, p, q);
}

__device__ inline void visc_v(const DevGrid &g, const double *mu, const double *str_t,
                              const double *str_s, const double *hq, const double *hh, int m, int n,
                              double &p, double &q)
{
    visc_v_at(CellView{&g}, mu, Plane{&g, str_t}, Plane{&g, str_s}, Plane{&g, hq}, Plane{&g, hh}, m, n, p, q);
}

// flux-form advection of uv_trans (Function/vel_ssh.f90:54-109) and the
// harmonic viscosity of uv_diff2 (:141-179) in one sweep.
__global__ void k_trans_diff2(DevGrid g, const double *__restrict__ u, const double *__restrict__ v,
                              const double *__restrict__ vort, const double *__restrict__ hq,
                              const double *__restrict__ hu, const double *__restrict__ hv,
                              const double *__restrict__ hh, const double *__restrict__ mu,
                              const double *__restrict__ str_t, const double *__restrict__ str_s,
                              double *__restrict__ radv, double *__restrict__ rady,
                              double *__restrict__ rdfx, double *__restrict__ rdfy)
{
    KERNEL_MN
    if (!IN_INTERIOR(g)) return;
    const size_t i = I(m, n);
    const double *dyh = g.dyh, *dxh = g.dxh;
    if (g.lcu[i] > 0.5f) {
        double fx_p = (u[i] * dyh[i] * hu[i] + u[I(m + 1, n)] * dyh[I(m + 1, n)] * hu[I(m + 1, n)]) /
                      2.0 * (u[i] + u[I(m + 1, n)]) / 2.0;
        double fx_m = (u[i] * dyh[i] * hu[i] + u[I(m - 1, n)] * dyh[I(m - 1, n)] * hu[I(m - 1, n)]) /
                      2.0 * (u[i] + u[I(m - 1, n)]) / 2.0;
        double fy_p = (v[i] * dxh[i] * hv[i] + v[I(m + 1, n)] * dxh[I(m + 1, n)] * hv[I(m + 1, n)]) /
                      2.0 * (u[I(m, n + 1)] + u[i]) / 2.0 * (double)g.luu[i];
        double fy_m = (v[I(m, n - 1)] * dxh[I(m, n - 1)] * hv[I(m, n - 1)] +
                       v[I(m + 1, n - 1)] * dxh[I(m + 1, n - 1)] * hv[I(m + 1, n - 1)]) /
                      2.0 * (u[I(m, n - 1)] + u[i]) / 2.0 * (double)g.luu[I(m, n - 1)];
        radv[i] = -(fx_p - fx_m + fy_p - fy_m) +
                  (vort[i] * hh[i] * (v[I(m + 1, n)] + v[i]) +
                   vort[I(m, n - 1)] * hh[I(m, n - 1)] * (v[I(m + 1, n - 1)] + v[I(m, n - 1)])) / 4.0;
        double p, q;
        visc_u(g, mu, str_t, str_s, hq, hh, m, n, p, q);
        rdfx[i] = p + q;
    }
    if (g.lcv[i] > 0.5f) {
        double fy_p = (v[i] * dxh[i] * hv[i] + v[I(m, n + 1)] * dxh[I(m, n + 1)] * hv[I(m, n + 1)]) /
                      2.0 * (v[i] + v[I(m, n + 1)]) / 2.0;
        double fy_m = (v[i] * dxh[i] * hv[i] + v[I(m, n - 1)] * dxh[I(m, n - 1)] * hv[I(m, n - 1)]) /
                      2.0 * (v[i] + v[I(m, n - 1)]) / 2.0;
        double fx_p = (u[i] * dyh[i] * hu[i] + u[I(m, n + 1)] * dyh[I(m, n + 1)] * hu[I(m, n + 1)]) /
                      2.0 * (v[I(m + 1, n)] + v[i]) / 2.0;
        double fx_m = (u[I(m - 1, n)] * dyh[I(m - 1, n)] * hu[I(m - 1, n)] +
                       u[I(m - 1, n + 1)] * dyh[I(m - 1, n + 1)] * hu[I(m - 1, n + 1)]) /
                      2.0 * (v[I(m - 1, n)] + v[i]) / 2.0;
        rady[i] = -(fx_p - fx_m + fy_p - fy_m) -
                  (vort[i] * hh[i] * (u[I(m, n + 1)] + u[i]) +
                   vort[I(m - 1, n)] * hh[I(m - 1, n)] * (u[I(m - 1, n + 1)] + u[I(m - 1, n)])) / 4.0;
        double p, q;
        visc_v(g, mu, str_t, str_s, hq, hh, m, n, p, q);
        rdfy[i] = -p + q;
    }
}

// first sweep of uv_diff4 (Function/vel_ssh.f90:216-257); fx, fy are
// zero-filled beforehand (:213-214).
__global__ void k_diff4_f(DevGrid g, const double *__restrict__ mu, const double *__restrict__ str_t,
                          const double *__restrict__ str_s, const double *__restrict__ hq,
                          const double *__restrict__ hu, const double *__restrict__ hv,
                          const double *__restrict__ hh, double *__restrict__ fx,
                          double *__restrict__ fy)
{
    KERNEL_MN
    if (!IN_INTERIOR(g)) return;
    const size_t i = I(m, n);
    double p, q;
    if (g.lcu[i] > 0.5f) {
        visc_u(g, mu, str_t, str_s, hq, hh, m, n, p, q);
        double f = p + q;
        fx[i] = -f / hu[i] / g.dxt[i] / g.dyh[i];
    }
    if (g.lcv[i] > 0.5f) {
        visc_v(g, mu, str_t, str_s, hq, hh, m, n, p, q);
        double f = -p + q;
        fy[i] = -f / hv[i] / g.dxh[i] / g.dyt[i];
    }
}

// last sweep of uv_diff4 (Function/vel_ssh.f90:274-312)
__global__ void k_diff4_add(DevGrid g, const double *__restrict__ mu, const double *__restrict__ str_t,
                            const double *__restrict__ str_s, const double *__restrict__ hq,
                            const double *__restrict__ hh, double *__restrict__ rdfx,
                            double *__restrict__ rdfy)
{
    KERNEL_MN
    if (!IN_INTERIOR(g)) return;
    const size_t i = I(m, n);
    double p, q;
    if (g.lcu[i] > 0.5f) {
        visc_u(g, mu, str_t, str_s, hq, hh, m, n, p, q);
        rdfx[i] = rdfx[i] + p + q;
    }
    if (g.lcv[i] > 0.5f) {
        visc_v(g, mu, str_t, str_s, hq, hh, m, n, p, q);
        rdfy[i] = rdfy[i] - p + q;
    }
}

// uv_diff4 (Function/vel_ssh.f90:203-312) in two sweeps for a caller that
// adds its terms itself in the reference's order (RHSx_dif + px + qx and
// RHSy_dif - py + qy), on a grid that is not periodic.  Each sweep forms
// what the stage-by-stage path keeps in planes (level-n depths of hh_update,
// Function/depth.f90:108-170; the stresses of stress_components) at the
// points its stencil reads, with the same expressions, and the value 0 where
// the reference does not compute them (outside their index range or mask),
// as the zero-initialised planes hold there.
// A view whose reads are moved into the array ([1, nx] x [1, ny]): the
// stencil of a point outside the interior can then be formed without a
// branch and discarded; at interior points nothing moves.
template <class V>
struct ClampView {
    V w;
    const DevGrid *g;
    __device__ int cx(int a) const { return min(max(a, 1), g->nx); }
    __device__ int cy(int b) const { return min(max(b, 1), g->ny); }
    __device__ size_t at(int a, int b) const { return w.at(cx(a), cy(b)); }
    __device__ double dx(int a, int b) const { return w.dx(cx(a), cy(b)); }
    __device__ double dy(int a, int b) const { return w.dy(cx(a), cy(b)); }
    __device__ double dxt(int a, int b) const { return w.dxt(cx(a), cy(b)); }
    __device__ double dyt(int a, int b) const { return w.dyt(cx(a), cy(b)); }
    __device__ double dxh(int a, int b) const { return w.dxh(cx(a), cy(b)); }
    __device__ double dyh(int a, int b) const { return w.dyh(cx(a), cy(b)); }
    __device__ double dxb(int a, int b) const { return w.dxb(cx(a), cy(b)); }
    __device__ double dyb(int a, int b) const { return w.dyb(cx(a), cy(b)); }
};

template <class V>
struct D4Depth {
    const V w;
    const double *sh;
    // hq = H + ssh at a T point
    __device__ double hq(int a, int b) const
    {
        const size_t j = w.at(a, b);
        return w.g->hr[j] + sh[j];
    }
    __device__ bool in_range(int a, int b) const
    {
        return a >= 2 && a <= w.g->nx - 2 && b >= 2 && b <= w.g->ny - 2;
    }
    __device__ double hu(int a, int b) const
    {
        if (!in_range(a, b) || !w.is_llu(a, b)) return 0.0;
        return avg_u_w(w, [this](int x, int y) { return hq(x, y); }, a, b);
    }
    __device__ double hv(int a, int b) const
    {
        if (!in_range(a, b) || !w.is_llv(a, b)) return 0.0;
        return avg_v_w(w, [this](int x, int y) { return hq(x, y); }, a, b);
    }
    __device__ double hh(int a, int b) const
    {
        if (!in_range(a, b) || !w.is_luh(a, b)) return 0.0;
        return avg_h_w(w, [this](int x, int y) { return hq(x, y); }, a, b);
    }
    __device__ double hu_nb(int a, int b) const
    {
        const int ac = min(max(a, 1), w.g->nx - 1), bc = min(max(b, 1), w.g->ny - 1);
        const double h = avg_u_w(w, [this](int x, int y) { return hq(x, y); }, ac, bc);
        return in_range(a, b) && w.is_llu(ac, bc) ? h : 0.0;
    }
    __device__ double hv_nb(int a, int b) const
    {
        const int ac = min(max(a, 1), w.g->nx - 1), bc = min(max(b, 1), w.g->ny - 1);
        const double h = avg_v_w(w, [this](int x, int y) { return hq(x, y); }, ac, bc);
        return in_range(a, b) && w.is_llv(ac, bc) ? h : 0.0;
    }
    // the same values without a branch: the average is formed at (a, b)
    // moved into [1, nx-1] x [1, ny-1] (every point it reads in the array;
    // unchanged where the value is used) and selected
    __device__ double hh_nb(int a, int b) const
    {
        const int ac = min(max(a, 1), w.g->nx - 1), bc = min(max(b, 1), w.g->ny - 1);
        const double h = avg_h_w(w, [this](int x, int y) { return hq(x, y); }, ac, bc);
        return in_range(a, b) && w.is_luh(ac, bc) ? h : 0.0;
    }
};

template <class V>
struct D4Stress {
    const V w;
    Plane u, v;
    __device__ bool interior(int a, int b) const
    {
        return a >= 3 && a <= w.g->nx - 2 && b >= 3 && b <= w.g->ny - 2;
    }
    __device__ double t(int a, int b) const
    {
        if (!interior(a, b) || !w.is_lu(a, b)) return 0.0;
        return tension_at(w, u, v, a, b);
    }
    __device__ double s(int a, int b) const
    {
        if (!interior(a, b) || !w.is_luu(a, b)) return 0.0;
        return shear_at(w, u, v, a, b);
    }
    // t and s without a branch, as D4Depth::hh_nb: tension reads (a-1..a,
    // b-1..b), shear (a..a+1, b..b+1)
    __device__ double t_nb(int a, int b) const
    {
        const int ac = min(max(a, 2), w.g->nx - 1), bc = min(max(b, 2), w.g->ny - 1);
        const double t = tension_at(w, u, v, ac, bc);
        return interior(a, b) && w.is_lu(ac, bc) ? t : 0.0;
    }
    __device__ double s_nb(int a, int b) const
    {
        const int ac = min(max(a, 1), w.g->nx - 1), bc = min(max(b, 1), w.g->ny - 1);
        const double sv = shear_at(w, u, v, ac, bc);
        return interior(a, b) && w.is_luu(ac, bc) ? sv : 0.0;
    }
};

// One sweep per launch over tiles of D4_TX x D4_TY points (one thread per
// column, D4_ROWS rows of threads, a wave per row of threads): the tile
// first forms, once per point, the tension on [m0, m0 + D4_TX] x
// [n0, n0 + D4_TY] and the shear and H-point depths on
// [m0 - 1, m0 + D4_TX - 1] x [n0 - 1, n0 + D4_TY - 1] in LDS, which is every
// point the viscous stencil of its own points reads.
// FIRST: the sweep of :216-257 from the stress of (a, b) = (up, vp), giving
// o0, o1 = fx, fy; else the sweep of :259-312 from the stress of
// (a, b) = (fx, fy), giving o0..o3 = px, qx, py, qy.  Every point of the
// array is written, 0 off lcu / lcv and outside the interior.  A wave's
// points share their row, so with RowView its metric reads are uniform.
constexpr int D4_TX = 64, D4_TY = 15, D4_ROWS = 4;
constexpr int D4_W = D4_TX + 1, D4_H = D4_TY + 1;

template <bool FIRST, class V>
__device__ __forceinline__ void d4_sweep(const V &w, const double *__restrict__ mu, const double *__restrict__ a,
                                         const double *__restrict__ b, const double *__restrict__ sh,
                                         double *__restrict__ o0, double *__restrict__ o1, double *__restrict__ o2,
                                         double *__restrict__ o3)
{
    __shared__ double Ts[D4_H][D4_W], Ss[D4_H][D4_W], Hs[D4_H][D4_W];
    const DevGrid &g = *w.g;
    int bx, by;
    xcd_tile((int)gridDim.x, (int)gridDim.y, (int)(blockIdx.y * gridDim.x + blockIdx.x), bx, by);
    const int m0 = 1 + bx * D4_TX, n0 = 1 + by * D4_TY;
    const int tx = (int)threadIdx.x, tid = (int)(threadIdx.y * D4_TX + threadIdx.x);
    const D4Depth<V> d{w, sh};
    const D4Stress<V> st{w, Plane{&g, a}, Plane{&g, b}};
    for (int r = (int)threadIdx.y; r < D4_H; r += D4_ROWS) {
        const int ru = __builtin_amdgcn_readfirstlane(r);  // one row per wave
        Ts[ru][tx] = st.t_nb(m0 + tx, n0 + ru);
        Ss[ru][tx] = st.s_nb(m0 - 1 + tx, n0 - 1 + ru);
        Hs[ru][tx] = d.hh_nb(m0 - 1 + tx, n0 - 1 + ru);
    }
    if (tid < D4_H) {  // the tile's last column
        Ts[tid][D4_TX] = st.t(m0 + D4_TX, n0 + tid);
        Ss[tid][D4_TX] = st.s(m0 - 1 + D4_TX, n0 - 1 + tid);
        Hs[tid][D4_TX] = d.hh(m0 - 1 + D4_TX, n0 - 1 + tid);
    }
    __syncthreads();
    auto tt = [&](int x, int y) { return Ts[y - n0][x - m0]; };
    auto ss = [&](int x, int y) { return Ss[y - n0 + 1][x - m0 + 1]; };
    auto hh = [&](int x, int y) { return Hs[y - n0 + 1][x - m0 + 1]; };
    // phase 2 forms both stencils at every point of the tile without a
    // branch (LDS reads stay in the tile; array reads go through cw) and
    // keeps them on lcu / lcv points of the interior
    const ClampView<V> cw{w, &g};
    auto hq = [&](int x, int y) { return d.hq(cw.cx(x), cw.cy(y)); };
    const int m = m0 + tx;
    if (m > g.nx) return;
    for (int j = (int)threadIdx.y; j < D4_TY; j += D4_ROWS) {
        const int n = __builtin_amdgcn_readfirstlane(n0 + j);  // one row per wave
        if (n > g.ny) return;
        const size_t i = w.at(m, n);
        double pu, qu, pv, qv;
        const bool in = IN_INTERIOR(g);
        const bool cu = in && w.is_lcu(m, n), cv = in && w.is_lcv(m, n);
        visc_u_at(cw, mu, tt, ss, hq, hh, m, n, pu, qu);
        visc_v_at(cw, mu, tt, ss, hq, hh, m, n, pv, qv);
        if constexpr (FIRST) {
            const double fu = pu + qu;
            const double vx = -fu / d.hu_nb(m, n) / cw.dxt(m, n) / cw.dyh(m, n);
            const double fv = -pv + qv;
            const double vy = -fv / d.hv_nb(m, n) / cw.dxh(m, n) / cw.dyt(m, n);
            o0[i] = cu ? vx : 0.0;
            o1[i] = cv ? vy : 0.0;
        } else {
            o0[i] = cu ? pu : 0.0;
            o1[i] = cu ? qu : 0.0;
            o2[i] = cv ? pv : 0.0;
            o3[i] = cv ? qv : 0.0;
        }
    }
}

// ROW: metrics from rm, the row table (9 x ny) of a grid whose metrics are
// uniform along x, and masks from mb, the packed mask bits; else the grid's
// per-cell planes
template <bool FIRST, bool ROW>
__global__ __launch_bounds__(D4_TX *D4_ROWS) void k_d4_sweep(DevGrid g, const double *__restrict__ rm,
                                                             const uint8_t *__restrict__ mb,
                                                             const double *__restrict__ mu,
                                                             const double *__restrict__ a,
                                                             const double *__restrict__ b,
                                                             const double *__restrict__ sh, double *__restrict__ o0,
                                                             double *__restrict__ o1, double *__restrict__ o2,
                                                             double *__restrict__ o3)
{
    if constexpr (ROW)
        d4_sweep<FIRST>(RowView{&g, rm, mb}, mu, a, b, sh, o0, o1, o2, o3);
    else
        d4_sweep<FIRST>(CellView{&g}, mu, a, b, sh, o0, o1, o2, o3);
}

// momentum with Coriolis, pressure gradient, Rayleigh friction and the
// time-mean accumulators (Function/vel_ssh.f90:531-573)
struct MomArgs {
    const double *RHSx, *RHSy, *rdis, *rdfx, *rdfy, *radv, *rady;
    const double *u, *v, *up, *vp, *ssh, *sshn;
    const double *hhu, *hhup, *hhun, *hhv, *hhvp, *hhvn, *hhh;
    double *un, *vn, *ubrtr_i, *vbrtr_i, *ssh4gx, *ssh4gy;
};

__global__ void k_momentum(DevGrid g, double tau_inner, double d4n, MomArgs a)
{
    KERNEL_MN
    if (!IN_INTERIOR(g)) return;
    const size_t i = I(m, n);
    const double G = g.g;
    if (g.lcu[i] > 0.5f) {
        double bp = a.hhun[i] * g.dxt[i] * g.dyh[i] / 2.0 / tau_inner;
        double bp0 = a.hhup[i] * g.dxt[i] * g.dyh[i] / 2.0 / tau_inner;
        double slx = -G * (a.ssh[I(m + 1, n)] - a.ssh[i]) * g.dyh[i] * a.hhu[i];
        double slxn = -G * (a.sshn[I(m + 1, n)] - a.sshn[i]) * g.dyh[i] * a.hhun[i];
        double grx = a.RHSx[i] + slx + a.rdfx[i] + a.radv[i] -
                     (a.rdis[i] + a.rdis[I(m + 1, n)]) / 2.0 * a.up[i] * g.dxt[i] * g.dyh[i] *
                         a.hhu[i] +
                     (g.rlh_s[i] * a.hhh[i] * g.dxb[i] * g.dyb[i] * (a.v[I(m + 1, n)] + a.v[i]) +
                      g.rlh_s[I(m, n - 1)] * a.hhh[I(m, n - 1)] * g.dxb[I(m, n - 1)] *
                          g.dyb[I(m, n - 1)] * (a.v[I(m + 1, n - 1)] + a.v[I(m, n - 1)])) /
                         4.0;
        double unew = (a.up[i] * bp0 + grx) / bp;
        a.un[i] = unew;
        a.ubrtr_i[i] = a.ubrtr_i[i] + (a.u[i] * a.hhu[i] + unew * a.hhun[i]) / d4n;
        a.ssh4gx[i] = a.ssh4gx[i] + (slx + slxn) / d4n;
    }
    if (g.lcv[i] > 0.5f) {
        double bp = a.hhvn[i] * g.dyt[i] * g.dxh[i] / 2.0 / tau_inner;
        double bp0 = a.hhvp[i] * g.dyt[i] * g.dxh[i] / 2.0 / tau_inner;
        double sly = -G * (a.ssh[I(m, n + 1)] - a.ssh[i]) * g.dxh[i] * a.hhv[i];
        double slyn = -G * (a.sshn[I(m, n + 1)] - a.sshn[i]) * g.dxh[i] * a.hhvn[i];
        double gry = a.RHSy[i] + sly + a.rdfy[i] + a.rady[i] -
                     (a.rdis[i] + a.rdis[I(m, n + 1)]) / 2.0 * a.vp[i] * g.dxh[i] * g.dyt[i] *
                         a.hhv[i] -
                     (g.rlh_s[i] * a.hhh[i] * g.dxb[i] * g.dyb[i] * (a.u[I(m, n + 1)] + a.u[i]) +
                      g.rlh_s[I(m - 1, n)] * a.hhh[I(m - 1, n)] * g.dxb[I(m - 1, n)] *
                          g.dyb[I(m - 1, n)] * (a.u[I(m - 1, n + 1)] + a.u[I(m - 1, n)])) /
                         4.0;
        double vnew = (a.vp[i] * bp0 + gry) / bp;
        a.vn[i] = vnew;
        a.vbrtr_i[i] = a.vbrtr_i[i] + (a.v[i] * a.hhv[i] + vnew * a.hhvn[i]) / d4n;
        a.ssh4gy[i] = a.ssh4gy[i] + (sly + slyn) / d4n;
    }
}

// ------------------------------------------------------------ fluxes.f90
// wind stress on U/V faces (Function/fluxes.f90:29-37)
__global__ void k_surface_stress(DevGrid g, const double *__restrict__ taux,
                                 const double *__restrict__ tauy, double *__restrict__ sx,
                                 double *__restrict__ sy)
{
    KERNEL_MN
    if (!IN_INTERIOR(g)) return;
    const size_t i = I(m, n);
    if (g.lcu[i] > 0.5f)
        sx[i] = (taux[i] * g.dx[i] * g.dy[i] + taux[I(m + 1, n)] * g.dx[I(m + 1, n)] * g.dy[I(m + 1, n)]) /
                g.rho0 / 2.0 / g.dxt[i] / g.dyh[i];
    if (g.lcv[i] > 0.5f)
        sy[i] = (tauy[i] * g.dx[i] * g.dy[i] + tauy[I(m, n + 1)] * g.dx[I(m, n + 1)] * g.dy[I(m, n + 1)]) /
                g.rho0 / 2.0 / g.dxh[i] / g.dyt[i];
}

// quadratic bottom friction rate on T points (Function/fluxes.f90:231-243)
__global__ void k_bottom_friction(DevGrid g, double cb_nl, double ebottom,
                                  const double *__restrict__ uu, const double *__restrict__ vv,
                                  const double *__restrict__ hhu, const double *__restrict__ hhv,
                                  const double *__restrict__ hhq, double *__restrict__ bf)
{
    KERNEL_MN
    if (!IN_INTERIOR(g)) return;
    const size_t i = I(m, n);
    if (g.lu[i] > 0.5f) {
        double ut = (uu[i] * g.dxt[i] * g.dyh[i] * hhu[i] +
                     uu[I(m - 1, n)] * g.dxt[I(m - 1, n)] * g.dyh[I(m - 1, n)] * hhu[I(m - 1, n)]) /
                    2.0 / g.dx[i] / g.dy[i] / hhq[i];
        double vt = (vv[i] * g.dxh[i] * g.dyt[i] * hhv[i] +
                     vv[I(m, n - 1)] * g.dxh[I(m, n - 1)] * g.dyt[I(m, n - 1)] * hhv[I(m, n - 1)]) /
                    2.0 / g.dx[i] / g.dy[i] / hhq[i];
        bf[i] = cb_nl * sqrt(ut * ut + vt * vt + ebottom);
    }
}

// bottom stress (Function/fluxes.f90:261-272) and the 2-D RHS assembly
// of Control/dyncall.f90:103-126 with every 3-D term zero.
__global__ void k_bottom_stress_rhs(DevGrid g, const double *__restrict__ bf,
                                    const double *__restrict__ uup, const double *__restrict__ vvp,
                                    const double *__restrict__ sx, const double *__restrict__ sy,
                                    double *__restrict__ bx, double *__restrict__ by,
                                    double *__restrict__ rhsx, double *__restrict__ rhsy)
{
    KERNEL_MN
    if (m > g.nx || n > g.ny) return;
    const size_t i = I(m, n);
    if (g.llu[i] > 0.5f) rhsx[i] = 0.0;
    if (g.llv[i] > 0.5f) rhsy[i] = 0.0;
    if (!IN_INTERIOR(g)) return;
    if (g.lcu[i] > 0.5f) {
        bx[i] = -(bf[i] + bf[I(m + 1, n)]) / 2.0 * uup[i];
        rhsx[i] = rhsx[i] + (sx[i] + bx[i]) * g.dxt[i] * g.dyh[i];
    }
    if (g.lcv[i] > 0.5f) {
        by[i] = -(bf[i] + bf[I(m, n + 1)]) / 2.0 * vvp[i];
        rhsy[i] = rhsy[i] + (sy[i] + by[i]) * g.dyt[i] * g.dxh[i];
    }
}

// ------------------------------------------------------------ host side
static inline dim3 grid2(const DevGrid &g) { return dim3((g.nx + 63) / 64, (g.ny + 3) / 4); }
static const dim3 BLK2(64, 4);

static void fill(inmsom_bt_ctx *c, double *a, double v)
{
    hipLaunchKernelGGL(k_fill, dim3((unsigned)((c->n + 255) / 256)), dim3(256), 0, c->stream, a, c->n, v);
}

void lit_copy(inmsom_bt_ctx *c, double *dst, const double *src)
{
    hipLaunchKernelGGL(k_copy, dim3((unsigned)((c->n + 255) / 256)), dim3(256), 0, c->stream, dst, src, c->n);
}

static void cyclize(inmsom_bt_ctx *c, double *f)
{
    if (c->g.px)
        hipLaunchKernelGGL(k_cyclize_x, dim3((c->ny + 255) / 256), dim3(256), 0, c->stream, c->g, f);
    if (c->g.py)
        hipLaunchKernelGGL(k_cyclize_y, dim3((c->nx + 255) / 256), dim3(256), 0, c->stream, c->g, f);
}

void lit_cyclize(inmsom_bt_ctx *c, double *f) { cyclize(c, f); }

void lit_hh_update(inmsom_bt_ctx *c, const double *sh, double *hqn, double *hun, double *hvn, double *hhn)
{
    hipLaunchKernelGGL(k_hh_update, grid2(c->g), BLK2, 0, c->stream, c->g, sh, hqn, hun, hvn, hhn);
    cyclize(c, hun);
    cyclize(c, hvn);
    cyclize(c, hhn);
}

void lit_diff4_terms(inmsom_bt_ctx *c, const double *up, const double *vp, const double *sh, const double *mu4,
                     double *const t[6], const double *rowmet, const uint8_t *mb)
{
    const DevGrid &g = c->g;
    const dim3 grid((unsigned)((g.nx + D4_TX - 1) / D4_TX), (unsigned)((g.ny + D4_TY - 1) / D4_TY));
    const dim3 blk(D4_TX, D4_ROWS);
    if (rowmet && mb) {
        hipLaunchKernelGGL((k_d4_sweep<true, true>), grid, blk, 0, c->stream, g, rowmet, mb, mu4, up, vp, sh, t[0],
                           t[1], nullptr, nullptr);
        hipLaunchKernelGGL((k_d4_sweep<false, true>), grid, blk, 0, c->stream, g, rowmet, mb, mu4, t[0], t[1], sh,
                           t[2], t[3], t[4], t[5]);
    } else {
        hipLaunchKernelGGL((k_d4_sweep<true, false>), grid, blk, 0, c->stream, g, nullptr, nullptr, mu4, up, vp, sh,
                           t[0], t[1], nullptr, nullptr);
        hipLaunchKernelGGL((k_d4_sweep<false, false>), grid, blk, 0, c->stream, g, nullptr, nullptr, mu4, t[0], t[1],
                           sh, t[2], t[3], t[4], t[5]);
    }
}

// Device array of the 12 depth pointers (k_hh_init and k_shift index
// it); written once when the context is created.
static inline double *const *dep_table(inmsom_bt_ctx *c)
{
    return reinterpret_cast<double *const *>(c->scal + DEP_TABLE_OFFSET);
}

void lit_hh_init(inmsom_bt_ctx *c, const double *ssh, const double *sshp)
{
    hipLaunchKernelGGL(k_hh_init, grid2(c->g), BLK2, 0, c->stream, c->g, ssh, sshp, dep_table(c));
    for (int k = 3; k < 12; ++k) cyclize(c, c->dep[k]);
}

// barotropic_dynamics (Function/vel_ssh.f90:379-679) on the context's
// device arrays: state c->st, accumulators c->acc, depths c->dep.
void lit_barotropic_dynamics(inmsom_bt_ctx *c, double tau, int nstep, int ksw4, const double *RHSx,
                             const double *RHSy, const double *wflux, const double *mu,
                             const double *mu4, const double *rdis)
{
    const DevGrid &g = c->g;
    double *u = c->w[0], *v = c->w[1], *ssh = c->w[2], *up = c->w[3], *vp = c->w[4];
    double *sshp = c->w[5], *un = c->w[6], *vn = c->w[7], *sshn = c->w[8];
    double **d = c->dep;
    const double tau_inner = tau / (double)nstep;
    const double d4n = (double)(4 * nstep), dn = (double)nstep;
    double *const *dt = dep_table(c);

    fill(c, un, 0.0);
    fill(c, vn, 0.0);
    fill(c, sshn, 0.0);
    lit_copy(c, up, c->st[1]);
    lit_copy(c, vp, c->st[3]);
    lit_copy(c, sshp, c->st[5]);
    lit_copy(c, u, c->st[0]);
    lit_copy(c, v, c->st[2]);
    lit_copy(c, ssh, c->st[4]);
    for (int k = 0; k < 4; ++k) fill(c, c->acc[k], 0.0);

    for (int step = 1; step <= 2 * nstep; ++step) {
        hipLaunchKernelGGL(k_continuity, grid2(g), BLK2, 0, c->stream, g, tau_inner, wflux, u, v, d[3],
                           d[6], sshp, sshn);
        cyclize(c, sshn);
        if (g.ffs > 0) {
            hipLaunchKernelGGL(k_hh_update, grid2(g), BLK2, 0, c->stream, g, sshn, d[2], d[5], d[8],
                               d[11]);
            cyclize(c, d[5]);
            cyclize(c, d[8]);
            cyclize(c, d[11]);
        }
        hipLaunchKernelGGL(k_stress, grid2(g), BLK2, 0, c->stream, g, up, vp, c->str_t, c->str_s);
        cyclize(c, c->str_t);
        cyclize(c, c->str_s);
        hipLaunchKernelGGL(k_vort, grid2(g), BLK2, 0, c->stream, g, u, v, c->vort);
        cyclize(c, c->vort);
        hipLaunchKernelGGL(k_trans_diff2, grid2(g), BLK2, 0, c->stream, g, u, v, c->vort, d[0], d[3],
                           d[6], d[9], mu, c->str_t, c->str_s, c->radv, c->rady, c->rdfx, c->rdfy);
        if (ksw4 > 0) {
            fill(c, c->fx, 0.0);
            fill(c, c->fy, 0.0);
            hipLaunchKernelGGL(k_diff4_f, grid2(g), BLK2, 0, c->stream, g, mu4, c->str_t, c->str_s,
                               d[0], d[3], d[6], d[9], c->fx, c->fy);
            cyclize(c, c->fx);
            cyclize(c, c->fy);
            hipLaunchKernelGGL(k_stress, grid2(g), BLK2, 0, c->stream, g, c->fx, c->fy, c->str_t,
                               c->str_s);
            cyclize(c, c->str_t);
            cyclize(c, c->str_s);
            hipLaunchKernelGGL(k_diff4_add, grid2(g), BLK2, 0, c->stream, g, mu4, c->str_t, c->str_s,
                               d[0], d[9], c->rdfx, c->rdfy);
        }
        MomArgs a{RHSx, RHSy, rdis, c->rdfx, c->rdfy, c->radv, c->rady, u, v, up, vp, ssh, sshn,
                  d[3], d[4], d[5], d[6], d[7], d[8], d[9], un, vn, c->acc[0], c->acc[1],
                  c->acc[2], c->acc[3]};
        hipLaunchKernelGGL(k_momentum, grid2(g), BLK2, 0, c->stream, g, tau_inner, d4n, a);
        cyclize(c, un);
        cyclize(c, vn);
        hipLaunchKernelGGL(k_shift, grid2(g), BLK2, 0, c->stream, g, dn, g.ffs > 0 ? 1 : 0, ssh, sshp,
                           sshn, u, up, un, v, vp, vn, dt);
        if (step == nstep) {
            lit_copy(c, c->st[1], up);
            lit_copy(c, c->st[3], vp);
            lit_copy(c, c->st[5], sshp);
            lit_copy(c, c->st[0], u);
            lit_copy(c, c->st[2], v);
            lit_copy(c, c->st[4], ssh);
        }
    }
    cyclize(c, c->acc[0]);
    cyclize(c, c->acc[1]);
    if (g.ffs > 0) lit_hh_init(c, c->st[4], c->st[5]);
}

// wind stress, bottom drag and RHS assembly of one 2-D model step
void lit_forcing(inmsom_bt_ctx *c)
{
    const DevGrid &g = c->g;
    hipLaunchKernelGGL(k_surface_stress, grid2(g), BLK2, 0, c->stream, g, c->taux, c->tauy, c->surf_x,
                       c->surf_y);
    hipLaunchKernelGGL(k_bottom_friction, grid2(g), BLK2, 0, c->stream, g, c->cb_nl, c->ebottom,
                       c->st[0], c->st[2], c->dep[3], c->dep[6], c->dep[0], c->bf);
    cyclize(c, c->bf);
    hipLaunchKernelGGL(k_bottom_stress_rhs, grid2(g), BLK2, 0, c->stream, g, c->bf, c->st[1], c->st[3],
                       c->surf_x, c->surf_y, c->bot_x, c->bot_y, c->rhsx, c->rhsy);
}

}  // namespace ibt
