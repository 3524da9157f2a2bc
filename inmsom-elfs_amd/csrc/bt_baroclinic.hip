// baroclinic_dynamics (Function/vel_ssh.f90:794-1027): for every open U
// (V) face the implicit vertical viscosity / vertical advection column of
// nz levels with Coriolis from the other component, Rayleigh drag and the
// surface and bottom stresses, solved by the sweep of factor8
// (Service/oc_algorithms.f90:494-524); then the Robert-Asselin step of the
// filtered level and uu = un, vv = vn.
//
// k_bc_solve: one thread per interior column (U or V; a workgroup's waves
// take the U and V columns of two rows of 64 consecutive m starting on a
// 64-column boundary, so every level's loads are whole 128-byte lines and
// the lines both components read are fetched once).  The sweep's
// coefficients are held in registers (nz fixed at compile time, or a
// generic instance with private arrays); each level's loads are issued one
// level ahead of its arithmetic, and the 3-D fields are addressed as
// buffers (a scalar level offset, one vector offset per neighbour).  The
// new level goes to a second array (uo, vo): a column's Coriolis term
// reads the old uu, vv of its neighbours, so the old level must stay until
// every column is solved.  The filtered level of a point is read by its own
// column only, so the backward sweep applies the Robert-Asselin step to it
// in place, from the old and filtered levels the forward sweep loaded
// (kept in registers for a fixed nz).  Interior points without an open face
// copy the old level.  k_bc_frame: the reference's update on the one-point
// ring around the interior (un taken at the point the periodic wrap copies
// it from, Service/oc_algorithms.f90:50-95, :118-135) and the old level
// copied on the outer frame.  The host-array entry keeps two staged level-n
// arrays and swaps them after each call; the device-address entry solves
// into a work array and copies it onto the caller's uu, vv.  Measured at
// 2048 x 1536 x 20 on MI355X: 3.06 ms per call with an un, vn work buffer
// and a separate update pass (14.7 GB moved), 1.88 ms in this form (8.2 GB).
// Reference expressions and operation order throughout (-ffp-contract=off):
// bit-identical to the reference.
#include <string>

#include "bt_internal.h"

namespace ibt {

namespace {

constexpr int BC_NZMAX = 128;  // levels of the generic instance
#ifndef BC_AHEAD
#define BC_AHEAD 1
#endif

struct BcParams {
    DevGrid g;
    int nz;
    double tau;
    int tiles_x, tiles_y;                      // 64 x 2 tiles of the solve (64-aligned columns)
    int bsx, bsy;                              // tiles per block of the solve's order
    const double *dz, *hzt;                    // nz, nz + 1 (device)
    const double *uu, *vv;                     // (nz, ny, nx) level n, read
    double *uup, *vvp;                         // (nz, ny, nx) filtered level, updated in place
    double *uo, *vo;                           // (nz, ny, nx) new level n (distinct from uu, vv)
    const double *ww, *nu;                     // (nz + 1, ny, nx)
    const double *rhsx, *rhsy, *rhsxt, *rhsyt; // (nz, ny, nx)
    const double *pgrx, *pgry, *rdis, *txs, *tys, *txb, *tyb;
    const double *hhu, *hhup, *hhun, *hhv, *hhvp, *hhvn, *hhh;
};

#define I(mm, nn) IDX(g, (mm), (nn))

// vel_ssh.f90:1006-1007 (1019-1020): the filtered level from the old and
// the new level
__device__ __forceinline__ double asselin(double u, double un, double up, double ts)
{
    return u + ts * (un - 2.0 * u + up) / 2.0;
}

typedef unsigned int u32x2 __attribute__((ext_vector_type(2)));

// One (nz, ny, nx) field of the solve, addressed by a byte offset of the
// column (per lane) and one of the level (uniform).  BUF: buffer loads and
// stores (the level's offset in a scalar register, one vector offset per
// neighbour for all levels; arrays below 2 GiB, checked by the host);
// otherwise 64-bit addresses.
template <bool BUF>
struct Field;
template <>
struct Field<true> {
    __amdgpu_buffer_rsrc_t r;
    __device__ __forceinline__ Field(const double *p, size_t bytes)
        : r(__builtin_amdgcn_make_buffer_rsrc(const_cast<double *>(p), 0, (int)bytes, 0x00020000))
    {
    }
    __device__ __forceinline__ double ld(size_t v, size_t s) const
    {
        return __builtin_bit_cast(double, __builtin_amdgcn_raw_buffer_load_b64(r, (int)v, (int)s, 0));
    }
    __device__ __forceinline__ void st(size_t v, size_t s, double x) const
    {
        __builtin_amdgcn_raw_buffer_store_b64(__builtin_bit_cast(u32x2, x), r, (int)v, (int)s, 0);
    }
};
template <>
struct Field<false> {
    char *q;
    __device__ __forceinline__ Field(const double *p, size_t) : q((char *)const_cast<double *>(p)) {}
    __device__ __forceinline__ double ld(size_t v, size_t s) const { return *(const double *)(q + s + v); }
    __device__ __forceinline__ void st(size_t v, size_t s, double x) const { *(double *)(q + s + v) = x; }
};

// The fields of one component's column: its own level n, filtered level,
// RHS pair and new level, the other component's level n (Coriolis), and
// the (nz + 1)-level ww, nu.
template <bool BUF>
struct ColFields {
    Field<BUF> w, o, wp, rx, rxt, wn, ww, nu;
};

// One column at (m, n): vel_ssh.f90:845-923 (U) or :925-998 (V), and its
// update :1006-1007 (:1019-1020).  The caller passes the byte offsets of
// the column (c0), of the cross-face neighbour (c1: m+1 for U, n+1 for V)
// and of the Coriolis points of the other component (U: vv at (m+1, n),
// (m, n), (m+1, n-1), (m, n-1); V: uu at (m, n+1), (m, n), (m-1, n+1),
// (m-1, n)), the column's 2-D coefficients, and sgn = +1 (U) or -1 (V)
// for the Coriolis term.  Operation order is the reference's.
// The 2-D inputs of one column (U: dxt, dyh, hhun, hhup at (m, n); rlh_s,
// hhh, dxb, dyb at (m, n) and (m, n-1); rdis, dx, dy at (m, n) and (m+1, n);
// hhu, pgrx, tx_surf, tx_bot at (m, n).  V: dxh, dyt, hhvn, hhvp; the
// Coriolis factors at (m, n) and (m-1, n); rdis, dx, dy at (m, n+1); hhv,
// pgry, ty_surf, ty_bot), loaded together before the sweep.
struct Col2D {
    double a0, a1, hn, hp, rl0, hh0, xb0, yb0, rl1, hh1, xb1, yb1, rd0, rd1, dx0, dy0, dx1, dy1, hw, pgr, ts, tb;
};

// One column at (m, n): vel_ssh.f90:845-923 (U) or :925-998 (V), and its
// update :1006-1007 (:1019-1020).  The caller passes the byte offsets of
// the column (c0), of the cross-face neighbour (c1: m+1 for U, n+1 for V)
// and of the Coriolis points of the other component (U: vv at (m+1, n),
// (m, n), (m+1, n-1), (m, n-1); V: uu at (m, n+1), (m, n), (m-1, n+1),
// (m-1, n)), the column's 2-D inputs, and is_u (the Coriolis term's sign).
// Operation order is the reference's.
template <int NZ, bool BUF>
__device__ __forceinline__ void solve_column(const BcParams &p, const ColFields<BUF> &F, size_t c0, size_t c1,
                                             const size_t (&co)[4], const Col2D &C, bool is_u)
{
    const double area0 = C.a0, area1 = C.a1;
    const double bp = C.hn * area0 * area1 / p.tau / 2.0;
    const double bp0 = C.hp * area0 * area1 / p.tau / 2.0;
    const double f0 = C.rl0 * C.hh0 * C.xb0 * C.yb0;
    const double f1 = C.rl1 * C.hh1 * C.xb1 * C.yb1;
    const double rd = (C.rd0 + C.rd1) / 2.0;
    // ww * dx * dy is (ww * dx) * dy: the metrics are not pre-multiplied
    const double dx0 = C.dx0, dy0 = C.dy0, dx1 = C.dx1, dy1 = C.dy1;
    const double hn = C.hn, hw = C.hw, pgr = C.pgr, ts_ = C.ts, tb_ = C.tb;
    const int nz = NZ > 0 ? NZ : p.nz;
    const size_t pl8 = (size_t)p.g.nx * p.g.ny * sizeof(double);
    constexpr int NA = NZ > 0 ? NZ : BC_NZMAX;
    constexpr int UR = NZ > 0 ? NZ : 1;  // full unrolling for a fixed nz
    constexpr bool HOLD = NZ > 0;        // old and filtered levels kept for the update
    double x[NA], y[NA];
    double w_old[HOLD ? NA : 1], wp_old[HOLD ? NA : 1];
    // the loads of level k (its k + 1 parts for k < nz), issued BC_AHEAD
    // levels before their use: the sweep's serial chain of divides then
    // waits on no memory latency
    struct LevelIn {
        double nu0, nu1, w1, ww0, ww1, wp, rx, rxt, o0, o1, o2, o3, dz, hzt;
    };
    auto load = [&](int k) {
        LevelIn L{};
        if (k > nz) return L;
        const size_t o = (size_t)(k - 1) * pl8;
        L.dz = p.dz[k - 1];
        if (k < nz) {
            const size_t o1 = o + pl8;
            L.hzt = p.hzt[k];
            L.nu0 = F.nu.ld(c0, o1);
            L.nu1 = F.nu.ld(c1, o1);
            L.w1 = F.w.ld(c0, o1);
            L.ww0 = F.ww.ld(c0, o1);
            L.ww1 = F.ww.ld(c1, o1);
        }
        L.wp = F.wp.ld(c0, o);
        L.rx = F.rx.ld(c0, o);
        L.rxt = F.rxt.ld(c0, o);
        L.o0 = F.o.ld(co[0], o);
        L.o1 = F.o.ld(co[1], o);
        L.o2 = F.o.ld(co[2], o);
        L.o3 = F.o.ld(co[3], o);
        return L;
    };
    LevelIn q[BC_AHEAD + 1];
    double wk = F.w.ld(c0, 0);
#pragma unroll
    for (int j = 0; j < BC_AHEAD; ++j) q[j] = load(1 + j);
    double dp = 0.0, fz_p = 0.0;
#pragma unroll UR
    for (int k = 1; k <= NA; ++k) {
        if (k > nz) break;
        q[BC_AHEAD] = load(k + BC_AHEAD);
        __builtin_amdgcn_sched_barrier(0);  // the loads go out before this level's arithmetic
        const LevelIn &L = q[0];
        const double dm = (k == 1) ? 0.0 : dp;
        const double fz_m = (k == 1) ? 0.0 : fz_p;
        const double dzk = L.dz;
        if constexpr (HOLD) w_old[k - 1] = wk;
        if (k < nz) {
            dp = (L.nu0 + L.nu1) / 2.0 / hn / L.hzt * area0 * area1;
            const double wk1 = L.w1;
            fz_p = (L.ww0 * dx0 * dy0 + L.ww1 * dx1 * dy1) / 2.0 * (wk + wk1) / 2.0;
            wk = wk1;
        } else {
            dp = 0.0;
            fz_p = 0.0;
        }
        const double c = (k < nz) ? -(dp / dzk) : 0.0;
        const double a = (k == 1) ? 0.0 : -(dm / dzk);
        const double b = (k == 1) ? bp + dp / dzk : (k == nz) ? bp + dm / dzk : bp + (dp + dm) / dzk;
        const double wp = L.wp;
        if constexpr (HOLD) wp_old[k - 1] = wp;
        const double cor = (f0 * (L.o0 + L.o1) + f1 * (L.o2 + L.o3)) / 4.0;
        double e = bp0 * wp + L.rx + L.rxt + pgr - (fz_p - fz_m) / dzk;
        e = is_u ? e + cor : e - cor;
        e = e - rd * wp * area0 * area1 * hw;
        if (k == 1) e = e + ts_ / dzk * area0 * area1;
        if (k == nz) e = e + tb_ / dzk * area0 * area1;
        // factor8 forward sweep
        if (k == 1) {
            x[0] = -(c / b);
            y[0] = e / b;
        } else {
            const double scr = 1.0 / (a * x[k - 2] + b);
            x[k - 1] = -(c * scr);
            y[k - 1] = (e - a * y[k - 2]) * scr;
        }
#pragma unroll
        for (int j = 0; j < BC_AHEAD; ++j) q[j] = q[j + 1];
        __builtin_amdgcn_sched_barrier(0);
    }
    // backward sweep, each level's update as soon as its new value is known
    const double ts = p.g.ts;
    double r = y[nz - 1];
    {
        const size_t o = (size_t)(nz - 1) * pl8;
        F.wp.st(c0, o, asselin(HOLD ? w_old[nz - 1] : F.w.ld(c0, o), r, HOLD ? wp_old[nz - 1] : F.wp.ld(c0, o), ts));
        F.wn.st(c0, o, r);
    }
#pragma unroll UR
    for (int k = NA - 1; k >= 1; --k) {
        if (k > nz - 1) continue;
        r = x[k - 1] * r + y[k - 1];
        const size_t o = (size_t)(k - 1) * pl8;
        F.wp.st(c0, o, asselin(HOLD ? w_old[k - 1] : F.w.ld(c0, o), r, HOLD ? wp_old[k - 1] : F.wp.ld(c0, o), ts));
        F.wn.st(c0, o, r);
    }
}

__device__ __forceinline__ void copy_column(const double *src, double *dst, size_t i, size_t pl, int nz)
{
    for (int k = 0; k < nz; ++k) dst[(size_t)k * pl + i] = src[(size_t)k * pl + i];
}

// Tiles in blocks of bsx x bsy = 64 tiles (bsx = min(16, tiles_x); one
// XCD's workgroups in flight at once: 2 per CU x 32 CUs), the blocks in
// row-major order and each XCD given a contiguous run of them: the rows
// above and below a tile and the columns beside it (the neighbours the
// columns read) are mostly fetched by a tile of the same block, in the
// same L2 at about the same time.  Measured at 2048 x 1536 x 20: FETCH_SIZE
// 6.08 GB per call against 6.86 GB in plain row-major order, 1 % slower.
__device__ __forceinline__ bool bc_tile(const BcParams &p, int b, int &bx, int &by)
{
    const int nbx = (p.tiles_x + p.bsx - 1) / p.bsx, nby = (p.tiles_y + p.bsy - 1) / p.bsy;
    int sx, sy;
    xcd_tile(nbx * p.bsx, nby * p.bsy, b, sx, sy);  // a position in the padded tile space ...
    const int pos = sy * nbx * p.bsx + sx;            // ... in XCD-contiguous order
    const int blk = pos / (p.bsx * p.bsy), w = pos - blk * (p.bsx * p.bsy);
    bx = (blk % nbx) * p.bsx + w % p.bsx;
    by = (blk / nbx) * p.bsy + w / p.bsx;
    return bx < p.tiles_x && by < p.tiles_y;
}

// A workgroup takes a 64 x 2 tile: waves 0 and 2 the U columns of its two
// rows, waves 1 and 3 the V columns, so the level-k lines of uu, vv, ww
// and nu that both components read are fetched once and meet in the cache.
template <int NZ, bool BUF>
__global__ __launch_bounds__(256, 2) void k_bc_solve(BcParams p)
{
    const DevGrid &g = p.g;
    int bx, by;
    if (!bc_tile(p, (int)blockIdx.x, bx, by)) return;
    const int w = __builtin_amdgcn_readfirstlane((int)(threadIdx.x >> 6));
    const int m = bx * 64 + (int)(threadIdx.x & 63) + 1;
    const int n = by * 2 + (w >> 1) + 3;
    if (m < 3 || m > g.nx - 2 || n > g.ny - 2) return;
    const size_t pl = (size_t)g.nx * g.ny, i = I(m, n);
    const size_t b3 = (size_t)p.nz * pl * sizeof(double), b4 = b3 + pl * sizeof(double);
    const size_t c0 = i * sizeof(double), row = (size_t)g.nx * sizeof(double), e8 = sizeof(double);
    if ((w & 1) == 0) {
        if (!(g.lcu[i] > 0.5f)) {
            copy_column(p.uu, p.uo, i, pl, p.nz);
            return;
        }
        // U column: vel_ssh.f90:845-923
        const size_t ie = I(m + 1, n), is = I(m, n - 1);
        const ColFields<BUF> F{{p.uu, b3}, {p.vv, b3}, {p.uup, b3}, {p.rhsx, b3}, {p.rhsxt, b3},
                               {p.uo, b3}, {p.ww, b4}, {p.nu, b4}};
        const size_t co[4] = {c0 + e8, c0, c0 - row + e8, c0 - row};
        const Col2D C{g.dxt[i],  g.dyh[i],  p.hhun[i], p.hhup[i], g.rlh_s[i], p.hhh[i], g.dxb[i], g.dyb[i],
                      g.rlh_s[is], p.hhh[is], g.dxb[is], g.dyb[is], p.rdis[i], p.rdis[ie], g.dx[i], g.dy[i],
                      g.dx[ie],  g.dy[ie],  p.hhu[i],  p.pgrx[i], p.txs[i],  p.txb[i]};
        __builtin_amdgcn_sched_barrier(0);  // every 2-D load in flight at once
        solve_column<NZ, BUF>(p, F, c0, c0 + e8, co, C, true);
    } else {
        if (!(g.lcv[i] > 0.5f)) {
            copy_column(p.vv, p.vo, i, pl, p.nz);
            return;
        }
        // V column: vel_ssh.f90:925-998
        const size_t in_ = I(m, n + 1), iw = I(m - 1, n);
        const ColFields<BUF> F{{p.vv, b3}, {p.uu, b3}, {p.vvp, b3}, {p.rhsy, b3}, {p.rhsyt, b3},
                               {p.vo, b3}, {p.ww, b4}, {p.nu, b4}};
        const size_t co[4] = {c0 + row, c0, c0 + row - e8, c0 - e8};
        const Col2D C{g.dxh[i],  g.dyt[i],  p.hhvn[i], p.hhvp[i], g.rlh_s[i], p.hhh[i], g.dxb[i], g.dyb[i],
                      g.rlh_s[iw], p.hhh[iw], g.dxb[iw], g.dyb[iw], p.rdis[i], p.rdis[in_], g.dx[i], g.dy[i],
                      g.dx[in_], g.dy[in_], p.hhv[i],  p.pgry[i], p.tys[i],  p.tyb[i]};
        __builtin_amdgcn_sched_barrier(0);
        solve_column<NZ, BUF>(p, F, c0, c0 + row, co, C, false);
    }
}

// The points around the interior: rows 1, 2, ny-1, ny and columns 1, 2,
// nx-1, nx.  On the ring (m = 2 or nx-1, n = 2 or ny-1) the update of
// vel_ssh.f90:1000-1023 with un (vn) zero, or on periodic boundaries the
// solved value at the point the wrap copies from; elsewhere, and where the
// face is closed, the old level.
__global__ __launch_bounds__(256) void k_bc_frame(BcParams p)
{
    const DevGrid &g = p.g;
    const int nx = g.nx, ny = g.ny;
    const int t = (int)(blockIdx.x * blockDim.x + threadIdx.x), k = (int)blockIdx.y;  // point, level
    int m, n;
    if (t < 4 * nx) {
        const int j = t / nx;
        n = j == 0 ? 1 : j == 1 ? 2 : j == 2 ? ny - 1 : ny;
        m = t - j * nx + 1;
    } else {
        const int u = t - 4 * nx, h = ny - 4;
        if (u >= 4 * h) return;
        const int j = u / h;
        m = j == 0 ? 1 : j == 1 ? 2 : j == 2 ? nx - 1 : nx;
        n = u - j * h + 3;
    }
    const size_t pl = (size_t)nx * ny, i = I(m, n), o = (size_t)k * pl;
    const bool ring = m >= 2 && m <= nx - 1 && n >= 2 && n <= ny - 1;
    const int mm = (g.px && m == 2) ? nx - 2 : (g.px && m == nx - 1) ? 3 : m;
    const int nn = (g.py && n == 2) ? ny - 2 : (g.py && n == ny - 1) ? 3 : n;
    const bool src_in = mm >= 3 && mm <= nx - 2 && nn >= 3 && nn <= ny - 2;
    const size_t is = I(mm, nn);
    const double ts = g.ts;
    if (ring && g.lcu[i] > 0.5f) {
        const double un = (src_in && g.lcu[is] > 0.5f) ? p.uo[o + is] : 0.0;
        p.uup[o + i] = asselin(p.uu[o + i], un, p.uup[o + i], ts);
        p.uo[o + i] = un;
    } else {
        p.uo[o + i] = p.uu[o + i];
    }
    if (ring && g.lcv[i] > 0.5f) {
        const double vn = (src_in && g.lcv[is] > 0.5f) ? p.vo[o + is] : 0.0;
        p.vvp[o + i] = asselin(p.vv[o + i], vn, p.vvp[o + i], ts);
        p.vo[o + i] = vn;
    } else {
        p.vo[o + i] = p.vv[o + i];
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
    c->bc_cur = 0;
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

// the staged level n of uu, vv (slot pair UU / UN and VV / VN, by bc_cur)
double *staged_u(inmsom_bt_ctx *c, int nz, bool next) { return s3(c, nz, (c->bc_cur != next) ? S_UN : S_UU); }
double *staged_v(inmsom_bt_ctx *c, int nz, bool next) { return s3(c, nz, (c->bc_cur != next) ? S_VN : S_VV); }

// the staged fields of one call: level n read from the current slots, the
// new level written to the others
BcParams staged(inmsom_bt_ctx *c, int nz, double tau)
{
    BcParams p{};
    p.g = c->g;
    p.nz = nz;
    p.tau = tau;
    p.dz = svert(c, nz);
    p.hzt = svert(c, nz) + nz + 1;
    p.uu = staged_u(c, nz, false);
    p.uo = staged_u(c, nz, true);
    p.uup = s3(c, nz, S_UUP);
    p.vv = staged_v(c, nz, false);
    p.vo = staged_v(c, nz, true);
    p.vvp = s3(c, nz, S_VVP);
    p.rhsx = s3(c, nz, S_RX);
    p.rhsy = s3(c, nz, S_RY);
    p.rhsxt = s3(c, nz, S_RXT);
    p.rhsyt = s3(c, nz, S_RYT);
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
    q.tiles_x = (c->nx + 63) / 64;
    q.tiles_y = (c->ny - 4 + 1) / 2;
    q.bsx = q.tiles_x < 16 ? q.tiles_x : 16;
    q.bsy = 64 / q.bsx;
    const int nbx = (q.tiles_x + q.bsx - 1) / q.bsx, nby = (q.tiles_y + q.bsy - 1) / q.bsy;
    const dim3 gs((unsigned)(nbx * nby * q.bsx * q.bsy)), bs(256);
    // buffer addressing while the (nz + 1)-level fields stay below 2 GiB
    const bool buf = (size_t)(q.nz + 1) * c->n * sizeof(double) < ((size_t)1 << 31);
    if (q.nz == 20 && buf)
        hipLaunchKernelGGL((k_bc_solve<20, true>), gs, bs, 0, c->stream, q);
    else if (q.nz == 20)
        hipLaunchKernelGGL((k_bc_solve<20, false>), gs, bs, 0, c->stream, q);
    else
        hipLaunchKernelGGL((k_bc_solve<0, false>), gs, bs, 0, c->stream, q);
    const int frame = 4 * c->nx + 4 * (c->ny - 4);
    hipLaunchKernelGGL(k_bc_frame, dim3((unsigned)((frame + 255) / 256), (unsigned)q.nz), dim3(256), 0, c->stream,
                       q);
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
    double *slot3[8] = {staged_u(c, nz, false), s3(c, nz, S_UUP), staged_v(c, nz, false), s3(c, nz, S_VVP),
                        s3(c, nz, S_RX),        s3(c, nz, S_RY),  s3(c, nz, S_RXT),         s3(c, nz, S_RYT)};
    for (int k = 0; k < 8; ++k) CHB(put(slot3[k], in3[k], (size_t)nz * pl));
    CHB(put(sw(c, nz, S_WW), ww, (size_t)(nz + 1) * pl));
    CHB(put(sw(c, nz, S_NU), nu, (size_t)(nz + 1) * pl));
    for (int k = 0; k < S_N2; ++k) CHB(put(s2(c, nz, k), in2[k], pl));
    CHB(put(svert(c, nz), dz, nz));
    CHB(put(svert(c, nz) + nz + 1, hzt, nz + 1));
    const BcParams p = staged(c, nz, tau);
    launch(c, p);
    CHB(hipGetLastError());
    c->bc_cur ^= 1;
    c->bc_nz = nz;
    double *out[4] = {uu, uup, vv, vvp};
    const double *from[4] = {staged_u(c, nz, false), s3(c, nz, S_UUP), staged_v(c, nz, false), s3(c, nz, S_VVP)};
    for (int k = 0; k < 4; ++k)
        CHB(hipMemcpyAsync(out[k], from[k], (size_t)nz * pl * sizeof(double), hipMemcpyDeviceToHost, c->stream));
    CHB(hipStreamSynchronize(c->stream));
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
    // the new uu, vv and the vertical grid
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
    p.uo = c->bc_dev;
    p.vo = c->bc_dev + (size_t)nz * c->n;
    launch(c, p);
    CHB(hipGetLastError());
    // every point of uo, vo was written: the new level onto the caller's arrays
    const size_t bytes = (size_t)nz * c->n * sizeof(double);
    CHB(hipMemcpyAsync(uu, p.uo, bytes, hipMemcpyDeviceToDevice, c->stream));
    CHB(hipMemcpyAsync(vv, p.vo, bytes, hipMemcpyDeviceToDevice, c->stream));
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
    hipEvent_t e0, e1;
    CHB(hipEventCreate(&e0));
    CHB(hipEventCreate(&e1));
    CHB(hipEventRecord(e0, c->stream));
    for (int r = 0; r < nrep; ++r) {
        launch(c, staged(c, nz, tau));
        c->bc_cur ^= 1;
    }
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
