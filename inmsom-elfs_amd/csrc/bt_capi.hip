// C ABI of inmsom_bt (declarations and contracts in include/inmsom_bt.h).
#include <cmath>
#include <cstring>
#include <string>
#include <vector>

#include "bt_dropin.h"
#include "bt_fused.h"
#include "bt_internal.h"
#include "bt_literal.h"

namespace ibt {

static thread_local std::string g_err;
void set_error(const std::string &msg) { g_err = msg; }

#define HIPCHECK(expr)                                                                  \
    do {                                                                                \
        hipError_t e_ = (expr);                                                         \
        if (e_ != hipSuccess) {                                                         \
            set_error(std::string(#expr) + ": " + hipGetErrorString(e_));               \
            return INMSOM_BT_EDEVICE;                                                   \
        }                                                                               \
    } while (0)

#define REQUIRE(cond, msg)                      \
    do {                                        \
        if (!(cond)) {                          \
            set_error(msg);                     \
            return INMSOM_BT_EINVAL;            \
        }                                       \
    } while (0)

static int dalloc(inmsom_bt_ctx *c, size_t bytes, void **p)
{
    hipError_t e = hipMalloc(p, bytes);
    if (e != hipSuccess) {
        set_error(std::string("hipMalloc: ") + hipGetErrorString(e));
        return INMSOM_BT_ENOMEM;
    }
    e = hipMemsetAsync(*p, 0, bytes, c->stream);
    if (e != hipSuccess) {
        set_error(std::string("hipMemset: ") + hipGetErrorString(e));
        return INMSOM_BT_EDEVICE;
    }
    return INMSOM_BT_OK;
}

#define DALLOC(ptr, bytes)                                                 \
    do {                                                                   \
        int r_ = dalloc(c, (bytes), reinterpret_cast<void **>(&(ptr)));    \
        if (r_ != INMSOM_BT_OK) return r_;                                 \
    } while (0)

static int upload(inmsom_bt_ctx *c, void *dst, const void *src, size_t bytes)
{
    HIPCHECK(hipMemcpyAsync(dst, src, bytes, hipMemcpyHostToDevice, c->stream));
    return INMSOM_BT_OK;
}

#define TRY(expr)                          \
    do {                                   \
        int r_ = (expr);                   \
        if (r_ != INMSOM_BT_OK) return r_; \
    } while (0)

// Service/grid_construction.f90:33-133 on one process.  Periodic wrap is
// the copy of a decomposed run (column mm -> mmm-1, mmm -> mm+1).
static void wrap_x(int nx, int ny, float *f)
{
    for (int n = 0; n < ny; ++n) {
        float a = f[(size_t)n * nx + nx - 3], b = f[(size_t)n * nx + 2];
        f[(size_t)n * nx + 1] = a;
        f[(size_t)n * nx + nx - 2] = b;
    }
}
static void wrap_y(int nx, int ny, float *f)
{
    for (int m = 0; m < nx; ++m) {
        float a = f[(size_t)(ny - 3) * nx + m], b = f[(size_t)2 * nx + m];
        f[(size_t)nx + m] = a;
        f[(size_t)(ny - 2) * nx + m] = b;
    }
}

static void gridcon(int nx, int ny, int px, int py, const int *lb, float *const mk[7])
{
    const size_t sz = (size_t)nx * ny;
    for (int k = 0; k < 7; ++k) std::memset(mk[k], 0, sz * sizeof(float));
    float *lu = mk[0], *luu = mk[1], *luh = mk[2], *lcu = mk[3], *lcv = mk[4], *llu = mk[5],
          *llv = mk[6];
    auto J = [nx](int m, int n) { return (size_t)(n - 1) * nx + (m - 1); };
    for (size_t i = 0; i < sz; ++i)
        if (lb[i] == 0) lu[i] = 1.0f;
    if (px) wrap_x(nx, ny, lu);
    if (py) wrap_y(nx, ny, lu);
    for (int n = 2; n <= ny - 2; ++n)
        for (int m = 2; m <= nx - 2; ++m) {
            float a = lu[J(m, n)], b = lu[J(m + 1, n)], cc = lu[J(m, n + 1)], d = lu[J(m + 1, n + 1)];
            if (a + b + cc + d > 0.5f) luh[J(m, n)] = 1.0f;
            if (a * b * cc * d > 0.5f) luu[J(m, n)] = 1.0f;
        }
    if (px) { wrap_x(nx, ny, luh); wrap_x(nx, ny, luu); }
    if (py) { wrap_y(nx, ny, luh); wrap_y(nx, ny, luu); }
    for (int n = 2; n <= ny - 2; ++n)
        for (int m = 2; m <= nx - 2; ++m) {
            float a = lu[J(m, n)], b = lu[J(m + 1, n)], cc = lu[J(m, n + 1)];
            if (a + b > 0.5f) llu[J(m, n)] = 1.0f;
            if (a + cc > 0.5f) llv[J(m, n)] = 1.0f;
            if (a * b > 0.5f) lcu[J(m, n)] = 1.0f;
            if (a * cc > 0.5f) lcv[J(m, n)] = 1.0f;
        }
    for (float *f : {lcu, llu, lcv, llv}) {
        if (px) wrap_x(nx, ny, f);
        if (py) wrap_y(nx, ny, f);
    }
}

static bool all_equal(const double *a, size_t n, double *v)
{
    for (size_t i = 1; i < n; ++i)
        if (a[i] != a[0]) return false;
    *v = a[0];
    return true;
}

}  // namespace ibt

using namespace ibt;

extern "C" {

const char *inmsom_bt_last_error(void) { return g_err.c_str(); }

int inmsom_bt_gridcon(int nx, int ny, int px, int py, const int *lb, float *lu, float *luu,
                      float *luh, float *lcu, float *lcv, float *llu, float *llv)
{
    REQUIRE(nx >= 5 && ny >= 5, "nx, ny must be at least 5 (interior plus two halo cells per side)");
    REQUIRE(lb && lu && luu && luh && lcu && lcv && llu && llv, "null pointer");
    float *mk[7] = {lu, luu, luh, lcu, lcv, llu, llv};
    gridcon(nx, ny, px, py, lb, mk);
    return INMSOM_BT_OK;
}

// Control/init_pars.f90:459-481
int inmsom_bt_rayleigh(const inmsom_bt_grid *gr, double r_fric, double *r_diss)
{
    REQUIRE(gr && r_diss && gr->hhq_rest && gr->dxt && gr->dyt, "null pointer");
    const int nx = gr->nx, ny = gr->ny;
    std::vector<float> own[7];
    const float *lu = gr->lu, *lcu = gr->lcu, *lcv = gr->lcv;
    if (!lu) {
        REQUIRE(gr->lbasins, "either lu.. or lbasins must be given");
        float *mk[7];
        for (int k = 0; k < 7; ++k) {
            own[k].assign((size_t)nx * ny, 0.0f);
            mk[k] = own[k].data();
        }
        gridcon(nx, ny, gr->periodicity_x, gr->periodicity_y, gr->lbasins, mk);
        lu = mk[0];
        lcu = mk[3];
        lcv = mk[4];
    }
    const double *h = gr->hhq_rest, *dxt = gr->dxt, *dyt = gr->dyt;
    auto J = [nx](int m, int n) { return (size_t)(n - 1) * nx + (m - 1); };
    for (int n = 3; n <= ny - 2; ++n)
        for (int m = 3; m <= nx - 2; ++m) {
            if (!(lu[J(m, n)] > 0.5f)) continue;
            double a = (h[J(m + 1, n)] - h[J(m, n)]) / dxt[J(m, n)];
            double b = (h[J(m, n)] - h[J(m - 1, n)]) / dxt[J(m - 1, n)];
            double hx2 = (a * a * (double)lcu[J(m, n)] + b * b * (double)lcu[J(m - 1, n)]) /
                         (double)(lcu[J(m, n)] + lcu[J(m - 1, n)]);
            double cc = (h[J(m, n + 1)] - h[J(m, n)]) / dyt[J(m, n)];
            double e = (h[J(m, n)] - h[J(m, n - 1)]) / dyt[J(m, n - 1)];
            double hy2 = (cc * cc * (double)lcv[J(m, n)] + e * e * (double)lcv[J(m, n - 1)]) /
                         (double)(lcv[J(m, n)] + lcv[J(m, n - 1)]);
            r_diss[J(m, n)] = r_fric * std::sqrt(hx2 + hy2);
        }
    if (gr->periodicity_x)
        for (int n = 1; n <= ny; ++n) {
            double a = r_diss[J(nx - 2, n)], b = r_diss[J(3, n)];
            r_diss[J(2, n)] = a;
            r_diss[J(nx - 1, n)] = b;
        }
    if (gr->periodicity_y)
        for (int m = 1; m <= nx; ++m) {
            double a = r_diss[J(m, ny - 2)], b = r_diss[J(m, 3)];
            r_diss[J(m, 2)] = a;
            r_diss[J(m, ny - 1)] = b;
        }
    return INMSOM_BT_OK;
}

int inmsom_bt_init(const inmsom_bt_grid *gr, int device, inmsom_bt_ctx **out)
{
    REQUIRE(gr && out, "null pointer");
    REQUIRE(gr->nx >= 5 && gr->ny >= 5, "nx, ny must be at least 5");
    REQUIRE(gr->dx && gr->dy && gr->dxt && gr->dyt && gr->dxh && gr->dyh && gr->dxb && gr->dyb,
            "metric pointer is null");
    REQUIRE(gr->rlh_s && gr->hhq_rest, "rlh_s / hhq_rest is null");
    REQUIRE(gr->lu || gr->lbasins, "either the masks or lbasins must be given");
    REQUIRE(gr->time_smooth >= 0.0 && gr->ref_den > 0.0, "bad physical constants");
    int ndev = 0;
    HIPCHECK(hipGetDeviceCount(&ndev));
    REQUIRE(device >= 0 && device < ndev, "device index out of range");
    HIPCHECK(hipSetDevice(device));

    auto *c = new inmsom_bt_ctx();
    *out = nullptr;
    c->device = device;
    c->nx = gr->nx;
    c->ny = gr->ny;
    c->n = (size_t)gr->nx * gr->ny;
    {
        hipError_t e = hipStreamCreateWithFlags(&c->stream, hipStreamNonBlocking);
        if (e != hipSuccess) {
            set_error(std::string("hipStreamCreate: ") + hipGetErrorString(e));
            delete c;
            return INMSOM_BT_EDEVICE;
        }
    }
    const size_t n = c->n, bd = n * sizeof(double), bf = n * sizeof(float);
    int rc = [&]() -> int {
        // masks
        std::vector<float> hm[7];
        const float *src[7] = {gr->lu, gr->luu, gr->luh, gr->lcu, gr->lcv, gr->llu, gr->llv};
        if (!gr->lu) {
            float *mk[7];
            for (int k = 0; k < 7; ++k) {
                hm[k].assign(n, 0.0f);
                mk[k] = hm[k].data();
                src[k] = mk[k];
            }
            gridcon(gr->nx, gr->ny, gr->periodicity_x, gr->periodicity_y, gr->lbasins, mk);
        } else {
            for (int k = 0; k < 7; ++k) REQUIRE(src[k], "mask pointer is null");
        }
        for (int k = 0; k < 7; ++k) {
            DALLOC(c->mask[k], bf);
            TRY(upload(c, c->mask[k], src[k], bf));
        }
        const double *met[8] = {gr->dx, gr->dy, gr->dxt, gr->dyt, gr->dxh, gr->dyh, gr->dxb, gr->dyb};
        bool uni = true;
        double v0 = 0.0;
        for (int k = 0; k < 8; ++k) {
            DALLOC(c->metric[k], bd);
            TRY(upload(c, c->metric[k], met[k], bd));
            double v;
            if (!all_equal(met[k], n, &v) || (k > 0 && v != v0)) uni = false;
            if (k == 0) v0 = v;
        }
        c->uniform_metrics = uni;
        c->metric0 = v0;
        DALLOC(c->rlh_s, bd);
        TRY(upload(c, c->rlh_s, gr->rlh_s, bd));
        DALLOC(c->hr, bd);
        TRY(upload(c, c->hr, gr->hhq_rest, bd));

        for (auto **p : {&c->taux, &c->tauy, &c->wflux, &c->mu, &c->mu4, &c->rdis}) DALLOC(*p, bd);
        for (int k = 0; k < 6; ++k) DALLOC(c->st[k], bd);
        for (int k = 0; k < 9; ++k) DALLOC(c->w[k], bd);
        for (int k = 0; k < 4; ++k) DALLOC(c->acc[k], bd);
        for (int k = 0; k < 12; ++k) DALLOC(c->dep[k], bd);
        for (auto **p : {&c->vort, &c->str_t, &c->str_s, &c->fx, &c->fy, &c->radv, &c->rady, &c->rdfx,
                         &c->rdfy, &c->rhsx, &c->rhsy, &c->bf, &c->surf_x, &c->surf_y, &c->bot_x,
                         &c->bot_y})
            DALLOC(*p, bd);
        DALLOC(c->scal, SCAL_DOUBLES * sizeof(double));
        DALLOC(c->cnt, 8 * sizeof(long long));
        TRY(upload(c, c->scal + DEP_TABLE_OFFSET, c->dep, sizeof(c->dep)));

        DevGrid &g = c->g;
        g.nx = gr->nx;
        g.ny = gr->ny;
        g.px = gr->periodicity_x;
        g.py = gr->periodicity_y;
        g.ffs = gr->full_free_surface;
        g.g = gr->free_fall_acc;
        g.rho0 = gr->ref_den;
        g.ts = gr->time_smooth;
        g.lu = c->mask[0];
        g.luu = c->mask[1];
        g.luh = c->mask[2];
        g.lcu = c->mask[3];
        g.lcv = c->mask[4];
        g.llu = c->mask[5];
        g.llv = c->mask[6];
        g.dx = c->metric[0];
        g.dy = c->metric[1];
        g.dxt = c->metric[2];
        g.dyt = c->metric[3];
        g.dxh = c->metric[4];
        g.dyh = c->metric[5];
        g.dxb = c->metric[6];
        g.dyb = c->metric[7];
        g.rlh_s = c->rlh_s;
        g.hr = c->hr;
        // depths of a state at rest until set_state is called
        lit_hh_init(c, c->st[4], c->st[5]);
        HIPCHECK(hipStreamSynchronize(c->stream));
        // what the fused barotropic_dynamics needs of the static grid, and the
        // address of the allocation-free entry's scratch on this device
        TRY(dropin_setup(c));
        TRY(noalloc_scratch_setup(device));
        return INMSOM_BT_OK;
    }();
    if (rc != INMSOM_BT_OK) {
        inmsom_bt_finalize(c);
        return rc;
    }
    *out = c;
    return INMSOM_BT_OK;
}

int inmsom_bt_finalize(inmsom_bt_ctx *c)
{
    if (!c) return INMSOM_BT_OK;
    hipSetDevice(c->device);
    if (c->stream) hipStreamSynchronize(c->stream);
    fused_release(c);
    dropin_release(c);
    auto fr = [](void *p) {
        if (p) hipFree(p);
    };
    for (auto *p : c->mask) fr(p);
    for (auto *p : c->metric) fr(p);
    for (auto *p : c->st) fr(p);
    for (auto *p : c->w) fr(p);
    for (auto *p : c->acc) fr(p);
    for (auto *p : c->dep) fr(p);
    for (auto *p : c->st_save) fr(p);
    for (auto *p : c->st_snap) fr(p);
    for (void *p : {(void *)c->rlh_s, (void *)c->hr, (void *)c->taux, (void *)c->tauy,
                    (void *)c->wflux, (void *)c->mu, (void *)c->mu4, (void *)c->rdis,
                    (void *)c->vort, (void *)c->str_t, (void *)c->str_s, (void *)c->fx,
                    (void *)c->fy, (void *)c->radv, (void *)c->rady, (void *)c->rdfx,
                    (void *)c->rdfy, (void *)c->rhsx, (void *)c->rhsy, (void *)c->bf,
                    (void *)c->surf_x, (void *)c->surf_y, (void *)c->bot_x, (void *)c->bot_y,
                    (void *)c->scal, (void *)c->cnt, (void *)c->cpl, (void *)c->cpl_new, (void *)c->f3,
                    (void *)c->f3_mb, (void *)c->f3_halo, (void *)c->f3_dz,
                    (void *)c->snap_dev, (void *)c->bc, (void *)c->bc_dev, (void *)c->vv, (void *)c->vv_dz,
                    (void *)c->pg, (void *)c->pg_z, (void *)c->eos, (void *)c->eos_z,
                    (void *)c->sm, (void *)c->sm_dz, (void *)c->sst})
        fr(p);
    if (c->snap_host) hipHostFree(c->snap_host);
    if (c->stream) hipStreamDestroy(c->stream);
    delete c;
    return INMSOM_BT_OK;
}

// Which path takes a barotropic_dynamics call: the fused drop-in where the
// context allows it and the caller did not ask for the stage-by-stage path.
static int dropin_route(inmsom_bt_ctx *c, int ksw4, bool *fused)
{
    std::string why;
    const bool ok = c->dropin_path != 1 && dropin_eligible(c, ksw4, &why);
    if (c->dropin_path == 2 && !ok) {
        set_error("fused barotropic_dynamics requested but not available: " + why);
        return INMSOM_BT_ESTATE;
    }
    *fused = ok;
    return INMSOM_BT_OK;
}

// dep[] back to hh_init of the context's state after fused drop-in calls
// changed it (the stage-by-stage path reads dep[] as the reference reads its
// module depths)
static void refresh_depths(inmsom_bt_ctx *c)
{
    if (!c->dep_stale) return;
    lit_hh_init(c, c->st[4], c->st[5]);
    c->dep_stale = false;
}

// One call on the context's device arrays (state st[], averages acc[]) with
// the inputs at the given device addresses, by either path.
static int dynamics_on_context(inmsom_bt_ctx *c, double tau, int nstep, int ksw4, const double *rx,
                               const double *ry, const double *wf, const double *mu, const double *mu4,
                               const double *rd)
{
    bool fused = false;
    TRY(dropin_route(c, ksw4, &fused));
    if (fused) {
        TRY(dropin_call(c, tau, nstep, c->st, c->acc, rx, ry, wf, mu, rd, nullptr, ksw4, mu4));
        c->dep_stale = true;
    } else {
        // the depths a reference caller holds at entry are hh_init(ssh_e, sshp_e):
        // every call ends with hh_init (vel_ssh.f90:668-675), as does set-up
        // (init_pars.f90:426-430)
        lit_hh_init(c, c->st[4], c->st[5]);
        c->dep_stale = false;
        lit_barotropic_dynamics(c, tau, nstep, ksw4, rx, ry, wf, mu, mu4, rd);
    }
    HIPCHECK(hipGetLastError());
    return INMSOM_BT_OK;
}

int inmsom_bt_barotropic_dynamics(inmsom_bt_ctx *c, double tau, int nstep, int ksw4,
                                  double *ubrtr_e, double *ubrtrp_e, double *vbrtr_e,
                                  double *vbrtrp_e, double *ssh_e, double *sshp_e,
                                  double *ubrtr_i, double *vbrtr_i, double *ssh4gradx,
                                  double *ssh4grady, const double *RHSx, const double *RHSy,
                                  const double *wflux, const double *mu, const double *mu4,
                                  const double *rdis)
{
    REQUIRE(c, "null context");
    REQUIRE(nstep >= 1 && nstep <= 1000, "nstep must be in 1..1000");
    REQUIRE(tau > 0.0, "tau must be positive");
    REQUIRE(ubrtr_e && ubrtrp_e && vbrtr_e && vbrtrp_e && ssh_e && sshp_e && ubrtr_i && vbrtr_i &&
                ssh4gradx && ssh4grady && RHSx && RHSy && wflux && mu && rdis,
            "null array pointer");
    REQUIRE(ksw4 <= 0 || mu4, "mu4 is required when ksw4 > 0");
    HIPCHECK(hipSetDevice(c->device));
    TRY(fused_flush(c));
    const size_t bd = c->n * sizeof(double);
    double *st_in[6] = {ubrtr_e, ubrtrp_e, vbrtr_e, vbrtrp_e, ssh_e, sshp_e};
    for (int k = 0; k < 6; ++k) TRY(upload(c, c->st[k], st_in[k], bd));
    // inputs go to scratch arrays that the stage-by-stage path does not otherwise use
    TRY(upload(c, c->rhsx, RHSx, bd));
    TRY(upload(c, c->rhsy, RHSy, bd));
    TRY(upload(c, c->surf_x, wflux, bd));
    TRY(upload(c, c->surf_y, mu, bd));
    TRY(upload(c, c->bot_x, mu4 ? mu4 : mu, bd));
    TRY(upload(c, c->bot_y, rdis, bd));
    TRY(dynamics_on_context(c, tau, nstep, ksw4, c->rhsx, c->rhsy, c->surf_x, c->surf_y, c->bot_x, c->bot_y));
    double *st_out[6] = {ubrtr_e, ubrtrp_e, vbrtr_e, vbrtrp_e, ssh_e, sshp_e};
    for (int k = 0; k < 6; ++k)
        HIPCHECK(hipMemcpyAsync(st_out[k], c->st[k], bd, hipMemcpyDeviceToHost, c->stream));
    double *acc_out[4] = {ubrtr_i, vbrtr_i, ssh4gradx, ssh4grady};
    for (int k = 0; k < 4; ++k)
        HIPCHECK(hipMemcpyAsync(acc_out[k], c->acc[k], bd, hipMemcpyDeviceToHost, c->stream));
    HIPCHECK(hipStreamSynchronize(c->stream));
    return INMSOM_BT_OK;
}

int inmsom_bt_barotropic_dynamics_dev(inmsom_bt_ctx *c, double tau, int nstep, int ksw4,
                                      double *ubrtr_e, double *ubrtrp_e, double *vbrtr_e,
                                      double *vbrtrp_e, double *ssh_e, double *sshp_e,
                                      double *ubrtr_i, double *vbrtr_i, double *ssh4gradx,
                                      double *ssh4grady, const double *RHSx, const double *RHSy,
                                      const double *wflux, const double *mu, const double *mu4,
                                      const double *rdis)
{
    REQUIRE(c, "null context");
    REQUIRE(nstep >= 1 && nstep <= 1000, "nstep must be in 1..1000");
    REQUIRE(tau > 0.0, "tau must be positive");
    REQUIRE(ubrtr_e && ubrtrp_e && vbrtr_e && vbrtrp_e && ssh_e && sshp_e && ubrtr_i && vbrtr_i &&
                ssh4gradx && ssh4grady && RHSx && RHSy && wflux && mu && rdis,
            "null array pointer");
    REQUIRE(ksw4 <= 0 || mu4, "mu4 is required when ksw4 > 0");
    HIPCHECK(hipSetDevice(c->device));
    bool fused = false;
    TRY(dropin_route(c, ksw4, &fused));
    double *st[6] = {ubrtr_e, ubrtrp_e, vbrtr_e, vbrtrp_e, ssh_e, sshp_e};
    double *acc[4] = {ubrtr_i, vbrtr_i, ssh4gradx, ssh4grady};
    if (fused) {
        TRY(dropin_call(c, tau, nstep, st, acc, RHSx, RHSy, wflux, mu, rdis, nullptr, ksw4, mu4));
    } else {
        // the stage-by-stage path runs on the context's own arrays: its
        // state is set aside, the caller's goes in and comes back by device
        // copies, then the context's own is put back (depths re-derived
        // from it when next needed)
        TRY(fused_flush(c));
        const size_t bd = c->n * sizeof(double);
        double *own[10], *theirs[10];
        for (int k = 0; k < 6; ++k) {
            own[k] = c->st[k];
            theirs[k] = st[k];
        }
        for (int k = 0; k < 4; ++k) {
            own[6 + k] = c->acc[k];
            theirs[6 + k] = acc[k];
        }
        for (int k = 0; k < 10; ++k) {
            if (!c->st_save[k]) DALLOC(c->st_save[k], bd);
            HIPCHECK(hipMemcpyAsync(c->st_save[k], own[k], bd, hipMemcpyDeviceToDevice, c->stream));
            if (k < 6) HIPCHECK(hipMemcpyAsync(own[k], theirs[k], bd, hipMemcpyDeviceToDevice, c->stream));
        }
        lit_hh_init(c, c->st[4], c->st[5]);
        lit_barotropic_dynamics(c, tau, nstep, ksw4, RHSx, RHSy, wflux, mu, mu4 ? mu4 : mu, rdis);
        for (int k = 0; k < 10; ++k) {
            HIPCHECK(hipMemcpyAsync(theirs[k], own[k], bd, hipMemcpyDeviceToDevice, c->stream));
            HIPCHECK(hipMemcpyAsync(own[k], c->st_save[k], bd, hipMemcpyDeviceToDevice, c->stream));
        }
        c->dep_stale = true;
    }
    HIPCHECK(hipGetLastError());
    HIPCHECK(hipStreamSynchronize(c->stream));
    return INMSOM_BT_OK;
}

int inmsom_bt_barotropic_dynamics_noalloc(inmsom_bt_ctx *c, double tau, int nstep, double *ubrtr_e,
                                          double *ubrtrp_e, double *vbrtr_e, double *vbrtrp_e, double *ssh_e,
                                          double *sshp_e, double *ubrtr_i, double *vbrtr_i, double *ssh4gradx,
                                          double *ssh4grady, const double *RHSx, const double *RHSy,
                                          const double *wflux, const double *mu, const double *rdis)
{
    REQUIRE(c, "null context");
    REQUIRE(nstep >= 1 && nstep <= 1000, "nstep must be in 1..1000");
    REQUIRE(tau > 0.0, "tau must be positive");
    REQUIRE(ubrtr_e && ubrtrp_e && vbrtr_e && vbrtrp_e && ssh_e && sshp_e && ubrtr_i && vbrtr_i &&
                ssh4gradx && ssh4grady && RHSx && RHSy && wflux && mu && rdis,
            "null array pointer");
    size_t have = 0;
    double *scr = noalloc_scratch(c->device, &have);
    if (!scr) {
        set_error("the scratch of the allocation-free entry is not set up on this device");
        return INMSOM_BT_ESTATE;
    }
    if (dropin_scratch_doubles(c) > have) {
        set_error("grid too large for the allocation-free entry: it needs 14 planes of nx*ny doubles "
                  "within its fixed scratch of " + std::to_string(have * 8 >> 20) + " MiB");
        return INMSOM_BT_EINVAL;
    }
    if (c->dropin_path == 1 || !dropin_eligible(c, 0, nullptr)) {
        set_error("the allocation-free entry runs the fused barotropic_dynamics only, which this context's grid "
                  "or path setting does not allow");
        return INMSOM_BT_ESTATE;
    }
    HIPCHECK(hipSetDevice(c->device));
    double *st[6] = {ubrtr_e, ubrtrp_e, vbrtr_e, vbrtrp_e, ssh_e, sshp_e};
    double *acc[4] = {ubrtr_i, vbrtr_i, ssh4gradx, ssh4grady};
    TRY(dropin_call(c, tau, nstep, st, acc, RHSx, RHSy, wflux, mu, rdis, scr));
    return INMSOM_BT_OK;
}

int inmsom_bt_set_dropin_path(inmsom_bt_ctx *c, int path)
{
    REQUIRE(c, "null context");
    REQUIRE(path >= 0 && path <= 2, "path must be 0 (automatic), 1 (stage by stage) or 2 (fused)");
    c->dropin_path = path;
    return INMSOM_BT_OK;
}

int inmsom_bt_get_dropin_path(inmsom_bt_ctx *c, int ksw4, int *path)
{
    REQUIRE(c && path, "null pointer");
    HIPCHECK(hipSetDevice(c->device));
    bool fused = false;
    TRY(dropin_route(c, ksw4, &fused));
    *path = fused ? 2 : 1;
    return INMSOM_BT_OK;
}

int inmsom_bt_dropin_timed(inmsom_bt_ctx *c, double tau, int nstep, int ksw4, int ncalls, double *ms)
{
    REQUIRE(c && ms, "null pointer");
    REQUIRE(nstep >= 1 && nstep <= 1000 && ncalls >= 0 && tau > 0.0, "bad tau / nstep / ncalls");
    REQUIRE(c->have_forcing, "inmsom_bt_set_forcing must be called first");
    HIPCHECK(hipSetDevice(c->device));
    TRY(fused_flush(c));
    refresh_depths(c);
    // the RHS of a model step (wind stress and bottom drag of the current state)
    lit_forcing(c);
    hipEvent_t e0, e1;
    HIPCHECK(hipEventCreate(&e0));
    HIPCHECK(hipEventCreate(&e1));
    HIPCHECK(hipEventRecord(e0, c->stream));
    for (int k = 0; k < ncalls; ++k)
        TRY(dynamics_on_context(c, tau, nstep, ksw4, c->rhsx, c->rhsy, c->wflux, c->mu, c->mu4, c->rdis));
    HIPCHECK(hipEventRecord(e1, c->stream));
    HIPCHECK(hipEventSynchronize(e1));
    float t = 0.f;
    HIPCHECK(hipEventElapsedTime(&t, e0, e1));
    hipEventDestroy(e0);
    hipEventDestroy(e1);
    *ms = t;
    return INMSOM_BT_OK;
}

int inmsom_bt_get_depths(inmsom_bt_ctx *c, double *const h[12])
{
    REQUIRE(c && h, "null pointer");
    HIPCHECK(hipSetDevice(c->device));
    TRY(fused_flush(c));
    refresh_depths(c);
    for (int k = 0; k < 12; ++k) {
        REQUIRE(h[k], "null depth pointer");
        HIPCHECK(hipMemcpyAsync(h[k], c->dep[k], c->n * sizeof(double), hipMemcpyDeviceToHost,
                                c->stream));
    }
    HIPCHECK(hipStreamSynchronize(c->stream));
    return INMSOM_BT_OK;
}

int inmsom_bt_set_forcing(inmsom_bt_ctx *c, const inmsom_bt_forcing *f)
{
    REQUIRE(c && f, "null pointer");
    REQUIRE(f->taux && f->tauy && f->wflux && f->mu && f->rdis, "null forcing array");
    HIPCHECK(hipSetDevice(c->device));
    const size_t bd = c->n * sizeof(double);
    TRY(upload(c, c->taux, f->taux, bd));
    TRY(upload(c, c->tauy, f->tauy, bd));
    TRY(upload(c, c->wflux, f->wflux, bd));
    TRY(upload(c, c->mu, f->mu, bd));
    if (f->mu4)
        TRY(upload(c, c->mu4, f->mu4, bd));
    TRY(upload(c, c->rdis, f->rdis, bd));
    c->cb_nl = f->cb_nl;
    c->ebottom = f->ebottom;
    c->have_forcing = true;
    c->fused_ready = false;  // per-row tables depend on the forcing
    HIPCHECK(hipStreamSynchronize(c->stream));
    return INMSOM_BT_OK;
}

int inmsom_bt_set_state(inmsom_bt_ctx *c, const double *u, const double *up, const double *v,
                        const double *vp, const double *ssh, const double *sshp)
{
    REQUIRE(c && u && up && v && vp && ssh && sshp, "null pointer");
    HIPCHECK(hipSetDevice(c->device));
    const size_t bd = c->n * sizeof(double);
    const double *in[6] = {u, up, v, vp, ssh, sshp};
    for (int k = 0; k < 6; ++k) TRY(upload(c, c->st[k], in[k], bd));
    c->state_on_fused = false;
    c->fused_mode = -1;
    lit_hh_init(c, c->st[4], c->st[5]);
    c->dep_stale = false;
    HIPCHECK(hipStreamSynchronize(c->stream));
    return INMSOM_BT_OK;
}

int inmsom_bt_save_state(inmsom_bt_ctx *c)
{
    REQUIRE(c, "null pointer");
    HIPCHECK(hipSetDevice(c->device));
    TRY(fused_flush(c));
    const size_t bd = c->n * sizeof(double);
    for (int k = 0; k < 6; ++k) {
        if (!c->st_snap[k]) HIPCHECK(hipMalloc(&c->st_snap[k], bd));
        HIPCHECK(hipMemcpyAsync(c->st_snap[k], c->st[k], bd, hipMemcpyDeviceToDevice, c->stream));
    }
    HIPCHECK(hipStreamSynchronize(c->stream));
    c->snap_ok = true;
    return INMSOM_BT_OK;
}

int inmsom_bt_restore_state(inmsom_bt_ctx *c)
{
    REQUIRE(c, "null pointer");
    if (!c->snap_ok) {
        set_error("no saved state: call inmsom_bt_save_state first");
        return INMSOM_BT_ESTATE;
    }
    HIPCHECK(hipSetDevice(c->device));
    const size_t bd = c->n * sizeof(double);
    for (int k = 0; k < 6; ++k)
        HIPCHECK(hipMemcpyAsync(c->st[k], c->st_snap[k], bd, hipMemcpyDeviceToDevice, c->stream));
    c->state_on_fused = false;
    c->fused_mode = -1;
    lit_hh_init(c, c->st[4], c->st[5]);
    c->dep_stale = false;
    HIPCHECK(hipStreamSynchronize(c->stream));
    return INMSOM_BT_OK;
}

int inmsom_bt_get_state(inmsom_bt_ctx *c, double *u, double *up, double *v, double *vp, double *ssh,
                        double *sshp)
{
    REQUIRE(c && u && up && v && vp && ssh && sshp, "null pointer");
    HIPCHECK(hipSetDevice(c->device));
    TRY(fused_flush(c));
    double *out[6] = {u, up, v, vp, ssh, sshp};
    for (int k = 0; k < 6; ++k)
        HIPCHECK(hipMemcpyAsync(out[k], c->st[k], c->n * sizeof(double), hipMemcpyDeviceToHost,
                                c->stream));
    HIPCHECK(hipStreamSynchronize(c->stream));
    return INMSOM_BT_OK;
}

int inmsom_bt_get_means(inmsom_bt_ctx *c, double *ubrtr_i, double *vbrtr_i, double *ssh4gradx,
                        double *ssh4grady)
{
    REQUIRE(c && ubrtr_i && vbrtr_i && ssh4gradx && ssh4grady, "null pointer");
    HIPCHECK(hipSetDevice(c->device));
    double *out[4] = {ubrtr_i, vbrtr_i, ssh4gradx, ssh4grady};
    for (int k = 0; k < 4; ++k)
        HIPCHECK(hipMemcpyAsync(out[k], c->acc[k], c->n * sizeof(double), hipMemcpyDeviceToHost,
                                c->stream));
    HIPCHECK(hipStreamSynchronize(c->stream));
    return INMSOM_BT_OK;
}

int inmsom_bt_model_steps(inmsom_bt_ctx *c, double tau, int nstep, int ksw4, int nsteps)
{
    REQUIRE(c, "null context");
    REQUIRE(c->have_forcing, "inmsom_bt_set_forcing must be called first");
    REQUIRE(nstep >= 1 && nsteps >= 0 && tau > 0.0, "bad tau / nstep / nsteps");
    HIPCHECK(hipSetDevice(c->device));
    TRY(fused_flush(c));
    for (int s = 0; s < nsteps; ++s) {
        // the forcing reads the depths hh_init left (Control/dyncall.f90:110-130)
        refresh_depths(c);
        lit_forcing(c);
        TRY(dynamics_on_context(c, tau, nstep, ksw4, c->rhsx, c->rhsy, c->wflux, c->mu, c->mu4, c->rdis));
    }
    HIPCHECK(hipGetLastError());
    HIPCHECK(hipStreamSynchronize(c->stream));
    return INMSOM_BT_OK;
}

int inmsom_bt_run(inmsom_bt_ctx *c, double tau, int nstep, int nsteps, int mode)
{
    REQUIRE(c, "null context");
    HIPCHECK(hipSetDevice(c->device));
    TRY(fused_run(c, tau, nstep, nsteps, mode, nullptr));
    return INMSOM_BT_OK;
}

int inmsom_bt_run_timed(inmsom_bt_ctx *c, double tau, int nstep, int nsteps, int mode, double *ms)
{
    REQUIRE(c && ms, "null pointer");
    HIPCHECK(hipSetDevice(c->device));
    TRY(fused_run(c, tau, nstep, nsteps, mode, ms));
    return INMSOM_BT_OK;
}

int inmsom_bt_set_pipeline(inmsom_bt_ctx *c, int stages)
{
    REQUIRE(c, "null context");
    TRY(fused_set_pipeline(c, stages));
    return INMSOM_BT_OK;
}

int inmsom_bt_get_pipeline(inmsom_bt_ctx *c, int mode, int *stages)
{
    REQUIRE(c && stages, "null pointer");
    REQUIRE(mode == INMSOM_BT_FP64_EXACT || mode == INMSOM_BT_FP64 || mode == INMSOM_BT_FP32, "unknown mode");
    *stages = pipe_stages(c, mode);
    return INMSOM_BT_OK;
}

int inmsom_bt_set_launch(inmsom_bt_ctx *c, int rows_per_wave)
{
    REQUIRE(c, "null context");
    TRY(fused_set_shape(c, rows_per_wave));
    return INMSOM_BT_OK;
}

int inmsom_bt_set_substeps_per_launch(inmsom_bt_ctx *c, int substeps)
{
    REQUIRE(c, "null context");
    TRY(fused_set_substeps(c, substeps));
    return INMSOM_BT_OK;
}

int inmsom_bt_get_substeps_per_launch(inmsom_bt_ctx *c, int mode, int *substeps)
{
    REQUIRE(c && substeps, "null pointer");
    REQUIRE(mode == INMSOM_BT_FP64_EXACT || mode == INMSOM_BT_FP64 || mode == INMSOM_BT_FP32, "unknown mode");
    const int k = pipe_stages(c, mode);
    *substeps = k > 0 ? k : subs_per_launch(c, mode);
    return INMSOM_BT_OK;
}

int inmsom_bt_get_launch(inmsom_bt_ctx *c, int *rows_per_wave)
{
    REQUIRE(c && rows_per_wave, "null pointer");
    *rows_per_wave = c->chunk;
    return INMSOM_BT_OK;
}

int inmsom_bt_range_check(inmsom_bt_ctx *c, long long *bad)
{
    REQUIRE(c && bad, "null pointer");
    HIPCHECK(hipSetDevice(c->device));
    TRY(fused_range_check(c, bad));
    return *bad > 0 ? INMSOM_BT_ERANGE : INMSOM_BT_OK;
}

int inmsom_bt_volume(inmsom_bt_ctx *c, double *vol)
{
    REQUIRE(c && vol, "null pointer");
    HIPCHECK(hipSetDevice(c->device));
    TRY(fused_volume(c, vol));
    return INMSOM_BT_OK;
}

int inmsom_bt_coarsen_state(inmsom_bt_ctx *c, int factor, double *u, double *v, double *ssh)
{
    REQUIRE(c && u && v && ssh, "null pointer");
    HIPCHECK(hipSetDevice(c->device));
    TRY(fused_coarsen(c, factor, u, v, ssh));
    return INMSOM_BT_OK;
}

int inmsom_bt_synchronize(inmsom_bt_ctx *c)
{
    REQUIRE(c, "null context");
    HIPCHECK(hipSetDevice(c->device));
    HIPCHECK(hipStreamSynchronize(c->stream));
    return INMSOM_BT_OK;
}

}  // extern "C"
