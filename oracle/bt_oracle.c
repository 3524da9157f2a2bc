/*
 * CPU restatement of INMSOM's barotropic hot path (TEST INFRASTRUCTURE:
 * the checker for the HIP implementation, never part of the product).
 *
 * Each function cites the reference routine it restates.  Expressions
 * keep the reference's operand order and parenthesisation, because
 * parity with the compiled reference is checked bit for bit; the file
 * must be compiled with -ffp-contract=off.
 *
 * Single-rank semantics: the reference's halo exchange
 * (Modules/mod_mpi_parallel_tools.f90:179-252) moves nothing when the
 * process grid is 1x1, so the sync calls disappear.  Periodic wrap
 * (Service/oc_algorithms.f90:50-135) is restated as the copy a
 * decomposed run performs: column mm -> mmm-1 and mmm -> mm+1 (rows
 * likewise).
 *
 * Pinning (what checks each function, all under tests/):
 *   - the barotropic step and its helpers: bit for bit against the
 *     compiled reference through oracle/ref_harness.f90 modes 0 and 1
 *     (test_oracle_golden.py, committed digests; test_oracle_reference_live.py)
 *     and by closed forms (test_oracle_invariants.py);
 *   - bto_remove_barotropic / bto_internal_update (mode 2), bto_forcing_3d
 *     (mode 3), bto_baroclinic_dynamics (mode 4), bto_vertical_velocity
 *     (mode 5), bto_pressure_gradients (mode 6), bto_density (mode 7): bit
 *     for bit against the reference's own routines
 *     (test_oracle_reference_live.py) and by closed forms or invariants
 *     (test_oracle_invariants.py);
 *   - bto_rayleigh_init: by closed forms (the reference computes it inside
 *     an initialisation that needs input files it does not ship): a
 *     linear slope, and a depth quadratic in the distances along x and y
 *     on a grid whose eight metrics all differ, with coasts of one and of
 *     two open faces (test_oracle_invariants.py; a wrong metric, face side
 *     or mask fails it);
 *   - the periodic wrap: by translation invariance and by the equality of
 *     a periodic basin with each copy of it tiled twice, bit for bit (the
 *     shipped reference has periodicity 0; a wrap that copies a column to
 *     its own neighbour or is off by one fails the latter).
 */
#include "bt_oracle.h"

#include <math.h>
#include <stdlib.h>
#include <string.h>
#ifdef _OPENMP
#include <omp.h>
#endif

#define IX(m, n) ((size_t)((n) - 1) * (size_t)g->nx + (size_t)((m) - 1))
#define NXS 3
#define NXE (g->nx - 2)
#define NYS 3
#define NYE (g->ny - 2)

/* Service/oc_algorithms.f90:50-95 (decomposed-run semantics). */
static void cyclize_x(const bto_grid *g, double *f)
{
    for (int n = 1; n <= g->ny; ++n) {
        double a = f[IX(NXE, n)], b = f[IX(NXS, n)];
        f[IX(NXS - 1, n)] = a;
        f[IX(NXE + 1, n)] = b;
    }
}

/* Service/oc_algorithms.f90:118-135. */
static void cyclize_y(const bto_grid *g, double *f)
{
    for (int m = 1; m <= g->nx; ++m) {
        double a = f[IX(m, NYE)], b = f[IX(m, NYS)];
        f[IX(m, NYS - 1)] = a;
        f[IX(m, NYE + 1)] = b;
    }
}

static void cyclize(const bto_grid *g, double *f)
{
    if (g->periodicity_x) cyclize_x(g, f);
    if (g->periodicity_y) cyclize_y(g, f);
}

static void cyclize_xf(int nx, int ny, float *f)
{
    for (int n = 0; n < ny; ++n) {
        float a = f[(size_t)n * nx + (nx - 3)], b = f[(size_t)n * nx + 2];
        f[(size_t)n * nx + 1] = a;
        f[(size_t)n * nx + (nx - 2)] = b;
    }
}

static void cyclize_yf(int nx, int ny, float *f)
{
    for (int m = 0; m < nx; ++m) {
        float a = f[(size_t)(ny - 3) * nx + m], b = f[(size_t)2 * nx + m];
        f[(size_t)1 * nx + m] = a;
        f[(size_t)(ny - 2) * nx + m] = b;
    }
}

void bto_gridcon(int nx, int ny, int px, int py, const int *lb, float *lu,
                 float *luu, float *luh, float *lcu, float *lcv, float *llu,
                 float *llv)
{
    size_t sz = (size_t)nx * ny;
    memset(lu, 0, sz * sizeof(float));
    memset(luu, 0, sz * sizeof(float));
    memset(luh, 0, sz * sizeof(float));
    memset(lcu, 0, sz * sizeof(float));
    memset(lcv, 0, sz * sizeof(float));
    memset(llu, 0, sz * sizeof(float));
    memset(llv, 0, sz * sizeof(float));
#define J(m, n) ((size_t)((n) - 1) * (size_t)nx + (size_t)((m) - 1))
    /* grid_construction.f90:33-39 */
    for (int n = 1; n <= ny; ++n)
        for (int m = 1; m <= nx; ++m)
            if (lb[J(m, n)] == 0) lu[J(m, n)] = 1.0f;
    if (px) cyclize_xf(nx, ny, lu);
    if (py) cyclize_yf(nx, ny, lu);
    /* :59-73 over nx_start-1..nx_end, ny_start-1..ny_end */
    for (int n = 2; n <= ny - 2; ++n)
        for (int m = 2; m <= nx - 2; ++m) {
            if (lu[J(m, n)] + lu[J(m + 1, n)] + lu[J(m, n + 1)] + lu[J(m + 1, n + 1)] > 0.5f)
                luh[J(m, n)] = 1.0f;
            if (lu[J(m, n)] * lu[J(m + 1, n)] * lu[J(m, n + 1)] * lu[J(m + 1, n + 1)] > 0.5f)
                luu[J(m, n)] = 1.0f;
        }
    if (px) { cyclize_xf(nx, ny, luh); cyclize_xf(nx, ny, luu); }
    if (py) { cyclize_yf(nx, ny, luh); cyclize_yf(nx, ny, luu); }
    /* :92-112 */
    for (int n = 2; n <= ny - 2; ++n)
        for (int m = 2; m <= nx - 2; ++m) {
            if (lu[J(m, n)] + lu[J(m + 1, n)] > 0.5f) llu[J(m, n)] = 1.0f;
            if (lu[J(m, n)] + lu[J(m, n + 1)] > 0.5f) llv[J(m, n)] = 1.0f;
            if (lu[J(m, n)] * lu[J(m + 1, n)] > 0.5f) lcu[J(m, n)] = 1.0f;
            if (lu[J(m, n)] * lu[J(m, n + 1)] > 0.5f) lcv[J(m, n)] = 1.0f;
        }
    if (px) {
        cyclize_xf(nx, ny, lcu); cyclize_xf(nx, ny, llu);
        cyclize_xf(nx, ny, lcv); cyclize_xf(nx, ny, llv);
    }
    if (py) {
        cyclize_yf(nx, ny, lcu); cyclize_yf(nx, ny, llu);
        cyclize_yf(nx, ny, lcv); cyclize_yf(nx, ny, llv);
    }
#undef J
}

/* Control/init_pars.f90:459-481 */
void bto_rayleigh_init(const bto_grid *g, double r_fric, double *r_diss)
{
    const double *h = g->hhq_rest;
#pragma omp parallel for schedule(static)
    for (int n = NYS; n <= NYE; ++n)
        for (int m = NXS; m <= NXE; ++m) {
            if (g->lu[IX(m, n)] > 0.5f) {
                double a = (h[IX(m + 1, n)] - h[IX(m, n)]) / g->dxt[IX(m, n)];
                double b = (h[IX(m, n)] - h[IX(m - 1, n)]) / g->dxt[IX(m - 1, n)];
                double hx2 = (a * a * (double)g->lcu[IX(m, n)] +
                              b * b * (double)g->lcu[IX(m - 1, n)]) /
                             (double)(g->lcu[IX(m, n)] + g->lcu[IX(m - 1, n)]);
                double c = (h[IX(m, n + 1)] - h[IX(m, n)]) / g->dyt[IX(m, n)];
                double e = (h[IX(m, n)] - h[IX(m, n - 1)]) / g->dyt[IX(m, n - 1)];
                double hy2 = (c * c * (double)g->lcv[IX(m, n)] +
                              e * e * (double)g->lcv[IX(m, n - 1)]) /
                             (double)(g->lcv[IX(m, n)] + g->lcv[IX(m, n - 1)]);
                r_diss[IX(m, n)] = r_fric * sqrt(hx2 + hy2);
            }
        }
    cyclize(g, r_diss);
}

/* Depth on U/V/H points from a T-point depth, Function/depth.f90:29-66. */
static inline double hu_of(const bto_grid *g, const double *hq, int m, int n)
{
    double slu = (double)(g->lu[IX(m, n)] + g->lu[IX(m + 1, n)]);
    return (hq[IX(m, n)] * g->dx[IX(m, n)] * g->dy[IX(m, n)] * (double)g->lu[IX(m, n)] +
            hq[IX(m + 1, n)] * g->dx[IX(m + 1, n)] * g->dy[IX(m + 1, n)] * (double)g->lu[IX(m + 1, n)]) /
           slu / g->dxt[IX(m, n)] / g->dyh[IX(m, n)];
}

static inline double hv_of(const bto_grid *g, const double *hq, int m, int n)
{
    double slu = (double)(g->lu[IX(m, n)] + g->lu[IX(m, n + 1)]);
    return (hq[IX(m, n)] * g->dx[IX(m, n)] * g->dy[IX(m, n)] * (double)g->lu[IX(m, n)] +
            hq[IX(m, n + 1)] * g->dx[IX(m, n + 1)] * g->dy[IX(m, n + 1)] * (double)g->lu[IX(m, n + 1)]) /
           slu / g->dxh[IX(m, n)] / g->dyt[IX(m, n)];
}

static inline double hh_of(const bto_grid *g, const double *hq, int m, int n)
{
    double slu = (double)(g->lu[IX(m, n)] + g->lu[IX(m + 1, n)] + g->lu[IX(m, n + 1)] +
                          g->lu[IX(m + 1, n + 1)]);
    return (hq[IX(m, n)] * g->dx[IX(m, n)] * g->dy[IX(m, n)] * (double)g->lu[IX(m, n)] +
            hq[IX(m + 1, n)] * g->dx[IX(m + 1, n)] * g->dy[IX(m + 1, n)] * (double)g->lu[IX(m + 1, n)] +
            hq[IX(m, n + 1)] * g->dx[IX(m, n + 1)] * g->dy[IX(m, n + 1)] * (double)g->lu[IX(m, n + 1)] +
            hq[IX(m + 1, n + 1)] * g->dx[IX(m + 1, n + 1)] * g->dy[IX(m + 1, n + 1)] *
                (double)g->lu[IX(m + 1, n + 1)]) /
           slu / g->dxb[IX(m, n)] / g->dyb[IX(m, n)];
}

/* Function/depth.f90:2-105 */
void bto_hh_init(const bto_grid *g, bto_depths *d, const double *ssh,
                 const double *sshp)
{
    size_t sz = (size_t)g->nx * g->ny;
    double ffs = (double)g->full_free_surface;
    double *hq = d->h[0], *hqp = d->h[1], *hqn = d->h[2];
    for (size_t i = 0; i < sz; ++i) {
        hq[i] = g->hhq_rest[i] + ssh[i] * ffs;
        hqp[i] = g->hhq_rest[i] + sshp[i] * ffs;
        hqn[i] = g->hhq_rest[i];
    }
#pragma omp parallel for schedule(static)
    for (int n = NYS - 1; n <= NYE; ++n)
        for (int m = NXS - 1; m <= NXE; ++m) {
            if (g->llu[IX(m, n)] > 0.5f) {
                d->h[3][IX(m, n)] = hu_of(g, hq, m, n);
                d->h[4][IX(m, n)] = hu_of(g, hqp, m, n);
                d->h[5][IX(m, n)] = hu_of(g, hqn, m, n);
            }
            if (g->llv[IX(m, n)] > 0.5f) {
                d->h[6][IX(m, n)] = hv_of(g, hq, m, n);
                d->h[7][IX(m, n)] = hv_of(g, hqp, m, n);
                d->h[8][IX(m, n)] = hv_of(g, hqn, m, n);
            }
            if (g->luh[IX(m, n)] > 0.5f) {
                d->h[9][IX(m, n)] = hh_of(g, hq, m, n);
                d->h[10][IX(m, n)] = hh_of(g, hqp, m, n);
                d->h[11][IX(m, n)] = hh_of(g, hqn, m, n);
            }
        }
    for (int k = 3; k < 12; ++k) cyclize(g, d->h[k]);
}

/* Function/depth.f90:108-170 */
void bto_hh_update(const bto_grid *g, double *hqn, double *hun, double *hvn,
                   double *hhn, const double *sh)
{
    size_t sz = (size_t)g->nx * g->ny;
    for (size_t i = 0; i < sz; ++i) hqn[i] = g->hhq_rest[i] + sh[i];
#pragma omp parallel for schedule(static)
    for (int n = NYS - 1; n <= NYE; ++n)
        for (int m = NXS - 1; m <= NXE; ++m) {
            if (g->llu[IX(m, n)] > 0.5f) hun[IX(m, n)] = hu_of(g, hqn, m, n);
            if (g->llv[IX(m, n)] > 0.5f) hvn[IX(m, n)] = hv_of(g, hqn, m, n);
            if (g->luh[IX(m, n)] > 0.5f) hhn[IX(m, n)] = hh_of(g, hqn, m, n);
        }
    cyclize(g, hun);
    cyclize(g, hvn);
    cyclize(g, hhn);
}

/* Function/depth.f90:174-220 */
void bto_hh_shift(const bto_grid *g, bto_depths *d, int nstep)
{
    const double ts = g->time_smooth, dn = (double)nstep;
#pragma omp parallel for schedule(static)
    for (int n = NYS - 1; n <= NYE + 1; ++n)
        for (int m = NXS - 1; m <= NXE + 1; ++m) {
            size_t i = IX(m, n);
            if (g->llu[i] > 0.5f) {
                double *h = d->h[3], *hp = d->h[4], *hn = d->h[5];
                hp[i] = h[i] + ts * (hn[i] - 2.0 * h[i] + hp[i]) / 2.0 / dn;
                h[i] = hn[i];
            }
            if (g->llv[i] > 0.5f) {
                double *h = d->h[6], *hp = d->h[7], *hn = d->h[8];
                hp[i] = h[i] + ts * (hn[i] - 2.0 * h[i] + hp[i]) / 2.0 / dn;
                h[i] = hn[i];
            }
            if (g->lu[i] > 0.5f) {
                double *h = d->h[0], *hp = d->h[1], *hn = d->h[2];
                hp[i] = h[i] + ts * (hn[i] - 2.0 * h[i] + hp[i]) / 2.0 / dn;
                h[i] = hn[i];
            }
            if (g->luh[i] > 0.5f) {
                double *h = d->h[9], *hp = d->h[10], *hn = d->h[11];
                hp[i] = h[i] + ts * (hn[i] - 2.0 * h[i] + hp[i]) / 2.0 / dn;
                h[i] = hn[i];
            }
        }
}

/* Function/mixing.f90:115-168 */
void bto_stress_components(const bto_grid *g, const double *u, const double *v,
                           double *str_t, double *str_s)
{
#pragma omp parallel for schedule(static)
    for (int n = NYS; n <= NYE; ++n)
        for (int m = NXS; m <= NXE; ++m)
            if (g->lu[IX(m, n)] > 0.5f)
                str_t[IX(m, n)] =
                    g->dy[IX(m, n)] / g->dx[IX(m, n)] *
                        (u[IX(m, n)] / g->dyh[IX(m, n)] - u[IX(m - 1, n)] / g->dyh[IX(m - 1, n)]) -
                    g->dx[IX(m, n)] / g->dy[IX(m, n)] *
                        (v[IX(m, n)] / g->dxh[IX(m, n)] - v[IX(m, n - 1)] / g->dxh[IX(m, n - 1)]);
#pragma omp parallel for schedule(static)
    for (int n = NYS; n <= NYE; ++n)
        for (int m = NXS; m <= NXE; ++m)
            if (g->luu[IX(m, n)] > 0.5f)
                str_s[IX(m, n)] =
                    g->dxb[IX(m, n)] / g->dyb[IX(m, n)] *
                        (u[IX(m, n + 1)] / g->dxt[IX(m, n + 1)] - u[IX(m, n)] / g->dxt[IX(m, n)]) +
                    g->dyb[IX(m, n)] / g->dxb[IX(m, n)] *
                        (v[IX(m + 1, n)] / g->dyt[IX(m + 1, n)] - v[IX(m, n)] / g->dyt[IX(m, n)]);
    cyclize(g, str_t);
    cyclize(g, str_s);
}

/* Function/vel_ssh.f90:2-114 */
void bto_uv_trans(const bto_grid *g, const double *u, const double *v,
                  double *vort, const double *hq, const double *hu,
                  const double *hv, const double *hh, double *rhsx,
                  double *rhsy)
{
    (void)hq;
    const double *dxt = g->dxt, *dyt = g->dyt, *dxb = g->dxb, *dyb = g->dyb;
    const double *dxh = g->dxh, *dyh = g->dyh;
#pragma omp parallel for schedule(static)
    for (int n = NYS; n <= NYE; ++n)
        for (int m = NXS; m <= NXE; ++m)
            if (g->luu[IX(m, n)] > 0.5f)
                vort[IX(m, n)] =
                    (v[IX(m + 1, n)] * dyt[IX(m + 1, n)] - v[IX(m, n)] * dyt[IX(m, n)]) -
                    (u[IX(m, n + 1)] * dxt[IX(m, n + 1)] - u[IX(m, n)] * dxt[IX(m, n)]) -
                    ((v[IX(m + 1, n)] - v[IX(m, n)]) * dyb[IX(m, n)] -
                     (u[IX(m, n + 1)] - u[IX(m, n)]) * dxb[IX(m, n)]);
    cyclize(g, vort);

#pragma omp parallel for schedule(static)
    for (int n = NYS; n <= NYE; ++n)
        for (int m = NXS; m <= NXE; ++m) {
            if (g->lcu[IX(m, n)] > 0.5f) {
                double fx_p = (u[IX(m, n)] * dyh[IX(m, n)] * hu[IX(m, n)] +
                               u[IX(m + 1, n)] * dyh[IX(m + 1, n)] * hu[IX(m + 1, n)]) / 2.0 *
                              (u[IX(m, n)] + u[IX(m + 1, n)]) / 2.0;
                double fx_m = (u[IX(m, n)] * dyh[IX(m, n)] * hu[IX(m, n)] +
                               u[IX(m - 1, n)] * dyh[IX(m - 1, n)] * hu[IX(m - 1, n)]) / 2.0 *
                              (u[IX(m, n)] + u[IX(m - 1, n)]) / 2.0;
                double fy_p = (v[IX(m, n)] * dxh[IX(m, n)] * hv[IX(m, n)] +
                               v[IX(m + 1, n)] * dxh[IX(m + 1, n)] * hv[IX(m + 1, n)]) / 2.0 *
                              (u[IX(m, n + 1)] + u[IX(m, n)]) / 2.0 * (double)g->luu[IX(m, n)];
                double fy_m = (v[IX(m, n - 1)] * dxh[IX(m, n - 1)] * hv[IX(m, n - 1)] +
                               v[IX(m + 1, n - 1)] * dxh[IX(m + 1, n - 1)] * hv[IX(m + 1, n - 1)]) / 2.0 *
                              (u[IX(m, n - 1)] + u[IX(m, n)]) / 2.0 * (double)g->luu[IX(m, n - 1)];
                rhsx[IX(m, n)] = -(fx_p - fx_m + fy_p - fy_m) +
                                 (vort[IX(m, n)] * hh[IX(m, n)] * (v[IX(m + 1, n)] + v[IX(m, n)]) +
                                  vort[IX(m, n - 1)] * hh[IX(m, n - 1)] *
                                      (v[IX(m + 1, n - 1)] + v[IX(m, n - 1)])) / 4.0;
            }
            if (g->lcv[IX(m, n)] > 0.5f) {
                double fy_p = (v[IX(m, n)] * dxh[IX(m, n)] * hv[IX(m, n)] +
                               v[IX(m, n + 1)] * dxh[IX(m, n + 1)] * hv[IX(m, n + 1)]) / 2.0 *
                              (v[IX(m, n)] + v[IX(m, n + 1)]) / 2.0;
                double fy_m = (v[IX(m, n)] * dxh[IX(m, n)] * hv[IX(m, n)] +
                               v[IX(m, n - 1)] * dxh[IX(m, n - 1)] * hv[IX(m, n - 1)]) / 2.0 *
                              (v[IX(m, n)] + v[IX(m, n - 1)]) / 2.0;
                double fx_p = (u[IX(m, n)] * dyh[IX(m, n)] * hu[IX(m, n)] +
                               u[IX(m, n + 1)] * dyh[IX(m, n + 1)] * hu[IX(m, n + 1)]) / 2.0 *
                              (v[IX(m + 1, n)] + v[IX(m, n)]) / 2.0;
                double fx_m = (u[IX(m - 1, n)] * dyh[IX(m - 1, n)] * hu[IX(m - 1, n)] +
                               u[IX(m - 1, n + 1)] * dyh[IX(m - 1, n + 1)] * hu[IX(m - 1, n + 1)]) / 2.0 *
                              (v[IX(m - 1, n)] + v[IX(m, n)]) / 2.0;
                rhsy[IX(m, n)] = -(fx_p - fx_m + fy_p - fy_m) -
                                 (vort[IX(m, n)] * hh[IX(m, n)] * (u[IX(m, n + 1)] + u[IX(m, n)]) +
                                  vort[IX(m - 1, n)] * hh[IX(m - 1, n)] *
                                      (u[IX(m - 1, n + 1)] + u[IX(m - 1, n)])) / 4.0;
            }
        }
}

/* The viscous sweeps of uv_diff2/uv_diff4 share one stencil
 * (Function/vel_ssh.f90:150-156, 166-172).  The tension part p and the
 * shear part q are returned separately because the reference adds them
 * to an existing RHS one after the other (:286-289, :302-305). */
static inline void visc_x(const bto_grid *g, const double *mu, const double *str_t,
                          const double *str_s, const double *hq, const double *hh,
                          int m, int n, double *p, double *q)
{
    double muh_p = (mu[IX(m, n)] + mu[IX(m + 1, n)] + mu[IX(m, n + 1)] + mu[IX(m + 1, n + 1)]) / 4.0;
    double muh_m = (mu[IX(m, n)] + mu[IX(m + 1, n)] + mu[IX(m, n - 1)] + mu[IX(m + 1, n - 1)]) / 4.0;
    double dy1 = g->dy[IX(m + 1, n)], dy0 = g->dy[IX(m, n)];
    double b1 = g->dxb[IX(m, n)], b0 = g->dxb[IX(m, n - 1)];
    *p = (dy1 * dy1 * mu[IX(m + 1, n)] * hq[IX(m + 1, n)] * str_t[IX(m + 1, n)] -
          dy0 * dy0 * mu[IX(m, n)] * hq[IX(m, n)] * str_t[IX(m, n)]) / g->dyh[IX(m, n)];
    *q = (b1 * b1 * muh_p * hh[IX(m, n)] * str_s[IX(m, n)] -
          b0 * b0 * muh_m * hh[IX(m, n - 1)] * str_s[IX(m, n - 1)]) / g->dxt[IX(m, n)];
}

/* p is the tension part the reference subtracts, q the shear part. */
static inline void visc_y(const bto_grid *g, const double *mu, const double *str_t,
                          const double *str_s, const double *hq, const double *hh,
                          int m, int n, double *p, double *q)
{
    double muh_p = (mu[IX(m, n)] + mu[IX(m + 1, n)] + mu[IX(m, n + 1)] + mu[IX(m + 1, n + 1)]) / 4.0;
    double muh_m = (mu[IX(m, n)] + mu[IX(m - 1, n)] + mu[IX(m, n + 1)] + mu[IX(m - 1, n + 1)]) / 4.0;
    double dx1 = g->dx[IX(m, n + 1)], dx0 = g->dx[IX(m, n)];
    double b1 = g->dyb[IX(m, n)], b0 = g->dyb[IX(m - 1, n)];
    *p = (dx1 * dx1 * mu[IX(m, n + 1)] * hq[IX(m, n + 1)] * str_t[IX(m, n + 1)] -
          dx0 * dx0 * mu[IX(m, n)] * hq[IX(m, n)] * str_t[IX(m, n)]) / g->dxh[IX(m, n)];
    *q = (b1 * b1 * muh_p * hh[IX(m, n)] * str_s[IX(m, n)] -
          b0 * b0 * muh_m * hh[IX(m - 1, n)] * str_s[IX(m - 1, n)]) / g->dyt[IX(m, n)];
}

/* Function/vel_ssh.f90:117-184 */
void bto_uv_diff2(const bto_grid *g, const double *mu, const double *str_t,
                  const double *str_s, const double *hq, const double *hu,
                  const double *hv, const double *hh, double *rhsx,
                  double *rhsy)
{
    (void)hu; (void)hv;
#pragma omp parallel for schedule(static)
    for (int n = NYS; n <= NYE; ++n)
        for (int m = NXS; m <= NXE; ++m) {
            double p, q;
            if (g->lcu[IX(m, n)] > 0.5f) {
                visc_x(g, mu, str_t, str_s, hq, hh, m, n, &p, &q);
                rhsx[IX(m, n)] = p + q;
            }
            if (g->lcv[IX(m, n)] > 0.5f) {
                visc_y(g, mu, str_t, str_s, hq, hh, m, n, &p, &q);
                rhsy[IX(m, n)] = -p + q;
            }
        }
}

/* Function/vel_ssh.f90:187-317 */
void bto_uv_diff4(const bto_grid *g, const double *mu, double *str_t,
                  double *str_s, double *fx, double *fy, const double *hq,
                  const double *hu, const double *hv, const double *hh,
                  double *rhsx, double *rhsy)
{
    size_t sz = (size_t)g->nx * g->ny;
    memset(fx, 0, sz * sizeof(double));
    memset(fy, 0, sz * sizeof(double));
#pragma omp parallel for schedule(static)
    for (int n = NYS; n <= NYE; ++n)
        for (int m = NXS; m <= NXE; ++m) {
            double p, q;
            if (g->lcu[IX(m, n)] > 0.5f) {
                visc_x(g, mu, str_t, str_s, hq, hh, m, n, &p, &q);
                double f = p + q;
                fx[IX(m, n)] = -f / hu[IX(m, n)] / g->dxt[IX(m, n)] / g->dyh[IX(m, n)];
            }
            if (g->lcv[IX(m, n)] > 0.5f) {
                visc_y(g, mu, str_t, str_s, hq, hh, m, n, &p, &q);
                double f = -p + q;
                fy[IX(m, n)] = -f / hv[IX(m, n)] / g->dxh[IX(m, n)] / g->dyt[IX(m, n)];
            }
        }
    cyclize(g, fx);
    cyclize(g, fy);
    bto_stress_components(g, fx, fy, str_t, str_s);
#pragma omp parallel for schedule(static)
    for (int n = NYS; n <= NYE; ++n)
        for (int m = NXS; m <= NXE; ++m) {
            double p, q;
            if (g->lcu[IX(m, n)] > 0.5f) {
                visc_x(g, mu, str_t, str_s, hq, hh, m, n, &p, &q);
                rhsx[IX(m, n)] = rhsx[IX(m, n)] + p + q;
            }
            if (g->lcv[IX(m, n)] > 0.5f) {
                visc_y(g, mu, str_t, str_s, hq, hh, m, n, &p, &q);
                rhsy[IX(m, n)] = rhsy[IX(m, n)] - p + q;
            }
        }
}

/* Function/vel_ssh.f90:379-679 */
void bto_barotropic_dynamics(const bto_grid *g, bto_depths *d, double tau,
                             int nstep, int ksw4, double *ubrtr_e,
                             double *ubrtrp_e, double *vbrtr_e,
                             double *vbrtrp_e, double *ssh_e, double *sshp_e,
                             double *ubrtr_i, double *vbrtr_i,
                             double *ssh4gradx, double *ssh4grady,
                             const double *RHSx, const double *RHSy,
                             const double *wflux, const double *mu,
                             const double *mu4, double *vort, double *str_t2d,
                             double *str_s2d, double *fx, double *fy,
                             const double *rdis, double *RHSx_adv,
                             double *RHSy_adv, double *RHSx_dif,
                             double *RHSy_dif)
{
    const size_t sz = (size_t)g->nx * g->ny;
    const size_t bytes = sz * sizeof(double);
    double *u = malloc(bytes), *v = malloc(bytes), *ssh = malloc(bytes);
    double *up = malloc(bytes), *vp = malloc(bytes), *sshp = malloc(bytes);
    double *un = calloc(sz, sizeof(double)), *vn = calloc(sz, sizeof(double));
    double *sshn = calloc(sz, sizeof(double));
    double *hhq_e = d->h[0], *hhqn_e = d->h[2];
    double *hhu_e = d->h[3], *hhup_e = d->h[4], *hhun_e = d->h[5];
    double *hhv_e = d->h[6], *hhvp_e = d->h[7], *hhvn_e = d->h[8];
    double *hhh_e = d->h[9], *hhhn_e = d->h[11];
    const double tau_inner = tau / (double)nstep;
    const double g_acc = g->free_fall_acc, rho0 = g->ref_den, ts = g->time_smooth;
    const double ffs = (double)g->full_free_surface;
    const double d4n = (double)(4 * nstep), dn = (double)nstep;

    memcpy(up, ubrtrp_e, bytes);
    memcpy(vp, vbrtrp_e, bytes);
    memcpy(sshp, sshp_e, bytes);
    memcpy(u, ubrtr_e, bytes);
    memcpy(v, vbrtr_e, bytes);
    memcpy(ssh, ssh_e, bytes);
    memset(ubrtr_i, 0, bytes);
    memset(vbrtr_i, 0, bytes);
    memset(ssh4gradx, 0, bytes);
    memset(ssh4grady, 0, bytes);

    for (int step = 1; step <= 2 * nstep; ++step) {
        /* continuity, :487-499 */
#pragma omp parallel for schedule(static)
        for (int n = NYS; n <= NYE; ++n)
            for (int m = NXS; m <= NXE; ++m)
                if (g->lu[IX(m, n)] > 0.5f)
                    sshn[IX(m, n)] =
                        sshp[IX(m, n)] +
                        2.0 * tau_inner *
                            (wflux[IX(m, n)] / rho0 * ffs -
                             (u[IX(m, n)] * hhu_e[IX(m, n)] * g->dyh[IX(m, n)] -
                              u[IX(m - 1, n)] * hhu_e[IX(m - 1, n)] * g->dyh[IX(m - 1, n)] +
                              v[IX(m, n)] * hhv_e[IX(m, n)] * g->dxh[IX(m, n)] -
                              v[IX(m, n - 1)] * hhv_e[IX(m, n - 1)] * g->dxh[IX(m, n - 1)]) /
                                 (g->dx[IX(m, n)] * g->dy[IX(m, n)]));
        cyclize(g, sshn);

        if (g->full_free_surface > 0) bto_hh_update(g, hhqn_e, hhun_e, hhvn_e, hhhn_e, sshn);

        bto_stress_components(g, up, vp, str_t2d, str_s2d);
        bto_uv_trans(g, u, v, vort, hhq_e, hhu_e, hhv_e, hhh_e, RHSx_adv, RHSy_adv);
        bto_uv_diff2(g, mu, str_t2d, str_s2d, hhq_e, hhu_e, hhv_e, hhh_e, RHSx_dif, RHSy_dif);
        if (ksw4 > 0)
            bto_uv_diff4(g, mu4, str_t2d, str_s2d, fx, fy, hhq_e, hhu_e, hhv_e, hhh_e,
                         RHSx_dif, RHSy_dif);

        /* momentum, :531-573 */
#pragma omp parallel for schedule(static)
        for (int n = NYS; n <= NYE; ++n)
            for (int m = NXS; m <= NXE; ++m) {
                size_t i = IX(m, n);
                if (g->lcu[i] > 0.5f) {
                    double bp = hhun_e[i] * g->dxt[i] * g->dyh[i] / 2.0 / tau_inner;
                    double bp0 = hhup_e[i] * g->dxt[i] * g->dyh[i] / 2.0 / tau_inner;
                    double slx = -g_acc * (ssh[IX(m + 1, n)] - ssh[i]) * g->dyh[i] * hhu_e[i];
                    double slxn = -g_acc * (sshn[IX(m + 1, n)] - sshn[i]) * g->dyh[i] * hhun_e[i];
                    double grx = RHSx[i] + slx + RHSx_dif[i] + RHSx_adv[i] -
                                 (rdis[i] + rdis[IX(m + 1, n)]) / 2.0 * up[i] * g->dxt[i] *
                                     g->dyh[i] * hhu_e[i] +
                                 (g->rlh_s[i] * hhh_e[i] * g->dxb[i] * g->dyb[i] *
                                      (v[IX(m + 1, n)] + v[i]) +
                                  g->rlh_s[IX(m, n - 1)] * hhh_e[IX(m, n - 1)] * g->dxb[IX(m, n - 1)] *
                                      g->dyb[IX(m, n - 1)] * (v[IX(m + 1, n - 1)] + v[IX(m, n - 1)])) /
                                     4.0;
                    un[i] = (up[i] * bp0 + grx) / bp;
                    ubrtr_i[i] = ubrtr_i[i] + (u[i] * hhu_e[i] + un[i] * hhun_e[i]) / d4n;
                    ssh4gradx[i] = ssh4gradx[i] + (slx + slxn) / d4n;
                }
                if (g->lcv[i] > 0.5f) {
                    double bp = hhvn_e[i] * g->dyt[i] * g->dxh[i] / 2.0 / tau_inner;
                    double bp0 = hhvp_e[i] * g->dyt[i] * g->dxh[i] / 2.0 / tau_inner;
                    double sly = -g_acc * (ssh[IX(m, n + 1)] - ssh[i]) * g->dxh[i] * hhv_e[i];
                    double slyn = -g_acc * (sshn[IX(m, n + 1)] - sshn[i]) * g->dxh[i] * hhvn_e[i];
                    double gry = RHSy[i] + sly + RHSy_dif[i] + RHSy_adv[i] -
                                 (rdis[i] + rdis[IX(m, n + 1)]) / 2.0 * vp[i] * g->dxh[i] *
                                     g->dyt[i] * hhv_e[i] -
                                 (g->rlh_s[i] * hhh_e[i] * g->dxb[i] * g->dyb[i] *
                                      (u[IX(m, n + 1)] + u[i]) +
                                  g->rlh_s[IX(m - 1, n)] * hhh_e[IX(m - 1, n)] * g->dxb[IX(m - 1, n)] *
                                      g->dyb[IX(m - 1, n)] * (u[IX(m - 1, n + 1)] + u[IX(m - 1, n)])) /
                                     4.0;
                    vn[i] = (vp[i] * bp0 + gry) / bp;
                    vbrtr_i[i] = vbrtr_i[i] + (v[i] * hhv_e[i] + vn[i] * hhvn_e[i]) / d4n;
                    ssh4grady[i] = ssh4grady[i] + (sly + slyn) / d4n;
                }
            }
        cyclize(g, un);
        cyclize(g, vn);

        /* time-level shift with the Robert-Asselin filter, :588-612 */
#pragma omp parallel for schedule(static)
        for (int n = NYS - 1; n <= NYE + 1; ++n)
            for (int m = NXS - 1; m <= NXE + 1; ++m) {
                size_t i = IX(m, n);
                if (g->lu[i] > 0.5f) {
                    sshp[i] = ssh[i] + ts * (sshn[i] - 2.0 * ssh[i] + sshp[i]) / 2.0 / dn;
                    ssh[i] = sshn[i];
                }
                if (g->lcu[i] > 0.5f) {
                    up[i] = u[i] + ts * (un[i] - 2.0 * u[i] + up[i]) / 2.0 / dn;
                    u[i] = un[i];
                }
                if (g->lcv[i] > 0.5f) {
                    vp[i] = v[i] + ts * (vn[i] - 2.0 * v[i] + vp[i]) / 2.0 / dn;
                    v[i] = vn[i];
                }
            }

        if (g->full_free_surface > 0) bto_hh_shift(g, d, nstep);

        if (step == nstep) {
            memcpy(ubrtrp_e, up, bytes);
            memcpy(vbrtrp_e, vp, bytes);
            memcpy(sshp_e, sshp, bytes);
            memcpy(ubrtr_e, u, bytes);
            memcpy(vbrtr_e, v, bytes);
            memcpy(ssh_e, ssh, bytes);
        }
    }
    cyclize(g, ubrtr_i);
    cyclize(g, vbrtr_i);
    if (g->full_free_surface > 0) bto_hh_init(g, d, ssh_e, sshp_e);
    (void)hhq_e;
    free(sshn); free(vn); free(un); free(sshp); free(vp); free(up);
    free(ssh); free(v); free(u);
}

/* Function/fluxes.f90:29-37 */
void bto_surface_stress(const bto_grid *g, const double *taux,
                        const double *tauy, double *surf_x, double *surf_y)
{
    const double rho0 = g->ref_den;
#pragma omp parallel for schedule(static)
    for (int n = NYS; n <= NYE; ++n)
        for (int m = NXS; m <= NXE; ++m) {
            size_t i = IX(m, n);
            if (g->lcu[i] > 0.5f)
                surf_x[i] = (taux[i] * g->dx[i] * g->dy[i] +
                             taux[IX(m + 1, n)] * g->dx[IX(m + 1, n)] * g->dy[IX(m + 1, n)]) /
                            rho0 / 2.0 / g->dxt[i] / g->dyh[i];
            if (g->lcv[i] > 0.5f)
                surf_y[i] = (tauy[i] * g->dx[i] * g->dy[i] +
                             tauy[IX(m, n + 1)] * g->dx[IX(m, n + 1)] * g->dy[IX(m, n + 1)]) /
                            rho0 / 2.0 / g->dxh[i] / g->dyt[i];
        }
}

/* Function/fluxes.f90:226-272 (type_fric = 2) */
void bto_bottom_stress(const bto_grid *g, double cb_nl, double ebottom,
                       const double *uu, const double *vv, const double *uup,
                       const double *vvp, const double *hhu, const double *hhv,
                       const double *hhq, double *bf, double *bot_x,
                       double *bot_y)
{
#pragma omp parallel for schedule(static)
    for (int n = NYS; n <= NYE; ++n)
        for (int m = NXS; m <= NXE; ++m) {
            size_t i = IX(m, n);
            if (g->lu[i] > 0.5f) {
                double ut = (uu[i] * g->dxt[i] * g->dyh[i] * hhu[i] +
                             uu[IX(m - 1, n)] * g->dxt[IX(m - 1, n)] * g->dyh[IX(m - 1, n)] *
                                 hhu[IX(m - 1, n)]) /
                            2.0 / g->dx[i] / g->dy[i] / hhq[i];
                double vt = (vv[i] * g->dxh[i] * g->dyt[i] * hhv[i] +
                             vv[IX(m, n - 1)] * g->dxh[IX(m, n - 1)] * g->dyt[IX(m, n - 1)] *
                                 hhv[IX(m, n - 1)]) /
                            2.0 / g->dx[i] / g->dy[i] / hhq[i];
                bf[i] = cb_nl * sqrt(ut * ut + vt * vt + ebottom);
            }
        }
    cyclize(g, bf);
#pragma omp parallel for schedule(static)
    for (int n = NYS; n <= NYE; ++n)
        for (int m = NXS; m <= NXE; ++m) {
            size_t i = IX(m, n);
            if (g->lcu[i] > 0.5f) bot_x[i] = -(bf[i] + bf[IX(m + 1, n)]) / 2.0 * uup[i];
            if (g->lcv[i] > 0.5f) bot_y[i] = -(bf[i] + bf[IX(m, n + 1)]) / 2.0 * vvp[i];
        }
}

/* Control/dyncall.f90:169-184 */
void bto_remove_barotropic(const bto_grid *g, int nz, double *uu, double *vv,
                           const double *uu2d, const double *vv2d)
{
    const size_t plane = (size_t)g->nx * g->ny;
#pragma omp parallel for schedule(static)
    for (int n = NYS - 1; n <= NYE + 1; ++n)
        for (int m = NXS - 1; m <= NXE + 1; ++m) {
            if (g->lcu[IX(m, n)] > 0.5f)
                for (int k = 0; k < nz; ++k) uu[k * plane + IX(m, n)] = uu[k * plane + IX(m, n)] - uu2d[IX(m, n)];
            if (g->lcv[IX(m, n)] > 0.5f)
                for (int k = 0; k < nz; ++k) vv[k * plane + IX(m, n)] = vv[k * plane + IX(m, n)] - vv2d[IX(m, n)];
        }
}

/* Control/dyncall.f90:188-231 (vertical_velocity excluded) with
 * ssh_internal, Function/vel_ssh.f90:682-751 */
void bto_internal_update(const bto_grid *g, double tau, int nz,
                         const double *ubrtr_i, const double *vbrtr_i,
                         const double *hhu, const double *hhv, const double *slpr,
                         const double *wflux, double *uu2d, double *vv2d,
                         double *pgrx, double *pgry, double *ssh_i, double *sshp_i,
                         double *hhqn, double *hhun, double *hhvn, double *hhhn,
                         double *uu, double *vv)
{
    const size_t plane = (size_t)g->nx * g->ny;
    const double rho0 = g->ref_den;
    /* dyncall.f90:188-206 */
#pragma omp parallel for schedule(static)
    for (int n = NYS - 1; n <= NYE + 1; ++n)
        for (int m = NXS - 1; m <= NXE + 1; ++m) {
            if (g->lcu[IX(m, n)] > 0.5f) {
                uu2d[IX(m, n)] = ubrtr_i[IX(m, n)] / hhu[IX(m, n)];
                pgrx[IX(m, n)] = pgrx[IX(m, n)] -
                                 (slpr[IX(m + 1, n)] - slpr[IX(m, n)]) * hhu[IX(m, n)] * g->dyh[IX(m, n)] / rho0;
            }
            if (g->lcv[IX(m, n)] > 0.5f) {
                vv2d[IX(m, n)] = vbrtr_i[IX(m, n)] / hhv[IX(m, n)];
                pgry[IX(m, n)] = pgry[IX(m, n)] -
                                 (slpr[IX(m, n + 1)] - slpr[IX(m, n)]) * hhv[IX(m, n)] * g->dxh[IX(m, n)] / rho0;
            }
        }
    /* ssh_internal: vel_ssh.f90:705-720 */
    double *ssh_new = (double *)calloc(plane, sizeof(double));
    const double ffs = (double)g->full_free_surface;
#pragma omp parallel for schedule(static)
    for (int n = NYS; n <= NYE; ++n)
        for (int m = NXS; m <= NXE; ++m)
            if (g->lu[IX(m, n)] > 0.5f)
                ssh_new[IX(m, n)] =
                    sshp_i[IX(m, n)] +
                    2.0 * tau *
                        (wflux[IX(m, n)] / rho0 * ffs -
                         (ubrtr_i[IX(m, n)] * g->dyh[IX(m, n)] - ubrtr_i[IX(m - 1, n)] * g->dyh[IX(m - 1, n)] +
                          vbrtr_i[IX(m, n)] * g->dxh[IX(m, n)] - vbrtr_i[IX(m, n - 1)] * g->dxh[IX(m, n - 1)]) /
                             (g->dx[IX(m, n)] * g->dy[IX(m, n)]));
    /* :722-735 */
    cyclize(g, ssh_new);
    if (g->full_free_surface > 0) bto_hh_update(g, hhqn, hhun, hhvn, hhhn, ssh_new);
    /* :737-748 */
    const double ts = g->time_smooth;
#pragma omp parallel for schedule(static)
    for (int n = NYS - 1; n <= NYE + 1; ++n)
        for (int m = NXS - 1; m <= NXE + 1; ++m)
            if (g->lu[IX(m, n)] > 0.5f) {
                sshp_i[IX(m, n)] = ssh_i[IX(m, n)] + ts * (ssh_new[IX(m, n)] - 2.0 * ssh_i[IX(m, n)] + sshp_i[IX(m, n)]) / 2.0;
                ssh_i[IX(m, n)] = ssh_new[IX(m, n)];
            }
    free(ssh_new);
    /* dyncall.f90:216-231 */
#pragma omp parallel for schedule(static)
    for (int n = NYS - 1; n <= NYE + 1; ++n)
        for (int m = NXS - 1; m <= NXE + 1; ++m) {
            if (g->lcu[IX(m, n)] > 0.5f)
                for (int k = 0; k < nz; ++k) uu[k * plane + IX(m, n)] = uu[k * plane + IX(m, n)] + uu2d[IX(m, n)];
            if (g->lcv[IX(m, n)] > 0.5f)
                for (int k = 0; k < nz; ++k) vv[k * plane + IX(m, n)] = vv[k * plane + IX(m, n)] + vv2d[IX(m, n)];
        }
}

/* Control/dyncall.f90:110-130 for a 2-D flow (all 3-D terms zero). */
void bto_rhs_assembly(const bto_grid *g, const double *surf_x,
                      const double *surf_y, const double *bot_x,
                      const double *bot_y, double *rhsx2d, double *rhsy2d)
{
    size_t sz = (size_t)g->nx * g->ny;
    /* depth_ave of a zero 3-D RHS leaves zero on llu/llv (dyncall.f90:103-104) */
    for (size_t i = 0; i < sz; ++i) {
        if (g->llu[i] > 0.5f) rhsx2d[i] = 0.0;
        if (g->llv[i] > 0.5f) rhsy2d[i] = 0.0;
    }
#pragma omp parallel for schedule(static)
    for (int n = NYS; n <= NYE; ++n)
        for (int m = NXS; m <= NXE; ++m) {
            size_t i = IX(m, n);
            if (g->lcu[i] > 0.5f)
                rhsx2d[i] = rhsx2d[i] + (surf_x[i] + bot_x[i]) * g->dxt[i] * g->dyh[i];
            if (g->lcv[i] > 0.5f)
                rhsy2d[i] = rhsy2d[i] + (surf_y[i] + bot_y[i]) * g->dyt[i] * g->dxh[i];
        }
}

/* Control/dyncall.f90:37-130 without the vertical mixing, Smagorinsky,
 * isopycnal slopes and biharmonic branches (ksw_vert = ksw_lat = 0,
 * ksw_lat4 = 0): stress_components on the level n-1 3-D and 2-D
 * velocities, sea_surface_fluxes and sea_bottom_fluxes (the bottom level
 * k = nz of the 3-D velocities), uv_trans and uv_diff2 at nz levels and on
 * the 2-D velocities, the caller's 3-D pressure-gradient terms (pressure_
 * gradients, NULL = none), depth_ave with dz and the assembly of
 * :109-126.  3-D arrays are (nz, ny, nx); rhsx2d, rhsy2d are read and
 * written (cells the reference does not write keep their values). */
void bto_forcing_3d(const bto_grid *g, int nz, const double *dz, double cb_nl, double ebottom,
                    const double *uu, const double *vv, const double *uup, const double *vvp,
                    const double *amuv, const double *uu2d, const double *vv2d,
                    const double *uup2d, const double *vvp2d, const double *amuv2d,
                    const double *hhq, const double *hhu, const double *hhv, const double *hhh,
                    const double *taux, const double *tauy, const double *slpr,
                    const double *pgx, const double *pgy, double *rhsx2d, double *rhsy2d)
{
    const size_t sz = (size_t)g->nx * g->ny, s3 = sz * (size_t)nz;
    double *st3 = calloc(s3, sizeof(double)), *ss3 = calloc(s3, sizeof(double));
    double *vort3 = calloc(s3, sizeof(double));
    double *tx3 = calloc(s3, sizeof(double)), *ty3 = calloc(s3, sizeof(double));
    double *rx3 = calloc(s3, sizeof(double)), *ry3 = calloc(s3, sizeof(double));
    double *w2[12];
    for (int k = 0; k < 12; ++k) w2[k] = calloc(sz, sizeof(double));
    double *st2 = w2[0], *ss2 = w2[1], *vort2 = w2[2], *txd = w2[3], *tyd = w2[4];
    double *dfx = w2[5], *dfy = w2[6], *sfx = w2[7], *sfy = w2[8], *bfx = w2[9], *bfy = w2[10];
    double *bf = w2[11];
    double *tx2 = calloc(sz, sizeof(double)), *ty2 = calloc(sz, sizeof(double));

    /* dyncall.f90:38-39 */
    for (int k = 0; k < nz; ++k)
        bto_stress_components(g, uup + k * sz, vvp + k * sz, st3 + k * sz, ss3 + k * sz);
    bto_stress_components(g, uup2d, vvp2d, st2, ss2);
    /* :53-59 */
    bto_surface_stress(g, taux, tauy, sfx, sfy);
    bto_bottom_stress(g, cb_nl, ebottom, uu + (size_t)(nz - 1) * sz, vv + (size_t)(nz - 1) * sz,
                      uup + (size_t)(nz - 1) * sz, vvp + (size_t)(nz - 1) * sz, hhu, hhv, hhq, bf, bfx, bfy);
    /* :65-81 (RHSx3d and RHSx2d_diff_disp start at zero, :65-66) */
    for (int k = 0; k < nz; ++k) {
        bto_uv_trans(g, uu + k * sz, vv + k * sz, vort3 + k * sz, hhq, hhu, hhv, hhh, tx3 + k * sz, ty3 + k * sz);
        bto_uv_diff2(g, amuv + k * sz, st3 + k * sz, ss3 + k * sz, hhq, hhu, hhv, hhh, rx3 + k * sz,
                     ry3 + k * sz);
    }
    bto_uv_trans(g, uu2d, vv2d, vort2, hhq, hhu, hhv, hhh, txd, tyd);
    bto_uv_diff2(g, amuv2d, st2, ss2, hhq, hhu, hhv, hhh, dfx, dfy);
    /* :92-98, the caller's pressure_gradients terms */
    if (pgx && pgy)
        for (size_t i = 0; i < s3; ++i) {
            rx3[i] = rx3[i] + pgx[i];
            ry3[i] = ry3[i] + pgy[i];
        }
    /* :100-104, depth_ave (Function/depth.f90:223-254) with kbcl = 0 */
    const double *src[4] = {rx3, ry3, tx3, ty3};
    double *dst[4] = {rhsx2d, rhsy2d, tx2, ty2};
    const float *msk[4] = {g->llu, g->llv, g->llu, g->llv};
    for (int q = 0; q < 4; ++q)
        for (int n = 2; n <= g->ny - 1; ++n)
            for (int m = 2; m <= g->nx - 1; ++m) {
                const size_t i = IX(m, n);
                if (msk[q][i] > 0.5f) {
                    double f = 0.0;
                    for (int k = 0; k < nz; ++k) f = f + src[q][k * sz + i] * dz[k];
                    dst[q][i] = f;
                }
            }
    /* :109-126 */
    const double rho0 = g->ref_den;
    for (int n = NYS; n <= NYE; ++n)
        for (int m = NXS; m <= NXE; ++m) {
            const size_t i = IX(m, n);
            if (g->lcu[i] > 0.5f)
                rhsx2d[i] = rhsx2d[i] - dfx[i] + tx2[i] - txd[i] + (sfx[i] + bfx[i]) * g->dxt[i] * g->dyh[i] -
                            (slpr[IX(m + 1, n)] - slpr[i]) * hhu[i] * g->dyh[i] / rho0;
            if (g->lcv[i] > 0.5f)
                rhsy2d[i] = rhsy2d[i] - dfy[i] + ty2[i] - tyd[i] + (sfy[i] + bfy[i]) * g->dyt[i] * g->dxh[i] -
                            (slpr[IX(m, n + 1)] - slpr[i]) * hhv[i] * g->dxh[i] / rho0;
        }
    free(st3); free(ss3); free(vort3); free(tx3); free(ty3); free(rx3); free(ry3);
    for (int k = 0; k < 12; ++k) free(w2[k]);
    free(tx2); free(ty2);
}

/* Service/oc_algorithms.f90:494-524 (factor8): the sweep for
 * a(j) f(j-1) + b(j) f(j) + c(j) f(j+1) = eta(j), j = 1..n. */
static void factor8(int n, const double *a, const double *b, const double *c, const double *eta, double *rksi,
                    double *x, double *y)
{
    x[0] = -(c[0] / b[0]);
    y[0] = eta[0] / b[0];
    for (int j = 1; j < n; ++j) {
        const double scr = 1.0 / (a[j] * x[j - 1] + b[j]);
        x[j] = -(c[j] * scr);
        y[j] = (eta[j] - a[j] * y[j - 1]) * scr;
    }
    rksi[n - 1] = y[n - 1];
    for (int j = n - 2; j >= 0; --j) rksi[j] = x[j] * rksi[j + 1] + y[j];
}

/* Function/vel_ssh.f90:794-1027 (baroclinic_dynamics): per open U (V) face
 * the implicit vertical viscosity/advection column of nz levels, solved
 * by factor8; un, vn wrapped on periodic boundaries; then the
 * Robert-Asselin step of uup, vvp and uu = un, vv = vn over the interior
 * and its one-point ring.  3-D arrays (nz, ny, nx); ww and nu (nz+1, ny,
 * nx); dz[nz], hzt[nz+1]. */
void bto_baroclinic_dynamics(const bto_grid *g, int nz, double tau, const double *dz, const double *hzt,
                             double *uu, double *uup, double *vv, double *vvp, const double *ww,
                             const double *nu, const double *rhsx, const double *rhsy,
                             const double *rhsx_tran, const double *rhsy_tran, const double *pgrx,
                             const double *pgry, const double *rdis, const double *tx_surf,
                             const double *ty_surf, const double *tx_bot, const double *ty_bot,
                             const double *hhu, const double *hhup, const double *hhun, const double *hhv,
                             const double *hhvp, const double *hhvn, const double *hhh)
{
    const size_t sz = (size_t)g->nx * g->ny;
    double *un = calloc(sz * nz, sizeof(double)), *vn = calloc(sz * nz, sizeof(double));
    double *a = malloc(6 * nz * sizeof(double));
    double *b = a + nz, *c = b + nz, *eta = c + nz, *x = eta + nz, *y = x + nz;
    double *rksi = malloc(nz * sizeof(double));
#define L3(f, m, n, k) (f)[(size_t)((k) - 1) * sz + IX(m, n)]
    for (int n = NYS; n <= NYE; ++n)
        for (int m = NXS; m <= NXE; ++m) {
            const size_t i = IX(m, n);
            if (g->lcu[i] > 0.5f) {
                const double bp = hhun[i] * g->dxt[i] * g->dyh[i] / tau / 2.0;
                const double bp0 = hhup[i] * g->dxt[i] * g->dyh[i] / tau / 2.0;
                double dm, dp = 0.0, fz_m, fz_p = 0.0;
                for (int k = 1; k <= nz; ++k) {
                    dm = (k == 1) ? 0.0 : dp;
                    fz_m = (k == 1) ? 0.0 : fz_p;
                    if (k < nz) {
                        dp = (L3(nu, m, n, k + 1) + L3(nu, m + 1, n, k + 1)) / 2.0 / hhun[i] / hzt[k] * g->dxt[i] *
                             g->dyh[i];
                        fz_p = (L3(ww, m, n, k + 1) * g->dx[i] * g->dy[i] +
                                L3(ww, m + 1, n, k + 1) * g->dx[IX(m + 1, n)] * g->dy[IX(m + 1, n)]) /
                               2.0 * (L3(uu, m, n, k) + L3(uu, m, n, k + 1)) / 2.0;
                    } else {
                        dp = 0.0;
                        fz_p = 0.0;
                    }
                    const double dzk = dz[k - 1];
                    c[k - 1] = (k < nz) ? -(dp / dzk) : 0.0;
                    a[k - 1] = (k == 1) ? 0.0 : -(dm / dzk);
                    b[k - 1] = (k == 1) ? bp + dp / dzk : (k == nz) ? bp + dm / dzk : bp + (dp + dm) / dzk;
                    const double up = L3(uup, m, n, k);
                    double e = bp0 * up + L3(rhsx, m, n, k) + L3(rhsx_tran, m, n, k) + pgrx[i] -
                               (fz_p - fz_m) / dzk +
                               (g->rlh_s[i] * hhh[i] * g->dxb[i] * g->dyb[i] *
                                    (L3(vv, m + 1, n, k) + L3(vv, m, n, k)) +
                                g->rlh_s[IX(m, n - 1)] * hhh[IX(m, n - 1)] * g->dxb[IX(m, n - 1)] *
                                    g->dyb[IX(m, n - 1)] * (L3(vv, m + 1, n - 1, k) + L3(vv, m, n - 1, k))) /
                                   4.0 -
                               (rdis[i] + rdis[IX(m + 1, n)]) / 2.0 * up * g->dxt[i] * g->dyh[i] * hhu[i];
                    if (k == 1) e = e + tx_surf[i] / dzk * g->dxt[i] * g->dyh[i];
                    if (k == nz) e = e + tx_bot[i] / dzk * g->dxt[i] * g->dyh[i];
                    eta[k - 1] = e;
                }
                factor8(nz, a, b, c, eta, rksi, x, y);
                for (int k = 1; k <= nz; ++k) L3(un, m, n, k) = rksi[k - 1];
            }
            if (g->lcv[i] > 0.5f) {
                const double bp = hhvn[i] * g->dxh[i] * g->dyt[i] / tau / 2.0;
                const double bp0 = hhvp[i] * g->dxh[i] * g->dyt[i] / tau / 2.0;
                double dm, dp = 0.0, fz_m, fz_p = 0.0;
                for (int k = 1; k <= nz; ++k) {
                    dm = (k == 1) ? 0.0 : dp;
                    fz_m = (k == 1) ? 0.0 : fz_p;
                    if (k < nz) {
                        dp = (L3(nu, m, n, k + 1) + L3(nu, m, n + 1, k + 1)) / 2.0 / hhvn[i] / hzt[k] * g->dxh[i] *
                             g->dyt[i];
                        fz_p = (L3(ww, m, n, k + 1) * g->dx[i] * g->dy[i] +
                                L3(ww, m, n + 1, k + 1) * g->dx[IX(m, n + 1)] * g->dy[IX(m, n + 1)]) /
                               2.0 * (L3(vv, m, n, k) + L3(vv, m, n, k + 1)) / 2.0;
                    } else {
                        dp = 0.0;
                        fz_p = 0.0;
                    }
                    const double dzk = dz[k - 1];
                    c[k - 1] = (k < nz) ? -(dp / dzk) : 0.0;
                    a[k - 1] = (k == 1) ? 0.0 : -(dm / dzk);
                    b[k - 1] = (k == 1) ? bp + dp / dzk : (k == nz) ? bp + dm / dzk : bp + (dp + dm) / dzk;
                    const double vp = L3(vvp, m, n, k);
                    double e = bp0 * vp + L3(rhsy, m, n, k) + L3(rhsy_tran, m, n, k) + pgry[i] -
                               (fz_p - fz_m) / dzk -
                               (g->rlh_s[i] * hhh[i] * g->dxb[i] * g->dyb[i] *
                                    (L3(uu, m, n + 1, k) + L3(uu, m, n, k)) +
                                g->rlh_s[IX(m - 1, n)] * hhh[IX(m - 1, n)] * g->dxb[IX(m - 1, n)] *
                                    g->dyb[IX(m - 1, n)] * (L3(uu, m - 1, n + 1, k) + L3(uu, m - 1, n, k))) /
                                   4.0 -
                               (rdis[i] + rdis[IX(m, n + 1)]) / 2.0 * vp * g->dxh[i] * g->dyt[i] * hhv[i];
                    if (k == 1) e = e + ty_surf[i] / dzk * g->dxh[i] * g->dyt[i];
                    if (k == nz) e = e + ty_bot[i] / dzk * g->dxh[i] * g->dyt[i];
                    eta[k - 1] = e;
                }
                factor8(nz, a, b, c, eta, rksi, x, y);
                for (int k = 1; k <= nz; ++k) L3(vn, m, n, k) = rksi[k - 1];
            }
        }
    for (int k = 0; k < nz; ++k) {
        cyclize(g, un + (size_t)k * sz);
        cyclize(g, vn + (size_t)k * sz);
    }
    const double ts = g->time_smooth;
    for (int n = NYS - 1; n <= NYE + 1; ++n)
        for (int m = NXS - 1; m <= NXE + 1; ++m) {
            const size_t i = IX(m, n);
            if (g->lcu[i] > 0.5f)
                for (int k = 1; k <= nz; ++k) {
                    L3(uup, m, n, k) = L3(uu, m, n, k) + ts * (L3(un, m, n, k) - 2.0 * L3(uu, m, n, k) +
                                                               L3(uup, m, n, k)) / 2.0;
                    L3(uu, m, n, k) = L3(un, m, n, k);
                }
            if (g->lcv[i] > 0.5f)
                for (int k = 1; k <= nz; ++k) {
                    L3(vvp, m, n, k) = L3(vv, m, n, k) + ts * (L3(vn, m, n, k) - 2.0 * L3(vv, m, n, k) +
                                                               L3(vvp, m, n, k)) / 2.0;
                    L3(vv, m, n, k) = L3(vn, m, n, k);
                }
        }
#undef L3
    free(un);
    free(vn);
    free(a);
    free(rksi);
}

/* Function/vel_ssh.f90:1030-1076 (vertical_velocity): w from the
 * continuity equation, integrated from the bottom (w(nz+1) = 0) on lu
 * interior columns; then the periodic wrap of all nz+1 levels.  u, v
 * (nz, ny, nx), w (nz+1, ny, nx) read and written (cells not computed
 * keep their values). */
void bto_vertical_velocity(const bto_grid *g, int nz, const double *dz, const double *u, const double *v, double *w,
                           const double *hu, const double *hv, const double *wflux)
{
    const size_t sz = (size_t)g->nx * g->ny;
    const double ffs = (double)g->full_free_surface;
    for (int n = NYS; n <= NYE; ++n)
        for (int m = NXS; m <= NXE; ++m) {
            const size_t i = IX(m, n), iw = IX(m - 1, n), is = IX(m, n - 1);
            if (g->lu[i] > 0.5f) {
                w[(size_t)nz * sz + i] = 0.0;
                for (int k = nz; k >= 1; --k) {
                    const size_t o = (size_t)(k - 1) * sz;
                    w[o + i] = w[o + sz + i] + dz[k - 1] / (g->dx[i] * g->dy[i]) *
                                                   (u[o + i] * hu[i] * g->dyh[i] - u[o + iw] * hu[iw] * g->dyh[iw] +
                                                    v[o + i] * hv[i] * g->dxh[i] - v[o + is] * hv[is] * g->dxh[is] +
                                                    wflux[i] / g->ref_den * ffs);
                }
            }
        }
    for (int k = 0; k <= nz; ++k) cyclize(g, w + (size_t)k * sz);
}

/* Function/mixing.f90:171-219 (smagorinsky_coeff) followed by
 * Function/depth.f90:223-254 (depth_ave(amuv, amuv2d, lu, kbcl = 0)), the
 * pair Control/dyncall.f90:40-43 runs when ksw_lat > 0.  On lu interior
 * points each level's viscosity is visc2 plus the Smagorinsky term
 * (kappa / pi)^2 dx dy sqrt(str_t^2 + <str_s^2>), <str_s^2> the mean of
 * the squared shear at the four H corners of the T cell (kappa = 0.4,
 * pi = 3.141592653589793, Inc/basinpar.fi:72,89), and the diffusivity is
 * dift plus a tenth of that term; both are wrapped on periodic
 * boundaries; then amuv2d = sum_k amuv dz on lu points of the interior and
 * its one-cell ring.  str_t, str_s, amts, amuv (nz, ny, nx); cells not
 * computed keep their values. */
void bto_smagorinsky(const bto_grid *g, int nz, double visc2, double dift, const double *str_t,
                     const double *str_s, const double *dz, double *amts, double *amuv, double *amuv2d)
{
    const size_t sz = (size_t)g->nx * g->ny;
    const double kp = 0.4 / 3.141592653589793;
    const double coef = kp * kp, prandtl = 0.1;
    for (int n = NYS; n <= NYE; ++n)
        for (int m = NXS; m <= NXE; ++m) {
            const size_t i = IX(m, n), iw = IX(m - 1, n), is = IX(m, n - 1), isw = IX(m - 1, n - 1);
            if (!(g->lu[i] > 0.5f)) continue;
            for (int k = 0; k < nz; ++k) {
                const size_t o = (size_t)k * sz;
                const double a = str_s[o + i], b = str_s[o + iw], c = str_s[o + is], d = str_s[o + isw];
                const double s2 = (a * a + b * b + c * c + d * d) / 4.0;
                const double t = str_t[o + i];
                const double mu = coef * g->dx[i] * g->dy[i] * sqrt(t * t + s2);
                amuv[o + i] = visc2 + mu;
                amts[o + i] = dift + mu * prandtl;
            }
        }
    for (int k = 0; k < nz; ++k) {
        cyclize(g, amuv + (size_t)k * sz);
        cyclize(g, amts + (size_t)k * sz);
    }
    for (int n = 2; n <= g->ny - 1; ++n)
        for (int m = 2; m <= g->nx - 1; ++m) {
            const size_t i = IX(m, n);
            if (!(g->lu[i] > 0.5f)) continue;
            double acc = 0.0;
            for (int k = 0; k < nz; ++k) acc = acc + amuv[(size_t)k * sz + i] * dz[k];
            amuv2d[i] = acc;
        }
}

/* Function/vel_ssh.f90:320-376 (pressure_gradients): the hydrostatic
 * pressure of density den (nz, ny, nx) integrated down the sigma levels
 * z[nz] between the two T columns of each lcu (x) and lcv (y) interior
 * point, the gradient subtracted from rhsx, rhsy (nz, ny, nx) in place;
 * hhq, hhu, hhv are the 3-D model's depths.  Cells not on lcu (lcv) keep
 * their values. */
void bto_pressure_gradients(const bto_grid *g, int nz, const double *z, const double *den, const double *hhq,
                            const double *hhu, const double *hhv, double *rhsx, double *rhsy)
{
    const size_t sz = (size_t)g->nx * g->ny;
    const double gr = g->free_fall_acc, rho = g->ref_den;
    for (int n = NYS; n <= NYE; ++n)
        for (int m = NXS; m <= NXE; ++m) {
            const size_t i = IX(m, n);
            for (int dir = 0; dir < 2; ++dir) {
                const float *mask = dir ? g->lcv : g->lcu;
                if (!(mask[i] > 0.5f)) continue;
                /* j: the other T column (east for u, north for v) */
                const size_t j = dir ? IX(m, n + 1) : IX(m + 1, n);
                const double h1 = hhq[j], h0 = hhq[i];
                const double fac = dir ? hhv[i] : hhu[i], met = dir ? g->dxh[i] : g->dyh[i];
                double *rhs = dir ? rhsy : rhsx;
                double ph = h1 * z[0] * den[j] - h0 * z[0] * den[i];
                double pl = h1 * z[0] * den[i] - h0 * z[0] * den[j];
                rhs[i] = rhs[i] - (ph - pl) * fac * met * gr / rho / 2.0;
                for (int k = 2; k <= nz; ++k) {
                    const size_t o = (size_t)(k - 1) * sz, op = (size_t)(k - 2) * sz;
                    ph = ph + (h1 * z[k - 1] * den[op + j] - h1 * z[k - 2] * den[o + j]) -
                         (h0 * z[k - 1] * den[op + i] - h0 * z[k - 2] * den[o + i]);
                    pl = h1 * z[k - 1] * den[o + i] - h0 * z[k - 1] * den[o + j];
                    rhs[o + i] = rhs[o + i] - (ph - pl) * fac * met * gr / rho / 2.0;
                }
            }
        }
}

/* Service/oc_algorithms.f90:138-161 (density): the Brydon et al. (1999,
 * J. Geophys. Res. 104(C1), 1537-1540) polynomial equation of state as the
 * reference evaluates it, the deviation from 1020 kg/m^3 at potential
 * temperature t [C], salinity s [psu] and pressure p_pa [Pa] (in MPa
 * inside; the pure-pressure term is left out there, as here).  The terms
 * are summed left to right as written in the reference. */
static double eos_brydon(double t, double s, double p_pa)
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

/* Control/dyncall.f90:437-449 (the density definition): on lu points of
 * m = 2..nx-1, n = 2..ny-1, every level, den at the pressure
 * FreeFallAcc * RefDen * hhq * z(k) and den_pot at zero pressure.  tt, ss,
 * den, den_pot (nz, ny, nx); cells off lu keep their values. */
void bto_density(const bto_grid *g, int nz, const double *z, const double *tt, const double *ss, const double *hhq,
                 double *den, double *den_pot)
{
    const size_t sz = (size_t)g->nx * g->ny;
    for (int n = NYS - 1; n <= NYE + 1; ++n)
        for (int m = NXS - 1; m <= NXE + 1; ++m) {
            const size_t i = IX(m, n);
            if (!(g->lu[i] > 0.5f)) continue;
            for (int k = 0; k < nz; ++k) {
                const size_t o = (size_t)k * sz + i;
                den[o] = eos_brydon(tt[o], ss[o], g->free_fall_acc * g->ref_den * hhq[i] * z[k]);
                den_pot[o] = eos_brydon(tt[o], ss[o], 0.0);
            }
        }
}

void bto_model_step(const bto_grid *g, bto_depths *d, double tau, int nstep,
                    int ksw4, double cb_nl, double ebottom, double *ubrtr_e,
                    double *ubrtrp_e, double *vbrtr_e, double *vbrtrp_e,
                    double *ssh_e, double *sshp_e, double *ubrtr_i,
                    double *vbrtr_i, double *ssh4gradx, double *ssh4grady,
                    const double *taux, const double *tauy,
                    const double *wflux, const double *mu, const double *mu4,
                    const double *rdis, double *work)
{
    const size_t sz = (size_t)g->nx * g->ny;
    double *surf_x = work, *surf_y = work + sz, *bot_x = work + 2 * sz;
    double *bot_y = work + 3 * sz, *bf = work + 4 * sz, *rhsx = work + 5 * sz;
    double *rhsy = work + 6 * sz, *vort = work + 7 * sz, *str_t = work + 8 * sz;
    double *str_s = work + 9 * sz, *fx = work + 10 * sz, *fy = work + 11 * sz;
    double *radv = work + 12 * sz, *rady = work + 13 * sz, *rdfx = work + 14 * sz;
    double *rdfy = work + 15 * sz;
    bto_surface_stress(g, taux, tauy, surf_x, surf_y);
    /* the drag sees the external-mode velocities and depths of this step */
    bto_bottom_stress(g, cb_nl, ebottom, ubrtr_e, vbrtr_e, ubrtrp_e, vbrtrp_e,
                      d->h[3], d->h[6], d->h[0], bf, bot_x, bot_y);
    bto_rhs_assembly(g, surf_x, surf_y, bot_x, bot_y, rhsx, rhsy);
    bto_barotropic_dynamics(g, d, tau, nstep, ksw4, ubrtr_e, ubrtrp_e, vbrtr_e,
                            vbrtrp_e, ssh_e, sshp_e, ubrtr_i, vbrtr_i,
                            ssh4gradx, ssh4grady, rhsx, rhsy, wflux, mu, mu4,
                            vort, str_t, str_s, fx, fy, rdis, radv, rady, rdfx,
                            rdfy);
}

double bto_volume(const bto_grid *g, const double *ssh)
{
    double total = 0.0;
    for (int n = NYS; n <= NYE; ++n) {
        double row = 0.0;
        for (int m = NXS; m <= NXE; ++m)
            if (g->lu[IX(m, n)] > 0.5f)
                row += ssh[IX(m, n)] * g->dx[IX(m, n)] * g->dy[IX(m, n)];
        total += row;
    }
    return total;
}

int bto_num_threads(void)
{
#ifdef _OPENMP
    return omp_get_max_threads();
#else
    return 1;
#endif
}
