// Fused barotropic step: one kernel launch advances the 2-D model one
// outer step (Control/dyncall.f90:52-164 with nstep = 1, ksw4 = 0):
// wind stress and quadratic drag (Function/fluxes.f90:29-37, 226-272),
// continuity, depth update, stress tensor, metric vorticity, flux-form
// advection, harmonic viscosity, momentum with Coriolis, pressure
// gradient and Rayleigh friction, and the Robert-Asselin shift
// (Function/vel_ssh.f90:484-648, depth.f90, mixing.f90:115-168).
//
// Work decomposition.  One 64-lane wave owns a strip of 60 output
// columns (lane L holds column m = 60*strip + L - 1, lanes 2..61 write)
// and marches down `chunk` output rows.  Everything a row needs from the
// rows above and below is kept in registers (a rolling window of input
// rows and of the T- and H-point intermediates); x-neighbours come from
// the adjacent lanes by wave shuffles.  No LDS tiles and no barriers:
// each input row is read once per strip and each intermediate is
// computed once per column.  Work items are dealt to the 8 XCDs in
// contiguous bands so that vertically adjacent chunks (which share three
// input rows) run on the same L2.
//
// Arithmetic modes (template parameter EXACT / the stencil type R):
//   FP64_EXACT  the reference's expressions, operation by operation, with
//               the uniform metric substituted; compiled without FMA
//               contraction it reproduces bt_literal.hip bit for bit;
//   FP64        the same algebra simplified for uniform metrics
//               (reciprocals, factored metric products, FMA);
//   FP32        FP64's algebra in FP32 for the level-n fields and all
//               stencils; the filtered (n-1) level, the continuity
//               accumulation and the Robert-Asselin filter stay FP64.
#include <cmath>
#include <cstring>
#include <vector>

#include "bt_fused.h"
#include "bt_fused_body.h"
#include "bt_literal.h"

namespace ibt {

template <typename R, bool EXACT, bool FORM>
__global__ void __launch_bounds__(WAVE) k_fused_step(FusedParams p)
{
    // XCD-aware work mapping: workgroup w runs on XCD w % 8; XCD k takes
    // the contiguous band of work items [k*per_xcd, (k+1)*per_xcd).
    const int w = blockIdx.x;
    const int t = (w % NXCD) * p.per_xcd + (w / NXCD);
    if (t >= p.total) return;
    const int chunk_id = t / p.strips, strip = t - chunk_id * p.strips;
    fused_chunk<R, EXACT, FORM>(p, chunk_id, strip);
}

// Substeps p.sub and p.sub + 1 in one launch (58-column strips).
template <typename R, bool EXACT, bool AFORM>
__global__ void __launch_bounds__(WAVE) k_fused_step2(FusedParams p)
{
    const int w = blockIdx.x;
    const int t = (w % NXCD) * p.per_xcd + (w / NXCD);
    if (t >= p.total) return;
    const int chunk_id = t / p.strips, strip = t - chunk_id * p.strips;
    fused_chunk2<R, EXACT, AFORM>(p, chunk_id, strip);
}

// K substeps per launch, one wave per substep (fused_pipe): workgroups of
// 64 K threads, one work item each.
template <typename R, int K>
__global__ void __launch_bounds__(WAVE * K) k_fused_pipe(FusedParams p, int form_mask, int save_forcing)
{
    __shared__ PipeSlot<R> ring[K - 1][PIPE_SLOTS];
    const int b = blockIdx.x;
    const int t = (b % NXCD) * p.per_xcd + (b / NXCD);
    if (t >= p.total) return;  // the whole workgroup
    const int chunk_id = t / p.strips, strip = t - chunk_id * p.strips;
    fused_pipe<R, K>(p, chunk_id, strip, form_mask, save_forcing != 0, ring);
}

// -------------------------------------------------------------- helpers
__global__ void k_pack_masks(int n, const float *lu, const float *luu, const float *luh,
                             const float *lcu, const float *lcv, const float *llu, const float *llv,
                             uint8_t *mb)
{
    size_t i = (size_t)blockIdx.x * blockDim.x + threadIdx.x;
    if (i >= (size_t)n) return;
    uint8_t b = 0;
    if (lu[i] > 0.5f) b |= MB_LU;
    if (lcu[i] > 0.5f) b |= MB_LCU;
    if (lcv[i] > 0.5f) b |= MB_LCV;
    if (luu[i] > 0.5f) b |= MB_LUU;
    if (luh[i] > 0.5f) b |= MB_LUH;
    if (llu[i] > 0.5f) b |= MB_LLU;
    if (llv[i] > 0.5f) b |= MB_LLV;
    mb[i] = b;
}

__global__ void k_to_f32(size_t n, const double *a, float *b)
{
    size_t i = (size_t)blockIdx.x * blockDim.x + threadIdx.x;
    if (i < n) b[i] = (float)a[i];
}
__global__ void k_to_f64(size_t n, const float *a, double *b)
{
    size_t i = (size_t)blockIdx.x * blockDim.x + threadIdx.x;
    if (i < n) b[i] = (double)a[i];
}
__global__ void k_copy_d(size_t n, const double *a, double *b)
{
    size_t i = (size_t)blockIdx.x * blockDim.x + threadIdx.x;
    if (i < n) b[i] = a[i];
}

// Volume anomaly: one block per row, fixed-shape tree in LDS, then the
// row sums are added in row order by one thread -> deterministic.
template <class R>
__global__ void k_row_volume(int nx, int ny, const R *ssh, const float *lu, const double *dx,
                             const double *dy, double *rows)
{
    __shared__ double part[256];
    const int n = blockIdx.x + 3;  // interior rows 3..ny-2
    double acc = 0.0;
    for (int m = threadIdx.x + 3; m <= nx - 2; m += blockDim.x) {
        size_t i = (size_t)(n - 1) * nx + (m - 1);
        if (lu[i] > 0.5f) acc += (double)ssh[i] * dx[i] * dy[i];
    }
    part[threadIdx.x] = acc;
    __syncthreads();
    for (int s = blockDim.x / 2; s > 0; s >>= 1) {
        if ((int)threadIdx.x < s) part[threadIdx.x] += part[threadIdx.x + s];
        __syncthreads();
    }
    if (threadIdx.x == 0) rows[blockIdx.x] = part[0];
}
__global__ void k_sum_rows(int nrows, const double *rows, double *out)
{
    if (threadIdx.x == 0 && blockIdx.x == 0) {
        double t = 0.0;
        for (int r = 0; r < nrows; ++r) t += rows[r];
        *out = t;
    }
}

// Range check of Control/dyncall.f90:452-516 on the 2-D state.
template <class R>
__global__ void k_range(int nx, int ny, const float *lu, const float *lcu, const float *lcv,
                        const R *ssh, const R *u, const R *v, unsigned long long *bad)
{
    size_t i = (size_t)blockIdx.x * blockDim.x + threadIdx.x;
    if (i >= (size_t)nx * ny) return;
    int m = (int)(i % nx) + 1, n = (int)(i / nx) + 1;
    if (m < 3 || m > nx - 2 || n < 3 || n > ny - 2) return;
    bool b = false;
    if (lu[i] > 0.5f) {
        double s = (double)ssh[i];
        b |= !(fabs(s) < 10.0);
    }
    if (lcu[i] > 0.5f) b |= !(fabs((double)u[i]) < 100.0);
    if (lcv[i] > 0.5f) b |= !(fabs((double)v[i]) < 100.0);
    if (b) atomicAdd(bad, 1ull);
}

// ------------------------------------------------------------- host side
static int hipfail(hipError_t e, const char *what)
{
    set_error(std::string(what) + ": " + hipGetErrorString(e));
    return INMSOM_BT_EDEVICE;
}
#define HC(expr)                                         \
    do {                                                 \
        hipError_t e_ = (expr);                          \
        if (e_ != hipSuccess) return hipfail(e_, #expr); \
    } while (0)

static inline unsigned nblk(size_t n) { return (unsigned)((n + 255) / 256); }

int fused_set_substeps(inmsom_bt_ctx *c, int k)
{
    if (k < 0 || k > 2) {
        set_error("substeps per launch must be 0 (automatic), 1 or 2");
        return INMSOM_BT_EINVAL;
    }
    c->subs_per_launch = k;
    return INMSOM_BT_OK;
}

int fused_set_pipeline(inmsom_bt_ctx *c, int k)
{
    if (k < 0 || k > 4) {
        set_error("pipeline stages must be 0 (automatic), 1 (off) or 2..4");
        return INMSOM_BT_EINVAL;
    }
    c->pipe_k = k;
    return INMSOM_BT_OK;
}

// Stages of the pipeline for the next launches (0 or 1: none): as set, or
// automatically 2 in FP64 and FP32 from TWO_SUBSTEP_CELLS cells on when the
// substeps per launch are automatic too.  Measured on MI355X at 2048 x 1536
// (us per substep, HIP events over 10^3 outer steps; two-substep march in
// one wave / 2-wave pipeline / 4-wave pipeline): FP64 73.2 / 60.7 / 68.4,
// FP32 51.3 / 50.0 / 52.3.
int pipe_stages(const inmsom_bt_ctx *c, int mode)
{
    if (mode == INMSOM_BT_FP64_EXACT) return 0;
    if (c->pipe_k > 0) return c->pipe_k >= 2 ? c->pipe_k : 0;
    return c->subs_per_launch == 0 && c->n >= TWO_SUBSTEP_CELLS ? 2 : 0;
}

int fused_set_shape(inmsom_bt_ctx *c, int chunk)
{
    if (chunk < 0 || chunk > 4096) {
        set_error("rows per wave out of range (0 = default, 1..4096)");
        return INMSOM_BT_EINVAL;
    }
    c->chunk = chunk;
    return INMSOM_BT_OK;
}

// Per-row tables and the packed mask; validates that the fused step
// applies (uniform metrics, forcing uniform along x, mu = nu * lu).
static int fused_prepare(inmsom_bt_ctx *c)
{
    if (c->fused_ready) return INMSOM_BT_OK;
    if (!c->uniform_metrics) {
        set_error("the fused step needs uniform metrics; use inmsom_bt_model_steps");
        return INMSOM_BT_ESTATE;
    }
    if (!c->have_forcing) {
        set_error("inmsom_bt_set_forcing must be called first");
        return INMSOM_BT_ESTATE;
    }
    const int nx = c->nx, ny = c->ny;
    const size_t n = c->n;
    std::vector<double> taux(n), tauy(n), wf(n), mu(n), f(n);
    std::vector<float> lu(n);
    HC(hipMemcpy(taux.data(), c->taux, n * 8, hipMemcpyDeviceToHost));
    HC(hipMemcpy(tauy.data(), c->tauy, n * 8, hipMemcpyDeviceToHost));
    HC(hipMemcpy(wf.data(), c->wflux, n * 8, hipMemcpyDeviceToHost));
    HC(hipMemcpy(mu.data(), c->mu, n * 8, hipMemcpyDeviceToHost));
    HC(hipMemcpy(f.data(), c->rlh_s, n * 8, hipMemcpyDeviceToHost));
    HC(hipMemcpy(lu.data(), c->mask[0], n * 4, hipMemcpyDeviceToHost));
    double nu = -1.0;
    for (size_t i = 0; i < n; ++i) {
        if (lu[i] > 0.5f) {
            if (nu < 0.0) nu = mu[i];
            if (mu[i] != nu) {
                set_error("the fused step needs mu = constant on wet cells and 0 on land");
                return INMSOM_BT_ESTATE;
            }
        } else if (mu[i] != 0.0) {
            set_error("the fused step needs mu = 0 on land");
            return INMSOM_BT_ESTATE;
        }
    }
    c->nu = nu < 0.0 ? 0.0 : nu;
    const double d = c->metric0, rho0 = c->g.rho0;
    std::vector<double> rf(ny), rsx(ny), rsy(ny), rwf(ny);
    for (int r = 0; r < ny; ++r) {
        const double *tx = &taux[(size_t)r * nx], *ty = &tauy[(size_t)r * nx];
        const double *w = &wf[(size_t)r * nx], *fr = &f[(size_t)r * nx];
        for (int m = 1; m < nx; ++m)
            if (tx[m] != tx[0] || ty[m] != ty[0] || w[m] != w[0] || fr[m] != fr[0]) {
                set_error("the fused step needs taux, tauy, wflux and rlh_s uniform along x");
                return INMSOM_BT_ESTATE;
            }
        rf[r] = fr[0];
        rwf[r] = w[0];
        // Function/fluxes.f90:30-31 and :35-36 with the row's uniform values
        rsx[r] = (tx[0] * d * d + tx[0] * d * d) / rho0 / 2.0 / d / d;
        rsy[r] = (ty[0] * d * d + ty[0] * d * d) / rho0 / 2.0 / d / d;
    }
    auto upl = [&](double **dst, const std::vector<double> &src) -> int {
        if (!*dst) HC(hipMalloc(dst, src.size() * 8));
        HC(hipMemcpy(*dst, src.data(), src.size() * 8, hipMemcpyHostToDevice));
        return INMSOM_BT_OK;
    };
    int rc;
    if ((rc = upl(&c->row_f, rf)) || (rc = upl(&c->row_wx, rsx)) || (rc = upl(&c->row_wy, rsy)) ||
        (rc = upl(&c->row_wf, rwf)))
        return rc;
    if (!c->mbits) HC(hipMalloc(&c->mbits, n));
    hipLaunchKernelGGL(k_pack_masks, dim3(nblk(n)), dim3(256), 0, c->stream, (int)n, c->mask[0],
                       c->mask[1], c->mask[2], c->mask[3], c->mask[4], c->mask[5], c->mask[6],
                       c->mbits);
    HC(hipGetLastError());
    HC(hipStreamSynchronize(c->stream));
    c->fused_ready = true;
    return INMSOM_BT_OK;
}

// Fused state buffers: level n (u, v, ssh) in the stencil type and the
// filtered level (up, vp, sshp) in FP64, twice (ping-pong), plus the
// static depth and Rayleigh coefficient in the stencil type.
struct FusedAlloc {
    void *lvl[2][3] = {};
    double *flt[2][3] = {};
    void *H = nullptr, *rd = nullptr;
    void *frx = nullptr, *fry = nullptr;     // frozen forcing (stencil type)
    // a second set: a pipeline launch that reads the current call's forcing
    // and forms the next call's writes the other set (no work item may
    // overwrite rows another still reads); frc_cur: the set holding it
    void *frx2 = nullptr, *fry2 = nullptr;
    int frc_cur = 0;
    double *hup_c = nullptr, *hvp_c = nullptr;  // carried filtered depths
    int mode = -1;
};
static FusedAlloc *falloc(inmsom_bt_ctx *c) { return c->fa; }

void fused_release(inmsom_bt_ctx *c)
{
    FusedAlloc *a = falloc(c);
    if (a) {
        for (int b = 0; b < 2; ++b)
            for (int k = 0; k < 3; ++k) {
                if (a->lvl[b][k]) hipFree(a->lvl[b][k]);
                if (a->flt[b][k]) hipFree(a->flt[b][k]);
            }
        for (void *q : {a->H, a->rd, a->frx, a->fry, a->frx2, a->fry2, (void *)a->hup_c, (void *)a->hvp_c})
            if (q) hipFree(q);
        delete a;
        c->fa = nullptr;
    }
    for (double **p : {&c->row_f, &c->row_wx, &c->row_wy, &c->row_wf})
        if (*p) {
            hipFree(*p);
            *p = nullptr;
        }
    if (c->mbits) hipFree(c->mbits);
    c->mbits = nullptr;
    c->fused_ready = false;
    c->state_on_fused = false;
}

// Moves the canonical FP64 state into the fused buffers of `mode`.
static int fused_load(inmsom_bt_ctx *c, int mode)
{
    FusedAlloc *a = falloc(c);
    const size_t n = c->n;
    const size_t es = (mode == INMSOM_BT_FP32) ? 4 : 8;
    if (a && a->mode != mode) {
        // different element size: reallocate
        fused_release(c);
        c->fused_ready = false;
        int rc = fused_prepare(c);
        if (rc) return rc;
        a = nullptr;
    }
    if (!a) {
        a = new FusedAlloc();
        c->fa = a;
        a->mode = mode;
        for (int b = 0; b < 2; ++b)
            for (int k = 0; k < 3; ++k) {
                HC(hipMalloc(&a->lvl[b][k], n * es));
                HC(hipMalloc(&a->flt[b][k], n * 8));
            }
        HC(hipMalloc(&a->H, n * es));
        HC(hipMalloc(&a->rd, n * es));
        HC(hipMalloc(&a->frx, n * es));
        HC(hipMalloc(&a->fry, n * es));
        HC(hipMalloc(&a->frx2, n * es));
        HC(hipMalloc(&a->fry2, n * es));
        HC(hipMalloc(&a->hup_c, n * 8));
        HC(hipMalloc(&a->hvp_c, n * 8));
        if (es == 4) {
            hipLaunchKernelGGL(k_to_f32, dim3(nblk(n)), dim3(256), 0, c->stream, n, c->hr, (float *)a->H);
            hipLaunchKernelGGL(k_to_f32, dim3(nblk(n)), dim3(256), 0, c->stream, n, c->rdis, (float *)a->rd);
        } else {
#pragma clang fp contract(fast)
            hipLaunchKernelGGL(k_copy_d, dim3(nblk(n)), dim3(256), 0, c->stream, n, c->hr, (double *)a->H);
            hipLaunchKernelGGL(k_copy_d, dim3(nblk(n)), dim3(256), 0, c->stream, n, c->rdis, (double *)a->rd);
        }
    } else {
#pragma clang fp contract(fast)
        // forcing may have changed: refresh the static copies
        if (es == 4)
            hipLaunchKernelGGL(k_to_f32, dim3(nblk(n)), dim3(256), 0, c->stream, n, c->rdis, (float *)a->rd);
        else
            hipLaunchKernelGGL(k_copy_d, dim3(nblk(n)), dim3(256), 0, c->stream, n, c->rdis, (double *)a->rd);
    }
    // canonical order u, up, v, vp, ssh, sshp -> level (u, v, ssh), filtered (up, vp, sshp)
    const int lv[3] = {0, 2, 4}, fl[3] = {1, 3, 5};
    for (int k = 0; k < 3; ++k) {
        if (es == 4)
            hipLaunchKernelGGL(k_to_f32, dim3(nblk(n)), dim3(256), 0, c->stream, n, c->st[lv[k]],
                               (float *)a->lvl[0][k]);
        else
            hipLaunchKernelGGL(k_copy_d, dim3(nblk(n)), dim3(256), 0, c->stream, n, c->st[lv[k]],
                               (double *)a->lvl[0][k]);
        hipLaunchKernelGGL(k_copy_d, dim3(nblk(n)), dim3(256), 0, c->stream, n, c->st[fl[k]], a->flt[0][k]);
    }
    HC(hipGetLastError());
    c->fz_cur = 0;
    c->fused_mode = mode;
    c->state_on_fused = true;
    return INMSOM_BT_OK;
}

int fused_flush(inmsom_bt_ctx *c)
{
    if (!c->state_on_fused) return INMSOM_BT_OK;
    FusedAlloc *a = falloc(c);
    const size_t n = c->n;
    const int b = c->fz_cur;
    const int lv[3] = {0, 2, 4}, fl[3] = {1, 3, 5};
    for (int k = 0; k < 3; ++k) {
        if (a->mode == INMSOM_BT_FP32)
            hipLaunchKernelGGL(k_to_f64, dim3(nblk(n)), dim3(256), 0, c->stream, n,
                               (const float *)a->lvl[b][k], c->st[lv[k]]);
        else
            hipLaunchKernelGGL(k_copy_d, dim3(nblk(n)), dim3(256), 0, c->stream, n,
                               (const double *)a->lvl[b][k], c->st[lv[k]]);
        hipLaunchKernelGGL(k_copy_d, dim3(nblk(n)), dim3(256), 0, c->stream, n, a->flt[b][k], c->st[fl[k]]);
    }
    // the reference's callers see depths re-initialised from the new state
    lit_hh_init(c, c->st[4], c->st[5]);
    HC(hipGetLastError());
    HC(hipStreamSynchronize(c->stream));
    c->state_on_fused = false;
    c->fused_mode = -1;
    return INMSOM_BT_OK;
}

// Parameters of a launch that reads buffer set b and writes 1 - b: substep
// `sub`, or substeps sub and sub + 1 when `two`.
// Substeps per launch: as set, or automatically two in FP64 and FP32 on
// grids large enough to be bandwidth-bound and one otherwise.  Measured on
// MI355X (us per substep, one / two per launch): 2048 x 1536 FP64 91.3 /
// 81.5, FP32 66.1 / 53.9, FP64_EXACT 153.8 / 183.1 (its IEEE divides make
// it compute-bound); 512 x 384 (latency-bound, rows per wave 2) FP32
// 8.7 / 9.5 per substep.
int subs_per_launch(const inmsom_bt_ctx *c, int mode)
{
    if (c->subs_per_launch > 0) return c->subs_per_launch;
    return mode != INMSOM_BT_FP64_EXACT && c->n >= TWO_SUBSTEP_CELLS ? 2 : 1;
}

// the kernel for (mode, two substeps, forming substep)
template <bool FORM>
static const void *fused_kernel(int mode, bool two)
{
    if (two) {
        if (mode == INMSOM_BT_FP64_EXACT) return (const void *)&k_fused_step2<double, true, FORM>;
        if (mode == INMSOM_BT_FP64) return (const void *)&k_fused_step2<double, false, FORM>;
        return (const void *)&k_fused_step2<float, false, FORM>;
    }
    if (mode == INMSOM_BT_FP64_EXACT) return (const void *)&k_fused_step<double, true, FORM>;
    if (mode == INMSOM_BT_FP64) return (const void *)&k_fused_step<double, false, FORM>;
    return (const void *)&k_fused_step<float, false, FORM>;
}

static const void *pipe_kernel(int mode, int K)
{
    if (mode == INMSOM_BT_FP64) {
        if (K == 2) return (const void *)&k_fused_pipe<double, 2>;
        if (K == 3) return (const void *)&k_fused_pipe<double, 3>;
        return (const void *)&k_fused_pipe<double, 4>;
    }
    if (K == 2) return (const void *)&k_fused_pipe<float, 2>;
    if (K == 3) return (const void *)&k_fused_pipe<float, 3>;
    return (const void *)&k_fused_pipe<float, 4>;
}

// Workgroups of the K-wave pipeline kernel the device holds at once.
static long long pipe_slots(inmsom_bt_ctx *c, int mode, int K)
{
    int cus = 0, per_cu = 0;
    if (hipDeviceGetAttribute(&cus, hipDeviceAttributeMultiprocessorCount, c->device) != hipSuccess || cus < 1)
        cus = 256;
    if (hipOccupancyMaxActiveBlocksPerMultiprocessor(&per_cu, pipe_kernel(mode, K), WAVE * K, 0) != hipSuccess ||
        per_cu < 1)
        per_cu = 1;
    return (long long)cus * per_cu;
}

// Waves of the fused kernel for (mode, two) that the device holds at once:
// CUs x waves per CU from the occupancy calculator, once per context.
static long long wave_slots(inmsom_bt_ctx *c, int mode, bool two)
{
    const int idx = (mode == INMSOM_BT_FP64_EXACT ? 0 : mode == INMSOM_BT_FP64 ? 1 : 2) + (two ? 3 : 0);
    if (c->slots[idx] > 0) return c->slots[idx];
    int cus = 0, per_cu = 0;
    if (hipDeviceGetAttribute(&cus, hipDeviceAttributeMultiprocessorCount, c->device) != hipSuccess || cus < 1)
        cus = 256;
    if (hipOccupancyMaxActiveBlocksPerMultiprocessor(&per_cu, fused_kernel<true>(mode, two), WAVE, 0) !=
            hipSuccess ||
        per_cu < 1)
        per_cu = 4;
    c->slots[idx] = (long long)cus * per_cu;
    return c->slots[idx];
}

static FusedParams make_params(inmsom_bt_ctx *c, FusedAlloc *a, int b, double tau, int nstep,
                               int sub, int mode, bool two)
{
    FusedParams p{};
    p.nx = c->nx;
    p.ny = c->ny;
    const int sw = two ? SW2 : SW;
    p.strips = (c->nx + 2 + sw - 1) / sw;  // columns 1..nx need strips covering m <= nx
    p.chunk = c->chunk > 0 ? c->chunk : fit_chunk(c->ny, p.strips, wave_slots(c, mode, two), two ? 7 : 3, 2);
    p.nchunks = (c->ny + p.chunk - 1) / p.chunk;
    p.total = p.strips * p.nchunks;
    p.per_xcd = (p.total + NXCD - 1) / NXCD;
    p.mb = c->mbits;
    p.H = a->H;
    p.rd = a->rd;
    p.u = a->lvl[b][0];
    p.v = a->lvl[b][1];
    p.s = a->lvl[b][2];
    p.up = a->flt[b][0];
    p.vp = a->flt[b][1];
    p.sp = a->flt[b][2];
    p.uo = a->lvl[1 - b][0];
    p.vo = a->lvl[1 - b][1];
    p.so = a->lvl[1 - b][2];
    p.upo = a->flt[1 - b][0];
    p.vpo = a->flt[1 - b][1];
    p.spo = a->flt[1 - b][2];
    p.row_f = c->row_f;
    p.row_sx = c->row_wx;
    p.row_sy = c->row_wy;
    p.row_wf = c->row_wf;
    p.g = c->g.g;
    p.rho0 = c->g.rho0;
    p.ts = c->g.ts;
    p.tau = tau / (double)nstep;  // tau_inner, vel_ssh.f90:465
    p.sub = sub;
    p.nsub = nstep;
    p.dn = (double)nstep;
    p.frx = a->frc_cur ? a->frx2 : a->frx;
    p.fry = a->frc_cur ? a->fry2 : a->fry;
    p.frx_in = p.frx;
    p.fry_in = p.fry;
    p.hup_c = a->hup_c;
    p.hvp_c = a->hvp_c;
    p.cb = c->cb_nl;
    p.eb = c->ebottom;
    p.nu = c->nu;
    p.ffs = (double)c->g.ffs;
    p.d = c->metric0;
    return p;
}

// the launch shape of the K-wave pipeline: strips of pipe_width(K) columns
static void pipe_shape(inmsom_bt_ctx *c, int mode, int K, FusedParams &p)
{
    const int sw = pipe_width(K);
    p.strips = (c->nx + sw - 1) / sw;
    long long &slots = c->pipe_slots[mode == INMSOM_BT_FP64 ? 0 : 1][K];
    if (slots <= 0) slots = pipe_slots(c, mode, K);
    p.chunk = c->chunk > 0 ? c->chunk : fit_chunk(c->ny, p.strips, slots, pipe_extra(K), 2);
    p.nchunks = (c->ny + p.chunk - 1) / p.chunk;
    p.total = p.strips * p.nchunks;
    p.per_xcd = (p.total + NXCD - 1) / NXCD;
}

static void launch_pipe(inmsom_bt_ctx *c, const FusedParams &p, int mode, int K, int form_mask, int save_forcing)
{
    dim3 grid((unsigned)(p.per_xcd * NXCD)), blk(WAVE * K);
    void *args[] = {const_cast<FusedParams *>(&p), &form_mask, &save_forcing};
    (void)hipLaunchKernel(pipe_kernel(mode, K), grid, blk, args, 0, c->stream);
}

static void launch(inmsom_bt_ctx *c, const FusedParams &p, int mode, bool two)
{
    dim3 grid((unsigned)(p.per_xcd * NXCD)), blk(WAVE);
    const void *k = p.sub == 1 ? fused_kernel<true>(mode, two) : fused_kernel<false>(mode, two);
    void *args[] = {const_cast<FusedParams *>(&p)};
    hipLaunchKernel(k, grid, blk, args, 0, c->stream);
}

int fused_run(inmsom_bt_ctx *c, double tau, int nstep, int nsteps, int mode, double *ms)
{
    if (nstep < 1 || nstep > 1000) {
        set_error("nstep must be in 1..1000");
        return INMSOM_BT_EINVAL;
    }
    if (mode != INMSOM_BT_FP64_EXACT && mode != INMSOM_BT_FP64 && mode != INMSOM_BT_FP32) {
        set_error("unknown mode");
        return INMSOM_BT_EINVAL;
    }
    if (!(tau > 0.0) || nsteps < 0) {
        set_error("bad tau / nsteps");
        return INMSOM_BT_EINVAL;
    }
    if (c->g.ffs != 1 || c->g.px || c->g.py) {
        set_error("the fused step covers full_free_surface = 1 without periodicity");
        return INMSOM_BT_ESTATE;
    }
    int rc = fused_prepare(c);
    if (rc) return rc;
    if (c->state_on_fused && c->fused_mode != mode) {
        rc = fused_flush(c);
        if (rc) return rc;
    }
    if (!c->state_on_fused) {
        rc = fused_load(c, mode);
        if (rc) return rc;
    }
    FusedAlloc *a = falloc(c);
    hipEvent_t e0 = nullptr, e1 = nullptr;
    if (ms) {
        HC(hipEventCreate(&e0));
        HC(hipEventCreate(&e1));
        HC(hipEventRecord(e0, c->stream));
    }
    // substeps of all outer steps in order; g = index (from 0) of the next
    // launch's first substep.  The K-wave pipeline may span outer steps;
    // one- and two-substep launches stay within one.
    const int K = pipe_stages(c, mode);
    const long long total = (long long)nsteps * nstep;
    for (long long g = 0; g < total;) {
        const int sub = (int)(g % nstep) + 1;
        if (K >= 2 && g + K <= total) {
            FusedParams p = make_params(c, a, c->fz_cur, tau, nstep, sub, mode, false);
            pipe_shape(c, mode, K, p);
            int form_mask = 0;
            for (int w = 0; w < K; ++w)
                if ((g + w) % nstep == 0) form_mask |= 1 << w;
            // the last stage's call continues after this launch: its forcing is
            // saved, into the other set when a stage of this launch formed it
            // (the current set may still be read by other work items' first
            // stage); a launch that forms none leaves the current set as it is
            int save = (int)((g + K - 1) % nstep) != nstep - 1;
            if (save && form_mask) {
                p.frx = a->frc_cur ? a->frx : a->frx2;
                p.fry = a->frc_cur ? a->fry : a->fry2;
                a->frc_cur = 1 - a->frc_cur;
            } else {
                save = 0;
            }
            launch_pipe(c, p, mode, K, form_mask, save);
            g += K;
        } else {
            const bool two = subs_per_launch(c, mode) == 2 && sub + 1 <= nstep;
            FusedParams p = make_params(c, a, c->fz_cur, tau, nstep, sub, mode, two);
            launch(c, p, mode, two);
            g += two ? 2 : 1;
        }
        c->fz_cur = 1 - c->fz_cur;
    }
    HC(hipGetLastError());
    if (ms) {
        HC(hipEventRecord(e1, c->stream));
        HC(hipEventSynchronize(e1));
        float t = 0.f;
        HC(hipEventElapsedTime(&t, e0, e1));
        *ms = t;
        hipEventDestroy(e0);
        hipEventDestroy(e1);
    } else {
        HC(hipStreamSynchronize(c->stream));
    }
    return INMSOM_BT_OK;
}

int fused_range_check(inmsom_bt_ctx *c, long long *bad)
{
    const size_t n = c->n;
    HC(hipMemsetAsync(c->cnt, 0, sizeof(long long), c->stream));
    unsigned long long *cnt = reinterpret_cast<unsigned long long *>(c->cnt);
    if (c->state_on_fused) {
        FusedAlloc *a = falloc(c);
        const int b = c->fz_cur;
        if (a->mode == INMSOM_BT_FP32)
            hipLaunchKernelGGL(k_range<float>, dim3(nblk(n)), dim3(256), 0, c->stream, c->nx, c->ny,
                               c->mask[0], c->mask[3], c->mask[4], (const float *)a->lvl[b][2],
                               (const float *)a->lvl[b][0], (const float *)a->lvl[b][1], cnt);
        else
            hipLaunchKernelGGL(k_range<double>, dim3(nblk(n)), dim3(256), 0, c->stream, c->nx, c->ny,
                               c->mask[0], c->mask[3], c->mask[4], (const double *)a->lvl[b][2],
                               (const double *)a->lvl[b][0], (const double *)a->lvl[b][1], cnt);
    } else {
#pragma clang fp contract(fast)
        hipLaunchKernelGGL(k_range<double>, dim3(nblk(n)), dim3(256), 0, c->stream, c->nx, c->ny,
                           c->mask[0], c->mask[3], c->mask[4], c->st[4], c->st[0], c->st[2], cnt);
    }
    HC(hipGetLastError());
    long long h = 0;
    HC(hipMemcpyAsync(&h, c->cnt, sizeof(long long), hipMemcpyDeviceToHost, c->stream));
    HC(hipStreamSynchronize(c->stream));
    *bad = h;
    return INMSOM_BT_OK;
}

// Block average of the state onto a grid `f` times coarser (interior
// blocks of f x f cells; two halo cells per side on both grids):
//   ssh: mean over the block's wet cells,
//   u:   mean of the f fine U faces on the block's east edge,
//   v:   mean of the f fine V faces on the block's north edge.
template <class T>
__global__ void k_coarsen(int nx, int ny, int f, int nxc, int nyc, const T *__restrict__ u,
                          const T *__restrict__ v, const T *__restrict__ s, const float *__restrict__ lu,
                          double *__restrict__ uc, double *__restrict__ vc, double *__restrict__ sc)
{
    const int i = blockIdx.x * blockDim.x + threadIdx.x;
    if (i >= nxc * nyc) return;
    const int mc = i % nxc, nc = i / nxc;
    double us = 0.0, vs = 0.0, ss = 0.0, ws = 0.0;
    if (mc >= 2 && mc < nxc - 2 && nc >= 2 && nc < nyc - 2) {
        const int m0 = 2 + (mc - 2) * f, n0 = 2 + (nc - 2) * f;
        for (int b = 0; b < f; ++b)
            for (int a = 0; a < f; ++a) {
                const size_t k = (size_t)(n0 + b) * nx + (m0 + a);
                ss += (double)s[k] * lu[k];
                ws += lu[k];
            }
        for (int b = 0; b < f; ++b) us += (double)u[(size_t)(n0 + b) * nx + (m0 + f - 1)];
        for (int a = 0; a < f; ++a) vs += (double)v[(size_t)(n0 + f - 1) * nx + (m0 + a)];
        us /= f;
        vs /= f;
        ss = ws > 0.0 ? ss / ws : 0.0;
    }
    uc[i] = us;
    vc[i] = vs;
    sc[i] = ss;
}

int fused_coarsen(inmsom_bt_ctx *c, int f, double *u, double *v, double *ssh)
{
    if (f < 1 || (c->nx - 4) % f || (c->ny - 4) % f) {
        set_error("the interior must divide into factor x factor blocks");
        return INMSOM_BT_EINVAL;
    }
    const int nxc = (c->nx - 4) / f + 4, nyc = (c->ny - 4) / f + 4;
    const size_t nc = (size_t)nxc * nyc;
    double *out = c->w[6];  // working arrays of barotropic_dynamics: free between calls
    double *uo = out, *vo = c->w[7], *so = c->w[8];
    if (c->state_on_fused && falloc(c)->mode == INMSOM_BT_FP32) {
        FusedAlloc *a = falloc(c);
        hipLaunchKernelGGL(k_coarsen<float>, dim3(nblk(nc)), dim3(256), 0, c->stream, c->nx, c->ny, f, nxc, nyc,
                           (const float *)a->lvl[c->fz_cur][0], (const float *)a->lvl[c->fz_cur][1],
                           (const float *)a->lvl[c->fz_cur][2], c->mask[0], uo, vo, so);
    } else {
        const double *su = c->st[0], *sv = c->st[2], *ss = c->st[4];
        if (c->state_on_fused) {
            FusedAlloc *a = falloc(c);
            su = (const double *)a->lvl[c->fz_cur][0];
            sv = (const double *)a->lvl[c->fz_cur][1];
            ss = (const double *)a->lvl[c->fz_cur][2];
        }
        hipLaunchKernelGGL(k_coarsen<double>, dim3(nblk(nc)), dim3(256), 0, c->stream, c->nx, c->ny, f, nxc, nyc,
                           su, sv, ss, c->mask[0], uo, vo, so);
    }
    HC(hipGetLastError());
    HC(hipMemcpyAsync(u, uo, nc * 8, hipMemcpyDeviceToHost, c->stream));
    HC(hipMemcpyAsync(v, vo, nc * 8, hipMemcpyDeviceToHost, c->stream));
    HC(hipMemcpyAsync(ssh, so, nc * 8, hipMemcpyDeviceToHost, c->stream));
    HC(hipStreamSynchronize(c->stream));
    return INMSOM_BT_OK;
}

int fused_volume(inmsom_bt_ctx *c, double *vol)
{
    const int rows = c->ny - 4;
    if (rows <= 0 || rows > DEP_TABLE_OFFSET * 1024) {
        set_error("grid too small");
        return INMSOM_BT_EINVAL;
    }
    double *rowbuf = nullptr;
    HC(hipMallocAsync((void **)&rowbuf, (size_t)rows * 8, c->stream));
    if (c->state_on_fused && falloc(c)->mode == INMSOM_BT_FP32)
        hipLaunchKernelGGL(k_row_volume<float>, dim3(rows), dim3(256), 0, c->stream, c->nx, c->ny,
                           (const float *)falloc(c)->lvl[c->fz_cur][2], c->mask[0], c->metric[0],
                           c->metric[1], rowbuf);
    else {
        const double *s = c->state_on_fused ? (const double *)falloc(c)->lvl[c->fz_cur][2] : c->st[4];
        hipLaunchKernelGGL(k_row_volume<double>, dim3(rows), dim3(256), 0, c->stream, c->nx, c->ny, s,
                           c->mask[0], c->metric[0], c->metric[1], rowbuf);
    }
    hipLaunchKernelGGL(k_sum_rows, dim3(1), dim3(64), 0, c->stream, rows, rowbuf, c->scal);
    HC(hipGetLastError());
    double h = 0.0;
    HC(hipMemcpyAsync(&h, c->scal, 8, hipMemcpyDeviceToHost, c->stream));
    HC(hipFreeAsync(rowbuf, c->stream));
    HC(hipStreamSynchronize(c->stream));
    *vol = h;
    return INMSOM_BT_OK;
}

}  // namespace ibt
