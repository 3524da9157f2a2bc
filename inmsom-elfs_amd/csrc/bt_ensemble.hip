// Batched ensembles: many independent basins of uneven size advanced by
// one persistent launch per substep, replayed from a HIP graph.
//
// Every member keeps its own fields (the fused step's layout, see
// bt_fused.hip) and a descriptor (FusedParams) in a device table, one
// table per ping-pong phase.  The work list enumerates (member, strip,
// chunk) items of all members; a fixed grid of waves walks it with a
// static stride, so every wave ends after its last item.  Items are dealt
// to the 8 XCDs in contiguous bands (a member's neighbouring chunks meet in
// one L2).  One outer step = nstep launches (substeps); its launches are
// captured once into a graph and the graph is replayed per outer step.
#include <cmath>
#include <cstring>
#include <string>
#include <vector>

#include "bt_cnn.h"
#include "bt_fused_body.h"
#include "bt_internal.h"

namespace ibt {

template <typename R, bool EXACT, bool FORM>
__global__ void __launch_bounds__(WAVE)
    k_fused_batch(const FusedParams *__restrict__ table, const int *__restrict__ item_member,
                  const int *__restrict__ item_local, int total, int per_xcd, int waves_per_xcd,
                  double tau_inner, int sub, int nsub)
{
    const int w = blockIdx.x;
    const int xcd = w % NXCD, j = w / NXCD;
    const int t0 = xcd * per_xcd, t1 = min(t0 + per_xcd, total);
    for (int t = t0 + j; t < t1; t += waves_per_xcd) {
        FusedParams p = table[item_member[t]];
        p.tau = tau_inner;
        p.sub = sub;
        p.nsub = nsub;
        p.dn = (double)nsub;
        const int local = item_local[t];
        const int chunk_id = local / p.strips, strip = local - chunk_id * p.strips;
        fused_chunk<R, EXACT, FORM>(p, chunk_id, strip);
    }
}

// Substeps sub and sub + 1 in one march per item (58-column strips).
template <typename R, bool EXACT, bool AFORM>
__global__ void __launch_bounds__(WAVE)
    k_fused_batch2(const FusedParams *__restrict__ table, const int *__restrict__ item_member,
                   const int *__restrict__ item_local, int total, int per_xcd, int waves_per_xcd,
                   double tau_inner, int sub, int nsub)
{
    const int w = blockIdx.x;
    const int xcd = w % NXCD, j = w / NXCD;
    const int t0 = xcd * per_xcd, t1 = min(t0 + per_xcd, total);
    for (int t = t0 + j; t < t1; t += waves_per_xcd) {
        FusedParams p = table[item_member[t]];
        p.tau = tau_inner;
        p.sub = sub;
        p.nsub = nsub;
        p.dn = (double)nsub;
        const int local = item_local[t];
        const int chunk_id = local / p.strips, strip = local - chunk_id * p.strips;
        fused_chunk2<R, EXACT, AFORM>(p, chunk_id, strip);
    }
}

// Substeps sub and sub + 1 of every item as a two-wave pipeline
// (fused_pipe, 58-column strips as the pair march): the two waves of a
// workgroup walk the work list together, one substep each.
static_assert(SW2 == pipe_width(2), "the pipeline and the pair march share the strip width");
template <typename R>
__global__ void __launch_bounds__(2 * WAVE)
    k_fused_batch_pipe(const FusedParams *__restrict__ table, const int *__restrict__ item_member,
                       const int *__restrict__ item_local, int total, int per_xcd, int groups_per_xcd,
                       double tau_inner, int sub, int nsub)
{
    __shared__ PipeSlot<R> ring[1][PIPE_SLOTS];
    const int w = blockIdx.x;
    const int xcd = w % NXCD, j = w / NXCD;
    const int t0 = xcd * per_xcd, t1 = min(t0 + per_xcd, total);
    // the first pair of a call forms the forcing; the later pairs read it,
    // so the first pair saves it when the call goes on
    const int form = sub == 1 ? 1 : 0;
    const bool save = sub == 1 && sub + 1 < nsub;
    for (int t = t0 + j; t < t1; t += groups_per_xcd) {
        FusedParams p = table[item_member[t]];
        p.tau = tau_inner;
        p.sub = sub;
        p.nsub = nsub;
        p.dn = (double)nsub;
        const int local = item_local[t];
        const int chunk_id = local / p.strips, strip = local - chunk_id * p.strips;
        fused_pipe<R, 2>(p, chunk_id, strip, form, save, ring);
    }
}

template <class R>
__global__ void k_member_range(int nx, int ny, const uint8_t *mb, const R *ssh, const R *u, const R *v,
                               unsigned long long *bad)
{
    size_t i = (size_t)blockIdx.x * blockDim.x + threadIdx.x;
    if (i >= (size_t)nx * ny) return;
    int m = (int)(i % nx) + 1, n = (int)(i / nx) + 1;
    if (m < 3 || m > nx - 2 || n < 3 || n > ny - 2) return;
    bool b = false;
    if (mb[i] & MB_LU) b |= !(fabs((double)ssh[i]) < 10.0);
    if (mb[i] & MB_LCU) b |= !(fabs((double)u[i]) < 100.0);
    if (mb[i] & MB_LCV) b |= !(fabs((double)v[i]) < 100.0);
    if (b) atomicAdd(bad, 1ull);
}

// per-row partial sums of ssh over wet interior cells (fixed order)
template <class R>
__global__ void k_member_rows(int nx, int ny, const uint8_t *mb, const R *ssh, double area,
                              double *rows)
{
    __shared__ double part[256];
    const int n = blockIdx.x + 3;
    double acc = 0.0;
    for (int m = threadIdx.x + 3; m <= nx - 2; m += blockDim.x) {
        size_t i = (size_t)(n - 1) * nx + (m - 1);
        if (mb[i] & MB_LU) acc += (double)ssh[i] * area;
    }
    part[threadIdx.x] = acc;
    __syncthreads();
    for (int s = blockDim.x / 2; s > 0; s >>= 1) {
        if ((int)threadIdx.x < s) part[threadIdx.x] += part[threadIdx.x + s];
        __syncthreads();
    }
    if (threadIdx.x == 0) rows[blockIdx.x] = part[0];
}

}  // namespace ibt

using namespace ibt;

struct EnsMember {
    int nx = 0, ny = 0;
    size_t n = 0;
    double d = 0.0, area = 0.0;
    uint8_t *mb = nullptr;
    void *H = nullptr, *rd = nullptr;
    void *lvl[2][3] = {};
    double *flt[2][3] = {};
    void *frx = nullptr, *fry = nullptr;
    double *hup_c = nullptr, *hvp_c = nullptr;
    double *row_f = nullptr, *row_sx = nullptr, *row_sy = nullptr, *row_wf = nullptr;
    double nu = 0.0;
    int items = 0;
};

struct inmsom_bt_ens {
    int device = 0;
    hipStream_t stream = nullptr;
    int mode = INMSOM_BT_FP32;
    double g = 9.81, rho0 = 1025.0, ts = 0.05, cb = 2.5e-3, eb = 25e-4, ffs = 1.0;
    std::vector<EnsMember> mem;
    FusedParams *table[2] = {};
    int *item_member = nullptr, *item_local = nullptr;
    int total = 0;
    int chunk = 0;          // rows per wave, 0 = per-mode default
    int waves_per_xcd = 0;  // persistent grid, 0 = default
    int subs = 0;           // substeps per launch: 0 = by mode, 1, 2
    int pipe = 0;           // substep pairs as a two-wave pipeline: 0 = by mode, 1 = off, 2 = on
    int built_chunk = -1;
    bool built_two = false;
    unsigned table_gen = 0;  // bumped whenever build_tables replaces table[] / item_*
    int phase = 0;          // buffer set holding the current state
    // graph of one outer step for (tau, nstep, start phase)
    // graphs of 1 and of GRAPH_STEPS outer steps, per start phase
    hipGraphExec_t gexec[2][2] = {};
    hipGraphExec_t gexec_plain[2] = {};  // one outer step without the closure (interval > 1)
    double g_tau = -1.0;
    int g_nstep = -1, g_chunk = -1, g_wpx = -1;
    bool g_two = false, g_pipe = false;
    // what the graphs' kernel nodes point at or hold by value: the tables'
    // generation, and the closure network and its version (0 / null: none)
    unsigned g_table_gen = 0, g_net_version = 0;
    const inmsom_bt_cnn *g_net = nullptr;
    double *rows = nullptr;
    unsigned long long *cnt = nullptr;
    // snapshot of every member's state (inmsom_bt_ens_save / _restore)
    std::vector<void *> snap;
    // learned closure applied after every outer step
    inmsom_bt_cnn *closure = nullptr;
    bool closure_lost = false;  // the attached network was destroyed: runs fail until set_closure
    double closure_ratio = 3.0;
    // the closure runs after every cl_every-th outer step (1, 2, 4 or 8),
    // applying cl_every steps' worth of its tendency; cl_phase = outer steps
    // since it last ran
    int cl_every = 1, cl_phase = 0, g_every = 1;
    ibt::cnn::Infer inf;
    ibt::cnn::ClosureMember *cl_mem[2] = {};  // device tables per buffer set
};

namespace {

int fail(const std::string &m, int code)
{
    set_error(m);
    return code;
}

#define EHC(expr)                                                                            \
    do {                                                                                     \
        hipError_t e_ = (expr);                                                              \
        if (e_ != hipSuccess) return fail(std::string(#expr) + ": " + hipGetErrorString(e_), \
                                          INMSOM_BT_EDEVICE);                                \
    } while (0)

size_t esz(int mode) { return mode == INMSOM_BT_FP32 ? 4 : 8; }

// two substeps per launch: as set, or by mode and size as for one basin
// (not in FP64_EXACT, whose IEEE divides make it compute-bound, nor below
// TWO_SUBSTEP_CELLS cells in all, where the launches are latency-bound)
bool two_per_launch(const inmsom_bt_ens *e)
{
    if (e->subs > 0) return e->subs == 2;
    long long cells = 0;
    for (const EnsMember &m : e->mem) cells += (long long)m.n;
    return e->mode != INMSOM_BT_FP64_EXACT && cells >= TWO_SUBSTEP_CELLS;
}

// substep pairs as a two-wave pipeline (FP64 and FP32): as set, or by default
bool pipe_pairs(const inmsom_bt_ens *e)
{
    if (e->mode == INMSOM_BT_FP64_EXACT || !two_per_launch(e)) return false;
    return e->pipe != 1;
}

// the batch kernel for (mode, two substeps, forming substep)
template <bool FORM>
const void *batch_kernel(int mode, bool two, bool pipe = false)
{
    if (two && pipe) {
        if (mode == INMSOM_BT_FP64) return (const void *)&k_fused_batch_pipe<double>;
        return (const void *)&k_fused_batch_pipe<float>;
    }
    if (two) {
        if (mode == INMSOM_BT_FP64_EXACT) return (const void *)&k_fused_batch2<double, true, FORM>;
        if (mode == INMSOM_BT_FP64) return (const void *)&k_fused_batch2<double, false, FORM>;
        return (const void *)&k_fused_batch2<float, false, FORM>;
    }
    if (mode == INMSOM_BT_FP64_EXACT) return (const void *)&k_fused_batch<double, true, FORM>;
    if (mode == INMSOM_BT_FP64) return (const void *)&k_fused_batch<double, false, FORM>;
    return (const void *)&k_fused_batch<float, false, FORM>;
}

// Workgroups of the batch kernel one XCD holds at once (one wave each, two
// for the pipeline): its CUs x the kernel's workgroups per CU from the
// occupancy calculator.
int default_waves(const inmsom_bt_ens *e, bool two)
{
    int cus = 0, per_cu = 0;
    const bool pipe = two && pipe_pairs(e);
    if (hipDeviceGetAttribute(&cus, hipDeviceAttributeMultiprocessorCount, e->device) != hipSuccess || cus < NXCD)
        cus = 256;
    if (hipOccupancyMaxActiveBlocksPerMultiprocessor(&per_cu, batch_kernel<true>(e->mode, two, pipe),
                                                     pipe ? 2 * WAVE : WAVE, 0) != hipSuccess ||
        per_cu < 1)
        per_cu = 4;
    return cus / NXCD * per_cu;
}

// Rows per wave: the chunk (2..64) that minimises rounds x (chunk + warm-up
// rows) over all members' items for the persistent grid's wave slots.
int default_chunk(const inmsom_bt_ens *e, bool two, int wpx)
{
    const int sw = two ? SW2 : SW, warm = two ? (pipe_pairs(e) ? pipe_extra(2) : 7) : 3;
    const long long slots = (long long)wpx * NXCD;
    int best = 16;
    long long best_cost = -1;
    for (int c = 2; c <= 64; ++c) {
        long long items = 0;
        for (const EnsMember &m : e->mem) items += (long long)((m.nx + 2 + sw - 1) / sw) * ((m.ny + c - 1) / c);
        const long long cost = (items + slots - 1) / slots * (c + warm);
        if (best_cost < 0 || cost < best_cost) {
            best_cost = cost;
            best = c;
        }
    }
    return best;
}

template <class T>
int upload_vec(inmsom_bt_ens *e, T **dst, const std::vector<T> &src)
{
    EHC(hipMalloc(dst, src.size() * sizeof(T)));
    EHC(hipMemcpy(*dst, src.data(), src.size() * sizeof(T), hipMemcpyHostToDevice));
    (void)e;
    return INMSOM_BT_OK;
}

int upload_field(inmsom_bt_ens *e, void **dst, const double *src, size_t n)
{
    if (e->mode == INMSOM_BT_FP32) {
        std::vector<float> t(n);
        for (size_t i = 0; i < n; ++i) t[i] = (float)src[i];
        return upload_vec(e, reinterpret_cast<float **>(dst), t);
    }
    std::vector<double> t(src, src + n);
    return upload_vec(e, reinterpret_cast<double **>(dst), t);
}

// Builds one member from its grid and forcing (host side, as fused_prepare).
int build_member(inmsom_bt_ens *e, const inmsom_bt_grid *gr, const inmsom_bt_forcing *f,
                 EnsMember &mb)
{
    if (!gr || !f) return fail("null member grid/forcing", INMSOM_BT_EINVAL);
    if (gr->nx < 5 || gr->ny < 5) return fail("member too small", INMSOM_BT_EINVAL);
    if (!gr->lbasins && !gr->lu) return fail("member needs lbasins or masks", INMSOM_BT_EINVAL);
    if (gr->periodicity_x || gr->periodicity_y || gr->full_free_surface != 1)
        return fail("ensemble members must be closed basins with full_free_surface = 1",
                    INMSOM_BT_EINVAL);
    const int nx = gr->nx, ny = gr->ny;
    const size_t n = (size_t)nx * ny;
    mb.nx = nx;
    mb.ny = ny;
    mb.n = n;
    std::vector<float> mk[7];
    const float *m7[7] = {gr->lu, gr->luu, gr->luh, gr->lcu, gr->lcv, gr->llu, gr->llv};
    if (!gr->lu) {
        float *p7[7];
        for (int k = 0; k < 7; ++k) {
            mk[k].assign(n, 0.0f);
            p7[k] = mk[k].data();
        }
        int rc = inmsom_bt_gridcon(nx, ny, 0, 0, gr->lbasins, p7[0], p7[1], p7[2], p7[3], p7[4], p7[5],
                                   p7[6]);
        if (rc) return rc;
        for (int k = 0; k < 7; ++k) m7[k] = p7[k];
    }
    std::vector<uint8_t> bits(n, 0);
    const uint8_t bit[7] = {MB_LU, MB_LUU, MB_LUH, MB_LCU, MB_LCV, MB_LLU, MB_LLV};
    for (size_t i = 0; i < n; ++i)
        for (int k = 0; k < 7; ++k)
            if (m7[k][i] > 0.5f) bits[i] |= bit[k];
    // uniform metrics
    const double *met[8] = {gr->dx, gr->dy, gr->dxt, gr->dyt, gr->dxh, gr->dyh, gr->dxb, gr->dyb};
    const double d = gr->dx[0];
    for (int k = 0; k < 8; ++k)
        for (size_t i = 0; i < n; ++i)
            if (met[k][i] != d) return fail("ensemble members need uniform metrics", INMSOM_BT_EINVAL);
    mb.d = d;
    mb.area = d * d;
    // viscosity = nu on wet cells
    double nu = -1.0;
    for (size_t i = 0; i < n; ++i) {
        if (bits[i] & MB_LU) {
            if (nu < 0.0) nu = f->mu[i];
            if (f->mu[i] != nu) return fail("member mu must be constant on wet cells", INMSOM_BT_EINVAL);
        } else if (f->mu[i] != 0.0) {
            return fail("member mu must be 0 on land", INMSOM_BT_EINVAL);
        }
    }
    mb.nu = nu < 0.0 ? 0.0 : nu;
    // row tables (Function/fluxes.f90:30-36 with x-uniform wind)
    std::vector<double> rf(ny), rsx(ny), rsy(ny), rwf(ny);
    for (int r = 0; r < ny; ++r) {
        const size_t o = (size_t)r * nx;
        for (int m = 1; m < nx; ++m)
            if (f->taux[o + m] != f->taux[o] || f->tauy[o + m] != f->tauy[o] ||
                f->wflux[o + m] != f->wflux[o] || gr->rlh_s[o + m] != gr->rlh_s[o])
                return fail("member wind, water flux and Coriolis must be uniform along x",
                            INMSOM_BT_EINVAL);
        rf[r] = gr->rlh_s[o];
        rwf[r] = f->wflux[o];
        rsx[r] = (f->taux[o] * d * d + f->taux[o] * d * d) / gr->ref_den / 2.0 / d / d;
        rsy[r] = (f->tauy[o] * d * d + f->tauy[o] * d * d) / gr->ref_den / 2.0 / d / d;
    }
    int rc;
    if ((rc = upload_vec(e, &mb.mb, bits)) || (rc = upload_vec(e, &mb.row_f, rf)) ||
        (rc = upload_vec(e, &mb.row_sx, rsx)) || (rc = upload_vec(e, &mb.row_sy, rsy)) ||
        (rc = upload_vec(e, &mb.row_wf, rwf)))
        return rc;
    if ((rc = upload_field(e, &mb.H, gr->hhq_rest, n)) || (rc = upload_field(e, &mb.rd, f->rdis, n)))
        return rc;
    const size_t es = esz(e->mode);
    for (int b = 0; b < 2; ++b)
        for (int k = 0; k < 3; ++k) {
            EHC(hipMalloc(&mb.lvl[b][k], n * es));
            EHC(hipMemset(mb.lvl[b][k], 0, n * es));
            EHC(hipMalloc(&mb.flt[b][k], n * 8));
            EHC(hipMemset(mb.flt[b][k], 0, n * 8));
        }
    EHC(hipMalloc(&mb.frx, n * es));
    EHC(hipMalloc(&mb.fry, n * es));
    EHC(hipMalloc(&mb.hup_c, n * 8));
    EHC(hipMalloc(&mb.hvp_c, n * 8));
    return INMSOM_BT_OK;
}

FusedParams member_params(const inmsom_bt_ens *e, const EnsMember &mb, int phase, int chunk, bool two)
{
    FusedParams p{};
    p.nx = mb.nx;
    p.ny = mb.ny;
    const int sw = two ? SW2 : SW;
    p.strips = (mb.nx + 2 + sw - 1) / sw;
    p.chunk = chunk;
    p.nchunks = (mb.ny + chunk - 1) / chunk;
    p.total = p.strips * p.nchunks;
    p.per_xcd = 0;
    p.mb = mb.mb;
    p.H = mb.H;
    p.rd = mb.rd;
    const int b = phase, o = 1 - phase;
    p.u = mb.lvl[b][0];
    p.v = mb.lvl[b][1];
    p.s = mb.lvl[b][2];
    p.up = mb.flt[b][0];
    p.vp = mb.flt[b][1];
    p.sp = mb.flt[b][2];
    p.uo = mb.lvl[o][0];
    p.vo = mb.lvl[o][1];
    p.so = mb.lvl[o][2];
    p.upo = mb.flt[o][0];
    p.vpo = mb.flt[o][1];
    p.spo = mb.flt[o][2];
    p.row_f = mb.row_f;
    p.row_sx = mb.row_sx;
    p.row_sy = mb.row_sy;
    p.row_wf = mb.row_wf;
    p.g = e->g;
    p.rho0 = e->rho0;
    p.ts = e->ts;
    p.cb = e->cb;
    p.eb = e->eb;
    p.nu = mb.nu;
    p.ffs = e->ffs;
    p.d = mb.d;
    p.frx = mb.frx;
    p.fry = mb.fry;
    p.frx_in = mb.frx;
    p.fry_in = mb.fry;
    p.hup_c = mb.hup_c;
    p.hvp_c = mb.hvp_c;
    return p;
}

// (Re)builds the descriptor tables and the work list for a launch shape.
int build_tables(inmsom_bt_ens *e, int chunk, bool two)
{
    if (e->built_chunk == chunk && e->built_two == two) return INMSOM_BT_OK;
    std::vector<FusedParams> t[2];
    std::vector<int> im, il;
    for (size_t k = 0; k < e->mem.size(); ++k) {
        for (int ph = 0; ph < 2; ++ph) t[ph].push_back(member_params(e, e->mem[k], ph, chunk, two));
        const int items = t[0].back().total;
        e->mem[k].items = items;
        for (int i = 0; i < items; ++i) {
            im.push_back((int)k);
            il.push_back(i);
        }
    }
    for (int ph = 0; ph < 2; ++ph) {
        if (e->table[ph]) EHC(hipFree(e->table[ph]));
        e->table[ph] = nullptr;
        int rc = upload_vec(e, &e->table[ph], t[ph]);
        if (rc) return rc;
    }
    if (e->item_member) EHC(hipFree(e->item_member));
    if (e->item_local) EHC(hipFree(e->item_local));
    e->item_member = e->item_local = nullptr;
    int rc;
    if ((rc = upload_vec(e, &e->item_member, im)) || (rc = upload_vec(e, &e->item_local, il))) return rc;
    e->total = (int)im.size();
    e->built_chunk = chunk;
    e->built_two = two;
    ++e->table_gen;
    return INMSOM_BT_OK;
}

void launch_batch(inmsom_bt_ens *e, int phase, double tau_inner, int sub, int nsub, int wpx, bool two)
{
    int per_xcd = (e->total + NXCD - 1) / NXCD;
    const bool pipe = two && pipe_pairs(e);
    dim3 grid((unsigned)(wpx * NXCD)), blk(pipe ? 2 * WAVE : WAVE);
    const FusedParams *tb = e->table[phase];
    const int *im = e->item_member, *il = e->item_local;
    int total = e->total;
    void *args[] = {&tb, &im, &il, &total, &per_xcd, &wpx, &tau_inner, &sub, &nsub};
    const void *k = sub == 1 ? batch_kernel<true>(e->mode, two, pipe) : batch_kernel<false>(e->mode, two, pipe);
    hipLaunchKernel(k, grid, blk, args, 0, e->stream);
}

// launches per outer step: substep pairs, and an odd last substep alone
int launches_per_step(int nstep, bool two) { return two ? nstep / 2 + nstep % 2 : nstep; }

// launches of one outer step starting from buffer set `phase`, the
// closure after it when `closure` (and one is attached)
void outer_step(inmsom_bt_ens *e, int phase, double tau, int nstep, int wpx, bool two, bool closure = true)
{
    for (int sub = 1; sub <= nstep;) {
        const bool pair = two && sub + 1 <= nstep;
        launch_batch(e, phase, tau / (double)nstep, sub, nstep, wpx, pair);
        phase = 1 - phase;
        sub += pair ? 2 : 1;
    }
    if (e->closure && closure)
        ibt::cnn::closure_launch(e->closure, e->stream, e->mode == INMSOM_BT_FP32, e->cl_mem[phase], e->inf,
                                 tau * e->cl_every, e->closure_ratio);
}

void drop_closure(inmsom_bt_ens *e)
{
    if (e->closure) ibt::cnn::net_detach(e->closure, e);
    ibt::cnn::infer_destroy(&e->inf);
    for (auto &t : e->cl_mem)
        if (t) {
            hipFree(t);
            t = nullptr;
        }
    e->closure = nullptr;
}

// Outer steps per replay of the long graph: one replay costs about 8 us
// on MI355X, the whole step of a small basin with the closure about 75 us.
constexpr int GRAPH_STEPS = 8;
static_assert(GRAPH_STEPS % 2 == 0, "a long graph must end in the phase it starts in");

void drop_graphs(inmsom_bt_ens *e)
{
    for (auto &gl : e->gexec)
        for (auto &g : gl)
            if (g) {
                hipGraphExecDestroy(g);
                g = nullptr;
            }
    for (auto &g : e->gexec_plain)
        if (g) {
            hipGraphExecDestroy(g);
            g = nullptr;
        }
    e->g_nstep = -1;
}

// The graphs' kernel nodes hold the table and work-list addresses of the
// build they were captured under, and the closure's normalisation by
// value: they are valid only while none of those has changed.
bool graphs_current(const inmsom_bt_ens *e, double tau, int nstep, int chunk, int wpx, bool two)
{
    const unsigned nv = e->closure ? ibt::cnn::net_version(e->closure) : 0u;
    return e->gexec[0][0] && e->g_tau == tau && e->g_nstep == nstep && e->g_chunk == chunk &&
           e->g_wpx == wpx && e->g_two == two && e->g_pipe == pipe_pairs(e) && e->g_table_gen == e->table_gen && e->g_net == e->closure &&
           e->g_net_version == nv && e->g_every == e->cl_every;
}

// inmsom_bt_cnn_destroy on the attached network: its memory is about to
// be freed, so the graphs that point at it and the inference buffers go now.
void closure_destroyed(void *owner)
{
    auto *e = static_cast<inmsom_bt_ens *>(owner);
    hipSetDevice(e->device);
    if (e->stream) hipStreamSynchronize(e->stream);
    drop_graphs(e);
    e->closure = nullptr;  // already detached by the network
    drop_closure(e);
    e->closure_lost = true;
}

// Graphs per starting buffer set: one outer step (with the closure), 8
// outer steps starting right after a closure application (the closure
// after every cl_every-th of them; cl_every divides 8), and with an
// interval above 1 one outer step without the closure.
int ensure_graphs(inmsom_bt_ens *e, double tau, int nstep, int chunk, int wpx, bool two)
{
    if (graphs_current(e, tau, nstep, chunk, wpx, two)) return INMSOM_BT_OK;
    drop_graphs(e);
    const bool flip = launches_per_step(nstep, two) % 2;
    const int nph = flip ? 2 : 1;
    const int kinds = (e->closure && e->cl_every > 1) ? 3 : 2;
    for (int kind = 0; kind < kinds; ++kind)
        for (int ph = 0; ph < nph; ++ph) {
            hipGraph_t graph = nullptr;
            EHC(hipStreamBeginCapture(e->stream, hipStreamCaptureModeThreadLocal));
            int phase = ph;
            const int len = kind == 1 ? GRAPH_STEPS : 1;
            for (int k = 0; k < len; ++k) {
                const bool cl = kind == 0 || (kind == 1 && (k + 1) % e->cl_every == 0);
                outer_step(e, phase, tau, nstep, wpx, two, cl);
                if (flip) phase = 1 - phase;
            }
            EHC(hipStreamEndCapture(e->stream, &graph));
            hipGraphExec_t *dst = kind == 2 ? &e->gexec_plain[ph] : &e->gexec[kind][ph];
            hipError_t r = hipGraphInstantiate(dst, graph, nullptr, nullptr, 0);
            hipGraphDestroy(graph);
            if (r != hipSuccess)
                return fail(std::string("hipGraphInstantiate: ") + hipGetErrorString(r), INMSOM_BT_EDEVICE);
        }
    e->g_tau = tau;
    e->g_nstep = nstep;
    e->g_chunk = chunk;
    e->g_wpx = wpx;
    e->g_two = two;
    e->g_pipe = pipe_pairs(e);
    e->g_table_gen = e->table_gen;
    e->g_net = e->closure;
    e->g_net_version = e->closure ? ibt::cnn::net_version(e->closure) : 0u;
    e->g_every = e->cl_every;
    return INMSOM_BT_OK;
}

// the launch shape of the next run: (two substeps per launch, waves per
// XCD, rows per wave), each as set or by default
void launch_shape(const inmsom_bt_ens *e, bool *two, int *wpx, int *chunk)
{
    *two = two_per_launch(e);
    *wpx = e->waves_per_xcd > 0 ? e->waves_per_xcd : default_waves(e, *two);
    *chunk = e->chunk > 0 ? e->chunk : default_chunk(e, *two, *wpx);
}

int run_steps(inmsom_bt_ens *e, double tau, int nstep, int nsteps, int use_graph)
{
    if (e->closure_lost)
        return fail("the closure network of this ensemble was destroyed: attach another or detach it "
                    "(inmsom_bt_ens_set_closure)", INMSOM_BT_ESTATE);
    bool two;
    int wpx, chunk;
    launch_shape(e, &two, &wpx, &chunk);
    int rc = build_tables(e, chunk, two);
    if (rc) return rc;
    const bool flip = launches_per_step(nstep, two) % 2;
    const int every = e->closure ? e->cl_every : 1;
    // one outer step; the closure after it when it completes an interval
    auto single = [&](bool graph) -> int {
        const bool cl = (e->cl_phase + 1) % every == 0;
        if (graph) {
            hipGraphExec_t g = cl ? e->gexec[0][flip ? e->phase : 0] : e->gexec_plain[flip ? e->phase : 0];
            EHC(hipGraphLaunch(g, e->stream));
        } else {
            outer_step(e, e->phase, tau, nstep, wpx, two, cl);
        }
        if (flip) e->phase = 1 - e->phase;
        e->cl_phase = (e->cl_phase + 1) % every;
        return INMSOM_BT_OK;
    };
    if (use_graph) {
        rc = ensure_graphs(e, tau, nstep, chunk, wpx, two);
        if (rc) return rc;
        int s = 0;
        for (; s < nsteps && e->cl_phase != 0; ++s)  // to the start of an interval
            if ((rc = single(true))) return rc;
        for (; s + GRAPH_STEPS <= nsteps; s += GRAPH_STEPS)  // GRAPH_STEPS is even: the phase comes back
            EHC(hipGraphLaunch(e->gexec[1][flip ? e->phase : 0], e->stream));
        for (; s < nsteps; ++s)
            if ((rc = single(true))) return rc;
    } else {
        for (int s = 0; s < nsteps; ++s)
            if ((rc = single(false))) return rc;
    }
    EHC(hipGetLastError());
    return INMSOM_BT_OK;
}

}  // namespace

extern "C" {

int inmsom_bt_ens_create(int nmembers, const inmsom_bt_grid *grids, const inmsom_bt_forcing *forcing,
                         int mode, int device, inmsom_bt_ens **out)
{
    if (!out || !grids || !forcing || nmembers < 1) return fail("bad arguments", INMSOM_BT_EINVAL);
    if (mode != INMSOM_BT_FP64_EXACT && mode != INMSOM_BT_FP64 && mode != INMSOM_BT_FP32)
        return fail("unknown mode", INMSOM_BT_EINVAL);
    *out = nullptr;
    int ndev = 0;
    EHC(hipGetDeviceCount(&ndev));
    if (device < 0 || device >= ndev) return fail("device index out of range", INMSOM_BT_EINVAL);
    EHC(hipSetDevice(device));
    auto *e = new inmsom_bt_ens();
    e->device = device;
    e->mode = mode;
    e->g = grids[0].free_fall_acc;
    e->rho0 = grids[0].ref_den;
    e->ts = grids[0].time_smooth;
    e->cb = forcing[0].cb_nl;
    e->eb = forcing[0].ebottom;
    int rc = INMSOM_BT_OK;
    if (hipStreamCreateWithFlags(&e->stream, hipStreamNonBlocking) != hipSuccess)
        rc = fail("hipStreamCreate failed", INMSOM_BT_EDEVICE);
    for (int k = 0; k < nmembers && rc == INMSOM_BT_OK; ++k) {
        if (grids[k].free_fall_acc != e->g || grids[k].ref_den != e->rho0 ||
            grids[k].time_smooth != e->ts || forcing[k].cb_nl != e->cb || forcing[k].ebottom != e->eb) {
            rc = fail("members must share g, rho0, time_smooth and the drag parameters", INMSOM_BT_EINVAL);
            break;
        }
        e->mem.emplace_back();
        rc = build_member(e, &grids[k], &forcing[k], e->mem.back());
    }
    if (rc == INMSOM_BT_OK && hipMalloc(&e->cnt, 8) != hipSuccess)
        rc = fail("hipMalloc failed", INMSOM_BT_ENOMEM);
    if (rc != INMSOM_BT_OK) {
        inmsom_bt_ens_destroy(e);
        return rc;
    }
    *out = e;
    return INMSOM_BT_OK;
}

int inmsom_bt_ens_destroy(inmsom_bt_ens *e)
{
    if (!e) return INMSOM_BT_OK;
    hipSetDevice(e->device);
    if (e->stream) hipStreamSynchronize(e->stream);
    drop_graphs(e);
    for (auto &mb : e->mem) {
        for (void *q : {(void *)mb.mb, mb.H, mb.rd, mb.frx, mb.fry, (void *)mb.hup_c, (void *)mb.hvp_c,
                        (void *)mb.row_f, (void *)mb.row_sx, (void *)mb.row_sy, (void *)mb.row_wf})
            if (q) hipFree(q);
        for (int b = 0; b < 2; ++b)
            for (int k = 0; k < 3; ++k) {
                if (mb.lvl[b][k]) hipFree(mb.lvl[b][k]);
                if (mb.flt[b][k]) hipFree(mb.flt[b][k]);
            }
    }
    for (void *q : e->snap)
        if (q) hipFree(q);
    drop_closure(e);
    for (void *q : {(void *)e->table[0], (void *)e->table[1], (void *)e->item_member,
                    (void *)e->item_local, (void *)e->rows, (void *)e->cnt})
        if (q) hipFree(q);
    if (e->stream) hipStreamDestroy(e->stream);
    delete e;
    return INMSOM_BT_OK;
}

int inmsom_bt_ens_size(inmsom_bt_ens *e, int *nmembers, long long *wet_cells, long long *cells)
{
    if (!e || !nmembers || !cells) return fail("null pointer", INMSOM_BT_EINVAL);
    *nmembers = (int)e->mem.size();
    long long c = 0;
    for (auto &mb : e->mem) c += (long long)mb.n;
    *cells = c;
    if (wet_cells) {
        long long w = 0;
        for (auto &mb : e->mem) {
            std::vector<uint8_t> bits(mb.n);
            EHC(hipMemcpy(bits.data(), mb.mb, mb.n, hipMemcpyDeviceToHost));
            for (int n = 3; n <= mb.ny - 2; ++n)
                for (int m = 3; m <= mb.nx - 2; ++m)
                    if (bits[(size_t)(n - 1) * mb.nx + (m - 1)] & MB_LU) ++w;
        }
        *wet_cells = w;
    }
    return INMSOM_BT_OK;
}

int inmsom_bt_ens_set_state(inmsom_bt_ens *e, int k, const double *u, const double *up, const double *v,
                            const double *vp, const double *ssh, const double *sshp)
{
    if (!e || k < 0 || k >= (int)e->mem.size()) return fail("bad member index", INMSOM_BT_EINVAL);
    e->cl_phase = 0;
    if (!u || !up || !v || !vp || !ssh || !sshp) return fail("null pointer", INMSOM_BT_EINVAL);
    EHC(hipSetDevice(e->device));
    EHC(hipStreamSynchronize(e->stream));
    EnsMember &mb = e->mem[k];
    const double *lv[3] = {u, v, ssh}, *fl[3] = {up, vp, sshp};
    const int b = e->phase;
    for (int q = 0; q < 3; ++q) {
        if (e->mode == INMSOM_BT_FP32) {
            std::vector<float> t(mb.n);
            for (size_t i = 0; i < mb.n; ++i) t[i] = (float)lv[q][i];
            EHC(hipMemcpy(mb.lvl[b][q], t.data(), mb.n * 4, hipMemcpyHostToDevice));
        } else {
            EHC(hipMemcpy(mb.lvl[b][q], lv[q], mb.n * 8, hipMemcpyHostToDevice));
        }
        EHC(hipMemcpy(mb.flt[b][q], fl[q], mb.n * 8, hipMemcpyHostToDevice));
    }
    return INMSOM_BT_OK;
}

int inmsom_bt_ens_get_state(inmsom_bt_ens *e, int k, double *u, double *up, double *v, double *vp,
                            double *ssh, double *sshp)
{
    if (!e || k < 0 || k >= (int)e->mem.size()) return fail("bad member index", INMSOM_BT_EINVAL);
    if (!u || !up || !v || !vp || !ssh || !sshp) return fail("null pointer", INMSOM_BT_EINVAL);
    EHC(hipSetDevice(e->device));
    EHC(hipStreamSynchronize(e->stream));
    EnsMember &mb = e->mem[k];
    double *lv[3] = {u, v, ssh}, *fl[3] = {up, vp, sshp};
    const int b = e->phase;
    for (int q = 0; q < 3; ++q) {
        if (e->mode == INMSOM_BT_FP32) {
            std::vector<float> t(mb.n);
            EHC(hipMemcpy(t.data(), mb.lvl[b][q], mb.n * 4, hipMemcpyDeviceToHost));
            for (size_t i = 0; i < mb.n; ++i) lv[q][i] = (double)t[i];
        } else {
            EHC(hipMemcpy(lv[q], mb.lvl[b][q], mb.n * 8, hipMemcpyDeviceToHost));
        }
        EHC(hipMemcpy(fl[q], mb.flt[b][q], mb.n * 8, hipMemcpyDeviceToHost));
    }
    return INMSOM_BT_OK;
}

int inmsom_bt_ens_set_launch(inmsom_bt_ens *e, int rows_per_wave, int waves_per_xcd)
{
    if (!e) return fail("null ensemble", INMSOM_BT_EINVAL);
    if (rows_per_wave < 0 || rows_per_wave > 4096 || waves_per_xcd < 0 || waves_per_xcd > 65536)
        return fail("launch shape out of range", INMSOM_BT_EINVAL);
    e->chunk = rows_per_wave;
    e->waves_per_xcd = waves_per_xcd;
    return INMSOM_BT_OK;
}

int inmsom_bt_ens_set_substeps_per_launch(inmsom_bt_ens *e, int substeps)
{
    if (!e) return fail("null ensemble", INMSOM_BT_EINVAL);
    if (substeps < 0 || substeps > 2) return fail("substeps per launch must be 0 (by mode), 1 or 2", INMSOM_BT_EINVAL);
    e->subs = substeps;
    return INMSOM_BT_OK;
}

int inmsom_bt_ens_set_pipeline(inmsom_bt_ens *e, int stages)
{
    if (!e) return fail("null ensemble", INMSOM_BT_EINVAL);
    if (stages < 0 || stages > 2)
        return fail("ensemble pipeline stages must be 0 (by mode), 1 (off) or 2", INMSOM_BT_EINVAL);
    e->pipe = stages;
    return INMSOM_BT_OK;
}

int inmsom_bt_ens_get_pipeline(inmsom_bt_ens *e, int *stages)
{
    if (!e || !stages) return fail("null pointer", INMSOM_BT_EINVAL);
    *stages = pipe_pairs(e) ? 2 : 0;
    return INMSOM_BT_OK;
}

int inmsom_bt_ens_get_substeps_per_launch(inmsom_bt_ens *e, int *substeps)
{
    if (!e || !substeps) return fail("null pointer", INMSOM_BT_EINVAL);
    *substeps = two_per_launch(e) ? 2 : 1;
    return INMSOM_BT_OK;
}

int inmsom_bt_ens_run(inmsom_bt_ens *e, double tau, int nstep, int nsteps, int use_graph)
{
    if (!e) return fail("null ensemble", INMSOM_BT_EINVAL);
    if (!(tau > 0.0) || nstep < 1 || nstep > 1000 || nsteps < 0)
        return fail("bad tau / nstep / nsteps", INMSOM_BT_EINVAL);
    EHC(hipSetDevice(e->device));
    int rc = run_steps(e, tau, nstep, nsteps, use_graph);
    if (rc) return rc;
    EHC(hipStreamSynchronize(e->stream));
    return INMSOM_BT_OK;
}

int inmsom_bt_ens_run_timed(inmsom_bt_ens *e, double tau, int nstep, int nsteps, int use_graph,
                            double *ms)
{
    if (!e || !ms) return fail("null pointer", INMSOM_BT_EINVAL);
    if (!(tau > 0.0) || nstep < 1 || nstep > 1000 || nsteps < 0)
        return fail("bad tau / nstep / nsteps", INMSOM_BT_EINVAL);
    EHC(hipSetDevice(e->device));
    // build tables and graphs before the timed region
    bool two;
    int wpx, chunk;
    launch_shape(e, &two, &wpx, &chunk);
    int rc = build_tables(e, chunk, two);
    if (rc) return rc;
    if (use_graph && (rc = ensure_graphs(e, tau, nstep, chunk, wpx, two))) return rc;
    hipEvent_t e0, e1;
    EHC(hipEventCreate(&e0));
    EHC(hipEventCreate(&e1));
    EHC(hipEventRecord(e0, e->stream));
    rc = run_steps(e, tau, nstep, nsteps, use_graph);
    if (rc) return rc;
    EHC(hipEventRecord(e1, e->stream));
    EHC(hipEventSynchronize(e1));
    float t = 0.f;
    EHC(hipEventElapsedTime(&t, e0, e1));
    hipEventDestroy(e0);
    hipEventDestroy(e1);
    *ms = t;
    return INMSOM_BT_OK;
}

int inmsom_bt_ens_save(inmsom_bt_ens *e)
{
    if (!e) return fail("null ensemble", INMSOM_BT_EINVAL);
    EHC(hipSetDevice(e->device));
    const size_t es = esz(e->mode);
    if (e->snap.empty()) {
        for (auto &mb : e->mem)
            for (int q = 0; q < 6; ++q) {
                void *p = nullptr;
                EHC(hipMalloc(&p, mb.n * (q < 3 ? es : 8)));
                e->snap.push_back(p);
            }
    }
    size_t j = 0;
    for (auto &mb : e->mem)
        for (int q = 0; q < 6; ++q, ++j) {
            const void *src = q < 3 ? mb.lvl[e->phase][q] : (const void *)mb.flt[e->phase][q - 3];
            EHC(hipMemcpyAsync(e->snap[j], src, mb.n * (q < 3 ? es : 8), hipMemcpyDeviceToDevice,
                               e->stream));
        }
    EHC(hipStreamSynchronize(e->stream));
    return INMSOM_BT_OK;
}

int inmsom_bt_ens_restore(inmsom_bt_ens *e)
{
    if (!e) return fail("null ensemble", INMSOM_BT_EINVAL);
    if (e->snap.empty()) return fail("inmsom_bt_ens_save was not called", INMSOM_BT_ESTATE);
    EHC(hipSetDevice(e->device));
    e->cl_phase = 0;
    const size_t es = esz(e->mode);
    size_t j = 0;
    for (auto &mb : e->mem)
        for (int q = 0; q < 6; ++q, ++j) {
            void *dst = q < 3 ? mb.lvl[e->phase][q] : (void *)mb.flt[e->phase][q - 3];
            EHC(hipMemcpyAsync(dst, e->snap[j], mb.n * (q < 3 ? es : 8), hipMemcpyDeviceToDevice,
                               e->stream));
        }
    EHC(hipStreamSynchronize(e->stream));
    return INMSOM_BT_OK;
}

int inmsom_bt_ens_range_check(inmsom_bt_ens *e, long long *bad)
{
    if (!e || !bad) return fail("null pointer", INMSOM_BT_EINVAL);
    EHC(hipSetDevice(e->device));
    EHC(hipMemsetAsync(e->cnt, 0, 8, e->stream));
    for (auto &mb : e->mem) {
        const int b = e->phase;
        dim3 g((unsigned)((mb.n + 255) / 256)), blk(256);
        if (e->mode == INMSOM_BT_FP32)
            hipLaunchKernelGGL(k_member_range<float>, g, blk, 0, e->stream, mb.nx, mb.ny, mb.mb,
                               (const float *)mb.lvl[b][2], (const float *)mb.lvl[b][0],
                               (const float *)mb.lvl[b][1], e->cnt);
        else
            hipLaunchKernelGGL(k_member_range<double>, g, blk, 0, e->stream, mb.nx, mb.ny, mb.mb,
                               (const double *)mb.lvl[b][2], (const double *)mb.lvl[b][0],
                               (const double *)mb.lvl[b][1], e->cnt);
    }
    EHC(hipGetLastError());
    unsigned long long h = 0;
    EHC(hipMemcpyAsync(&h, e->cnt, 8, hipMemcpyDeviceToHost, e->stream));
    EHC(hipStreamSynchronize(e->stream));
    *bad = (long long)h;
    return h ? INMSOM_BT_ERANGE : INMSOM_BT_OK;
}

int inmsom_bt_ens_volume(inmsom_bt_ens *e, int k, double *vol)
{
    if (!e || !vol || k < 0 || k >= (int)e->mem.size()) return fail("bad arguments", INMSOM_BT_EINVAL);
    EHC(hipSetDevice(e->device));
    EnsMember &mb = e->mem[k];
    const int rows = mb.ny - 4;
    double *rb = nullptr;
    EHC(hipMalloc(&rb, (size_t)rows * 8));
    const int b = e->phase;
    if (e->mode == INMSOM_BT_FP32)
        hipLaunchKernelGGL(k_member_rows<float>, dim3(rows), dim3(256), 0, e->stream, mb.nx, mb.ny, mb.mb,
                           (const float *)mb.lvl[b][2], mb.area, rb);
    else
        hipLaunchKernelGGL(k_member_rows<double>, dim3(rows), dim3(256), 0, e->stream, mb.nx, mb.ny, mb.mb,
                           (const double *)mb.lvl[b][2], mb.area, rb);
    std::vector<double> h(rows);
    EHC(hipMemcpyAsync(h.data(), rb, (size_t)rows * 8, hipMemcpyDeviceToHost, e->stream));
    EHC(hipStreamSynchronize(e->stream));
    hipFree(rb);
    double t = 0.0;
    for (double x : h) t += x;
    *vol = t;
    return INMSOM_BT_OK;
}

int inmsom_bt_ens_set_closure_interval(inmsom_bt_ens *e, int every)
{
    if (!e) return fail("null ensemble", INMSOM_BT_EINVAL);
    if (every != 1 && every != 2 && every != 4 && every != 8)
        return fail("the closure interval must be 1, 2, 4 or 8 outer steps", INMSOM_BT_EINVAL);
    EHC(hipSetDevice(e->device));
    EHC(hipStreamSynchronize(e->stream));
    drop_graphs(e);
    e->cl_every = every;
    e->cl_phase = 0;
    return INMSOM_BT_OK;
}

int inmsom_bt_ens_set_closure(inmsom_bt_ens *e, inmsom_bt_cnn *net, double ratio)
{
    if (!e) return fail("null ensemble", INMSOM_BT_EINVAL);
    if (net && !(ratio > 0.0)) return fail("ratio must be positive", INMSOM_BT_EINVAL);
    EHC(hipSetDevice(e->device));
    EHC(hipStreamSynchronize(e->stream));
    drop_graphs(e);
    drop_closure(e);
    e->closure_lost = false;
    e->cl_phase = 0;
    if (!net) return INMSOM_BT_OK;
    if (ibt::cnn::net_device(net) != e->device) return fail("network and ensemble on different devices",
                                                            INMSOM_BT_EINVAL);
    if (e->mode == INMSOM_BT_FP64_EXACT)
        return fail("the closure runs with the FP64 and FP32 modes", INMSOM_BT_EINVAL);
    const int nm = (int)e->mem.size();
    std::vector<int> nx(nm), ny(nm);
    for (int k = 0; k < nm; ++k) {
        nx[k] = e->mem[k].nx;
        ny[k] = e->mem[k].ny;
    }
    int rc = ibt::cnn::infer_create(e->device, nm, nx.data(), ny.data(), &e->inf);
    if (rc) {
        drop_closure(e);
        return rc;
    }
    long long pix = 0, row = 0;
    for (int b = 0; b < 2; ++b) {
        std::vector<ibt::cnn::ClosureMember> t(nm);
        pix = row = 0;
        for (int k = 0; k < nm; ++k) {
            const EnsMember &mb = e->mem[k];
            t[k] = ibt::cnn::ClosureMember{mb.nx, mb.ny, pix, row, mb.lvl[b][0], mb.lvl[b][1], mb.lvl[b][2],
                                           mb.H, mb.mb, mb.d, mb.nu};
            pix += (long long)mb.nx * mb.ny;
            row += mb.ny;
        }
        if ((rc = upload_vec(e, &e->cl_mem[b], t))) {
            drop_closure(e);
            return rc;
        }
    }
    e->closure = net;
    ibt::cnn::net_attach(net, e, closure_destroyed);
    e->closure_ratio = ratio;
    return INMSOM_BT_OK;
}

int inmsom_bt_ens_closure_stats(inmsom_bt_ens *e, int *nfallback, double *norms)
{
    if (!e) return fail("null ensemble", INMSOM_BT_EINVAL);
    if (!e->closure) return fail("no closure attached", INMSOM_BT_ESTATE);
    EHC(hipSetDevice(e->device));
    EHC(hipStreamSynchronize(e->stream));
    const size_t nm = e->mem.size();
    if (nfallback) EHC(hipMemcpy(nfallback, e->inf.nfall, nm * sizeof(int), hipMemcpyDeviceToHost));
    if (norms) EHC(hipMemcpy(norms, e->inf.norms, nm * 2 * sizeof(double), hipMemcpyDeviceToHost));
    return INMSOM_BT_OK;
}

}  // extern "C"
