// Internal declarations shared by the inmsom_bt HIP sources.
#pragma once

#include <hip/hip_runtime.h>

#include <cstddef>
#include <cstdint>
#include <string>

#include "inmsom_bt.h"

namespace ibt {

// Static grid as the device kernels see it.  Indices are the reference's
// 1-based (m, n); element (m, n) lives at (n-1)*nx + (m-1).
struct DevGrid {
    int nx, ny;
    int px, py;  // periodicity flags
    int ffs;     // full_free_surface
    double g, rho0, ts;
    const float *lu, *luu, *luh, *lcu, *lcv, *llu, *llv;
    const double *dx, *dy, *dxt, *dyt, *dxh, *dyh, *dxb, *dyb;
    const double *rlh_s, *hr;
};

__host__ __device__ inline size_t IDX(const DevGrid &g, int m, int n)
{
    return (size_t)(n - 1) * (size_t)g.nx + (size_t)(m - 1);
}

// Tiles of a tiles_x x tiles_y grid in row-major order, dealt to the eight
// XCDs (the hardware hands workgroup b to XCD b % 8) as contiguous runs: the
// tiles one XCD holds at once are neighbours, so the cache lines a tile
// shares with the tiles beside it (rows are not whole 128-byte lines) and
// the rows above and below stay in that XCD's L2.  A launch of exactly
// tiles_x * tiles_y workgroups visits every tile once.
__device__ __forceinline__ void xcd_tile(int tiles_x, int tiles_y, int b, int &bx, int &by)
{
    const int T = tiles_x * tiles_y, xcd = b % 8, q = b / 8;
    const int per = T / 8, extra = T % 8;
    const int pos = xcd * per + (xcd < extra ? xcd : extra) + q;
    by = pos / tiles_x;
    bx = pos % tiles_x;
}

// Mask bits of the fused kernels (one byte per cell).
enum : uint8_t {
    MB_LU = 1, MB_LCU = 2, MB_LCV = 4, MB_LUU = 8, MB_LUH = 16, MB_LLU = 32, MB_LLV = 64
};

void set_error(const std::string &msg);

}  // namespace ibt

namespace ibt { struct FusedAlloc; struct DropinAlloc; }

// Context: every device allocation the library makes for one basin.
struct inmsom_bt_ctx {
    int device = 0;
    hipStream_t stream = nullptr;
    int nx = 0, ny = 0;
    size_t n = 0;  // nx*ny
    ibt::DevGrid g{};

    // static grid (device)
    float *mask[7] = {};
    double *metric[8] = {};
    double *rlh_s = nullptr, *hr = nullptr;
    bool uniform_metrics = false;
    double metric0 = 0.0;

    // forcing (device)
    double *taux = nullptr, *tauy = nullptr, *wflux = nullptr, *mu = nullptr, *mu4 = nullptr,
           *rdis = nullptr;
    double cb_nl = 2.5e-3, ebottom = 25.0e-4;
    bool have_forcing = false;

    // external-mode state: u, up, v, vp, ssh, sshp
    double *st[6] = {};
    // barotropic_dynamics working arrays: u, v, ssh, up, vp, sshp, un, vn, sshn
    double *w[9] = {};
    double *acc[4] = {};   // ubrtr_i, vbrtr_i, ssh4gradx, ssh4grady
    double *dep[12] = {};  // hhq_e .. hhhn_e
    double *vort = nullptr, *str_t = nullptr, *str_s = nullptr, *fx = nullptr, *fy = nullptr;
    double *radv = nullptr, *rady = nullptr, *rdfx = nullptr, *rdfy = nullptr;
    double *rhsx = nullptr, *rhsy = nullptr, *bf = nullptr;
    double *surf_x = nullptr, *surf_y = nullptr, *bot_x = nullptr, *bot_y = nullptr;

    // fused path
    int fused_mode = -1;           // mode the fused state is held in, -1 = none
    bool fused_ready = false;      // per-row tables built
    bool state_on_fused = false;   // authoritative state lives in the fused buffers
    uint8_t *mbits = nullptr;
    double *row_f = nullptr;       // Coriolis per H row
    double *row_wx = nullptr;      // wind stress on U faces per row (fluxes.f90:30-31)
    double *row_wy = nullptr;      // wind stress on V faces per row
    double nu = 0.0;               // constant viscosity on wet cells
    double *row_wf = nullptr;      // surface water flux per T row
    ibt::FusedAlloc *fa = nullptr;  // ping-pong state buffers of the fused step
    int fz_cur = 0;
    int chunk = 0;                 // output rows per wave of the fused step (0 = per-mode default)
    int subs_per_launch = 0;       // fused substeps per launch (0 = by mode, 1, 2 = two in one march)
    int pipe_k = 0;                // K substeps per launch as a K-wave pipeline (0 = off)
    long long pipe_slots[2][5] = {};  // workgroups of the pipeline kernel the device holds (FP64 / FP32, K)
    long long slots[6] = {};       // waves the device holds at once, per fused kernel (0 = not queried)
    double *scal = nullptr;        // device scalars (reductions)
    double *cpl = nullptr;         // staging of the 2-D -> 3-D coupling (bt_coupling.hip)
    size_t cpl_n = 0;              // doubles in cpl
    double *cpl_new = nullptr;     // one plane: ssh_new of the device-pointer coupling
    double *f3 = nullptr;          // staging of the 3-D forcing producer (bt_forcing3d.hip)
    size_t f3_n = 0;               // doubles in f3
    int f3_nz = 0;                 // levels of the staged fields (0 = none)
    bool f3_pg = false;            // staged pressure-gradient terms
    uint8_t *f3_mb = nullptr;      // mask bits of the forcing (lu, luu, lcu, lcv, llu, llv)
    double *f3_halo = nullptr;     // periodic halo table of the forcing
    double *f3_dz = nullptr;       // dz of the device-address forcing (up to 1000 levels)
    double *bc = nullptr;          // staging of baroclinic_dynamics (bt_baroclinic.hip)
    size_t bc_n = 0;               // doubles in bc
    int bc_nz = 0;                 // levels of the staged fields (0 = none)
    int bc_cur = 0;                // staged level n of uu, vv: 0 = slots UU, VV; 1 = UN, VN
    double *bc_dev = nullptr;      // new uu, vv and the vertical grid of the device-address call
    size_t bc_dev_n = 0;           // doubles in bc_dev
    double *vv = nullptr;          // staging of vertical_velocity (bt_vertical.hip)
    size_t vv_n = 0;               // doubles in vv
    int vv_nz = 0;                 // levels of the staged fields (0 = none)
    double *vv_dz = nullptr;       // dz of the device-address call (up to 1000 levels)
    double *pg = nullptr;          // staging of pressure_gradients (bt_pressure.hip)
    size_t pg_n = 0;               // doubles in pg
    int pg_nz = 0;                 // levels of the staged fields (0 = none)
    double *pg_z = nullptr;        // z of the device-address call (up to 1000 levels)
    double *eos = nullptr;         // staging of the density definition (bt_density.hip)
    size_t eos_n = 0;              // doubles in eos
    int eos_nz = 0;                // levels of the staged fields (0 = none)
    double *eos_z = nullptr;       // z of the device-address call (up to 1000 levels)
    double *sm = nullptr;          // staging of the Smagorinsky viscosity (bt_smagorinsky.hip)
    size_t sm_n = 0;               // doubles in sm
    int sm_nz = 0;                 // levels of the staged fields (0 = none)
    double sm_visc2 = 0.0, sm_dift = 0.0;  // coefficients of the staged call
    double *sm_dz = nullptr;       // dz of the device-address call (up to 1000 levels)
    double *sst = nullptr;         // staging of the 3-D stress_components (bt_smagorinsky.hip)
    size_t sst_n = 0;              // doubles in sst
    int sst_nz = 0;                // levels of the staged fields (0 = none)
    float *snap_dev = nullptr;     // packed snapshot fields (bt_snapshot.hip)
    float *snap_host = nullptr;    // pinned copy of snap_dev
    size_t snap_n = 0;             // floats in snap_dev / snap_host
    size_t f3_halo_n = 0;          // doubles in f3_halo
    long long *cnt = nullptr;
    // fused drop-in of barotropic_dynamics (bt_dropin.hip)
    ibt::DropinAlloc *da = nullptr;
    int dropin_path = 0;           // 0 = fused where eligible, 1 = stage-by-stage, 2 = fused (error if not eligible)
    bool dep_stale = false;        // dep[] no longer hh_init(st[4], st[5]) (a fused drop-in call changed st)
    double *st_save[10] = {};      // the context's state and averages kept across a stage-by-stage device-address call
    double *st_snap[6] = {};       // inmsom_bt_save_state's copy of the state
    bool snap_ok = false;          // st_snap holds a saved state
};
