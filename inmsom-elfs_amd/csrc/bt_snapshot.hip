// Snapshot of the external-mode state in the reference's output formats
// (Control/output.f90:115-154 for SSH, :155-190 and :230-260 for the
// velocities on the model grid): DIOGIN float32 records with land points
// as undef, one record per snapshot, and a GrADS descriptor per file.
//
// The fields are packed on the device (FP64 state -> float32 over the
// output box, undef where the mask is land), copied once into pinned host
// memory and written as records.
#include <string>

#include "bt_fused.h"
#include "bt_internal.h"
#include "bt_io.h"

namespace ibt {

namespace {

constexpr float SNAP_UNDEF = -1.0e32f;  // Inc/reclen.fi: undef

// one output box: dst[(n - n1) * rx + (m - m1)] = float(src(m, n)), or
// undef where |mask| < 0.5 (wdstd, Service/input_output_data.f90:268-283)
__global__ void k_pack_box(const double *__restrict__ src, const float *__restrict__ mask, int nx, int m1, int rx,
                           int n1, int ry, float undef, float *__restrict__ dst)
{
    const int q = blockIdx.x * blockDim.x + threadIdx.x;
    if (q >= rx * ry) return;
    const int i = q % rx, j = q / rx;
    const size_t a = (size_t)(n1 - 1 + j) * nx + (size_t)(m1 - 1 + i);
    dst[q] = fabsf(mask[a]) < 0.5f ? undef : (float)src[a];
}

struct Box {
    const char *file;
    int state;  // index into c->st
    int mask;   // index into c->mask
    int m1, m2, n1, n2;
    double dx0, dy0;  // axis origin offset in steps
    const char *title, *var;
};

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

}  // namespace

}  // namespace ibt

using namespace ibt;

extern "C" int inmsom_bt_write_snapshot(inmsom_bt_ctx *c, const char *path, int nrec, double xt0, double dxst,
                                        double yt0, double dyst, int calendar, int year, int month, int day, int hour,
                                        int minute, float tstep, double *ms)
{
    if (!c || !path) return sfail("null pointer", INMSOM_BT_EINVAL);
    if (nrec < 1) return sfail("nrec must be >= 1", INMSOM_BT_EINVAL);
    if (c->nx < 5 || c->ny < 5) return sfail("the snapshot needs nx, ny >= 5", INMSOM_BT_EINVAL);
    const int nx = c->nx, ny = c->ny;
    // masks: lu for SSH on T points, llu / llv for the velocities (the
    // reference writes uu on m1loc-1..m2loc with llu, vv on n1loc-1..n2loc
    // with llv); mask order lu, luu, luh, lcu, lcv, llu, llv
    const Box boxes[3] = {
        {"ssh.dat", 4, 0, 3, nx - 2, 3, ny - 2, 0.0, 0.0, "SSH, m", "ssh"},
        {"ubrtr.dat", 0, 5, 2, nx - 2, 3, ny - 2, -0.5, 0.0, "zonal barotropic transport, m2/s", "ubrtr"},
        {"vbrtr.dat", 2, 6, 3, nx - 2, 2, ny - 2, 0.0, -0.5, "meridional barotropic transport, m2/s", "vbrtr"},
    };
    size_t off[4] = {0, 0, 0, 0};
    for (int b = 0; b < 3; ++b)
        off[b + 1] = off[b] + (size_t)(boxes[b].m2 - boxes[b].m1 + 1) * (boxes[b].n2 - boxes[b].n1 + 1);
    CHS(hipSetDevice(c->device));
    int rc = fused_flush(c);
    if (rc != INMSOM_BT_OK) return rc;
    if (!c->snap_dev || c->snap_n < off[3]) {
        if (c->snap_dev) CHS(hipFree(c->snap_dev));
        if (c->snap_host) CHS(hipHostFree(c->snap_host));
        c->snap_dev = nullptr;
        c->snap_host = nullptr;
        c->snap_n = 0;
        CHS(hipMalloc(&c->snap_dev, off[3] * sizeof(float)));
        CHS(hipHostMalloc(&c->snap_host, off[3] * sizeof(float), hipHostMallocDefault));
        c->snap_n = off[3];
    }
    hipEvent_t e0, e1;
    CHS(hipEventCreate(&e0));
    CHS(hipEventCreate(&e1));
    CHS(hipEventRecord(e0, c->stream));
    for (int b = 0; b < 3; ++b) {
        const Box &x = boxes[b];
        const int rx = x.m2 - x.m1 + 1, ry = x.n2 - x.n1 + 1;
        hipLaunchKernelGGL(k_pack_box, dim3((unsigned)((rx * ry + 255) / 256)), dim3(256), 0, c->stream,
                           c->st[x.state], c->mask[x.mask], nx, x.m1, rx, x.n1, ry, SNAP_UNDEF, c->snap_dev + off[b]);
    }
    CHS(hipGetLastError());
    CHS(hipMemcpyAsync(c->snap_host, c->snap_dev, off[3] * sizeof(float), hipMemcpyDeviceToHost, c->stream));
    CHS(hipEventRecord(e1, c->stream));
    CHS(hipEventSynchronize(e1));
    float t = 0.0f;
    CHS(hipEventElapsedTime(&t, e0, e1));
    CHS(hipEventDestroy(e0));
    CHS(hipEventDestroy(e1));
    if (ms) *ms = t;
    for (int b = 0; b < 3; ++b) {
        const Box &x = boxes[b];
        const int rx = x.m2 - x.m1 + 1, ry = x.n2 - x.n1 + 1;
        rc = io::write_record(path, x.file, nrec, c->snap_host + off[b], (size_t)rx * ry * sizeof(float));
        if (rc != INMSOM_BT_OK) return rc;
        rc = io::write_ctl(path, x.file, SNAP_UNDEF, rx, ry, nrec, xt0 + x.dx0 * dxst, dxst, yt0 + x.dy0 * dyst,
                           dyst, calendar, year, month, day, hour, minute, tstep, x.title, x.var);
        if (rc != INMSOM_BT_OK) return rc;
    }
    return INMSOM_BT_OK;
}
