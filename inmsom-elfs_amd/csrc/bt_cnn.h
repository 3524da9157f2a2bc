// Internal interface of the CNN closure shared by the training object
// (bt_cnn.hip) and the batched ensemble step (bt_ensemble.hip).
#pragma once

#include <hip/hip_runtime.h>

#include <cstddef>
#include <cstdint>

#include "inmsom_bt.h"

namespace ibt {
namespace cnn {

constexpr int NL = 4;    // layers
constexpr int CH = 32;   // hidden (and padded output) channels
constexpr int CIN0 = 8;  // input channels, 4 used (u, v, eta, depth)
constexpr int NIN = 4;
constexpr int NOUT = 2;  // output channels used (u and v tendencies)

// One image of a (possibly ragged) batch: rows H, columns W, first pixel.
struct Img {
    int H, W;
    long long pix;
};

// Input normalisation x_c = (field_c - shift_c) * scale_c for (u, v, eta,
// depth); the network's outputs times oscale are the u and v tendencies
// [m/s^2] (kind OUT_UV), or output 0 times oscale[0] is a stream-function
// tendency psi on H points whose differences give them (kind OUT_PSI):
//   du(m, n) = -(psi(m, n) - psi(m, n-1)),  dv(m, n) = psi(m, n) - psi(m-1, n)
// (divergence-free; the grid step is folded into oscale).
enum { OUT_UV = 0, OUT_PSI = 1 };
struct Norm {
    float shift[NIN], scale[NIN], oscale[NOUT];
    int kind;
};

// u and v tendencies (normalised units) at pixel (x, y) of an H x W image
// whose channel-0 output is psi at stride `st` floats per pixel.
__host__ __device__ inline void psi_uv(const float *psi, int st, int H, int W, int x, int y, float *du, float *dv)
{
    (void)H;
    const float c = psi[((long long)y * W + x) * st];
    const float s = y > 0 ? psi[((long long)(y - 1) * W + x) * st] : 0.0f;
    const float w = x > 0 ? psi[((long long)y * W + x - 1) * st] : 0.0f;
    *du = -(c - s);
    *dv = c - w;
}

// One basin the closure acts on.  u, v, s, H are the basin's (ny, nx)
// arrays (float or double as the owning step stores them); the closure
// image of the basin is the whole array.
struct ClosureMember {
    int nx, ny;
    long long pix;  // first pixel in the packed activations
    long long row0; // first row in the per-row partial sums
    void *u, *v;
    const void *s, *H;
    const uint8_t *mb;  // mask bits (bit 1: U point, bit 2: V point)
    double d, nu;       // grid step [m], lateral viscosity [m^2/s]
};

// Device buffers of batched inference over a set of basins.
struct Infer {
    int nmem = 0;
    long long npix = 0;
    int nrows = 0, ntiles = 0;
    Img *imgs = nullptr;
    int *tile_img = nullptr, *tile_local = nullptr;
    int *row_member = nullptr, *row_y = nullptr;
    __bf16 *a0 = nullptr, *a1 = nullptr;
    float *tend = nullptr;
    double *part = nullptr;   // [nrows][2]: closure and viscous squared norms
    double *norms = nullptr;  // [nmem][2]: closure and viscous tendency norms of the last step
    int *active = nullptr;    // [nmem]: closure applied in the last step
    int *nfall = nullptr;     // [nmem]: steps that fell back to plain viscosity
};

int infer_create(int device, int nmem, const int *nx, const int *ny, Infer *inf);
void infer_destroy(Infer *inf);

// One application of the closure after an outer step of length dt: the
// network's tendency is added to u and v (dt * tendency on U/V points) of
// every member whose closure-tendency norm is at most ratio times the norm of
// its plain viscous tendency nu * Laplacian(u, v); a member above the limit
// keeps the plain step for this step (counted in nfall).  members is a device
// table of inf.nmem entries.
void closure_launch(const inmsom_bt_cnn *net, hipStream_t s, bool f32, ClosureMember *members, Infer &inf,
                    double dt, double ratio);
int net_device(const inmsom_bt_cnn *net);
// Changes whenever something that closure_launch passes by value (the
// normalisation and output kind) changes: graphs captured under an older
// version replay stale arguments and must be captured again.
unsigned net_version(const inmsom_bt_cnn *net);
// Holders of the network (ensembles that use it as their closure): on
// inmsom_bt_cnn_destroy every holder still attached is told through
// detach(owner) before the network's memory is freed.
void net_attach(inmsom_bt_cnn *net, void *owner, void (*detach)(void *owner));
void net_detach(inmsom_bt_cnn *net, void *owner);

}  // namespace cnn
}  // namespace ibt
