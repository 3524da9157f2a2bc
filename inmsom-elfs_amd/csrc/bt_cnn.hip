// Learned lateral-momentum closure: a 4-layer 3x3 CNN (4 -> 32 -> 32 -> 32 -> 2
// channels, ReLU) on the coarse C-grid, with BF16 weights and activations and
// FP32 accumulation on the gfx950 matrix cores (v_mfma_f32_32x32x16_bf16),
// trained on the GPU with hand-written forward, backward and Adam kernels.
//
// Convolutions are implicit GEMMs: one wave computes 32 consecutive pixels of
// one image row x 32 output channels; the reduction runs over K = 9 taps x
// input channels in steps of 16.  Activations are NHWC (channels innermost)
// so every lane's 8-element operand slice is one 16-byte load.
//   forward            y[p, co]  = sum_{t, ci} W[co][t][ci] x[p + o_t, ci]
//   backward (data)    dx[q, ci] = sum_{t, co} W[co][t][ci] dy[q - o_t, co]
//                      (the forward kernel with negated tap offsets and the
//                      weights transposed to [ci][t][co])
//   backward (weights) dW[co][t][ci] = sum_p dy[p, co] x[p + o_t, ci]
//                      (pixels are the reduction axis: per-chunk partial sums,
//                      then a fixed-order reduction -> deterministic)
// MFMA operand layout (checked on the device): A[i][k] with i = lane % 32,
// k = 8 (lane / 32) + e; B[k][j] with j = lane % 32; D[i][j] in acc[r] with
// i = 8 (r / 4) + 4 (lane / 32) + r % 4, j = lane % 32.
#include <hip/hip_bf16.h>

#include <cmath>
#include <type_traits>
#include <cstring>
#include <string>
#include <vector>

#include "bt_cnn.h"
#include "bt_internal.h"

typedef __bf16 bf16;
typedef __attribute__((__vector_size__(8 * sizeof(__bf16)))) __bf16 bf16x8;
typedef __attribute__((__vector_size__(16 * sizeof(float)))) float f32x16;
typedef __attribute__((__vector_size__(4 * sizeof(__bf16)))) __bf16 bf16x4;
typedef __attribute__((__vector_size__(2 * sizeof(float)))) float f32x2;

namespace ibt {
namespace cnn {

struct Layer {
    int cin;        // padded input channels (8 or 32)
    int kpad;       // 9*cin rounded up to 16
    size_t woff;    // offset of W[32][kpad] in the parameter vector
    size_t boff;    // offset of b[32]
};

enum { EPI_RELU_BF16 = 0, EPI_OUT_F32 = 1, EPI_MASK_BF16 = 2 };

// Rows of one block tile (RW rows per wave, 4 waves) and its column width.
constexpr int RW = 2;
constexpr int TR = 4 * RW;
// Inference batches of at most this many tiles run the four layers in one
// pass (k_closure4) unless the network's fuse setting says otherwise.
constexpr int FUSED_MAX_TILES = 4096;
constexpr int TW = 32;

// LDS pixel stride (bf16 elements) of a staged input tile: 32-channel
// pixels are padded to 80 bytes so that the 16-byte reads of 8 consecutive
// lanes (8 consecutive pixels) fall on distinct banks.
template <int CIN>
constexpr int lds_stride() { return CIN == 8 ? 8 : CIN + 8; }

// One 3x3 convolution layer over a batch of images, NHWC BF16 in.  A block
// (4 waves) owns a tile of TR rows x 32 columns of one image: it stages the
// (TR + 2) x 34 input halo tile in LDS, and each wave multiplies RW rows of
// 32 pixels by the 32 output channels with its weight fragments held in
// registers for the whole (persistent) launch.
//   forward:   out = relu(z + b) in BF16 (EPI_RELU_BF16) or z + b, channels
//              0..1 in FP32 (EPI_OUT_F32); w = W[co][kpad]
//   backward:  NEG taps, w = W^T[ci][t*32 + co] (kpad 288), out = dx masked
//              by act_prev > 0 in BF16 (EPI_MASK_BF16)
template <int CIN, bool NEG, int EPI>
__global__ void __launch_bounds__(256)
    k_conv(const Img *__restrict__ imgs, const int *__restrict__ tile_img,
           const int *__restrict__ tile_local, int ntiles, const bf16 *__restrict__ in,
           const bf16 *__restrict__ w, const float *__restrict__ bias,
           const bf16 *__restrict__ act_prev, void *__restrict__ out)
{
    constexpr int KPAD = NEG ? 9 * CH : (9 * CIN + 15) / 16 * 16;
    constexpr int KS = KPAD / 16;
    constexpr int LS = lds_stride<CIN>();
    constexpr int CPP = CIN / 8;                 // 16-byte chunks per pixel
    constexpr int NCHUNK = (TR + 2) * (TW + 2) * CPP;
    __shared__ __attribute__((aligned(16))) bf16 tile_in[(TR + 2) * (TW + 2) * LS];

    const int wv = threadIdx.x >> 6, lane = threadIdx.x & 63, i = lane & 31, half = lane >> 5;
    bf16x8 b[KS];
#pragma unroll
    for (int s = 0; s < KS; ++s) b[s] = *reinterpret_cast<const bf16x8 *>(w + (size_t)i * KPAD + 16 * s + 8 * half);
    // bias of the lane's 16 output channels (8 (r / 4) + 4 half + r % 4)
    float bco[16];
#pragma unroll
    for (int r = 0; r < 16; ++r) bco[r] = EPI == EPI_MASK_BF16 ? 0.0f : bias[8 * (r >> 2) + 4 * half + (r & 3)];

    // Staging is software-pipelined: the global loads of the block's next
    // tile are issued into registers before the current tile's matrix
    // products and committed to LDS after them.
    constexpr int NPT = (NCHUNK + 255) / 256;  // staged chunks per thread
    bf16x8 raw[NPT];
    unsigned valid = 0;
    auto fetch = [&](int t) {
        const Img im = imgs[tile_img[t]];
        const int tl = tile_local[t];
        const int tx = (im.W + TW - 1) / TW;
        const int ty = tl / tx;
        const int y0 = ty * TR, x0 = (tl - ty * tx) * TW;
        valid = 0;
#pragma unroll
        for (int j = 0; j < NPT; ++j) {
            const int c = threadIdx.x + 256 * j;
            const int px = c / CPP, part = c - px * CPP;
            const int ry = px / (TW + 2), rx = px - ry * (TW + 2);
            const int y = y0 - 1 + ry, x = x0 - 1 + rx;
            if (c < NCHUNK && y >= 0 && y < im.H && x >= 0 && x < im.W) {
                valid |= 1u << j;
                raw[j] = *reinterpret_cast<const bf16x8 *>(in + ((im.pix + (long long)y * im.W + x) * CIN + part * 8));
            }
        }
    };
    auto commit = [&]() {
#pragma unroll
        for (int j = 0; j < NPT; ++j) {
            const int c = threadIdx.x + 256 * j;
            if (c >= NCHUNK) continue;
            const int px = c / CPP, part = c - px * CPP;
            bf16x8 v = {};
            if (valid & (1u << j)) v = raw[j];
            *reinterpret_cast<bf16x8 *>(tile_in + px * LS + part * 8) = v;
        }
    };
    if (blockIdx.x < (unsigned)ntiles) fetch(blockIdx.x);
    for (int tile = blockIdx.x; tile < ntiles; tile += gridDim.x) {
        const Img im = imgs[tile_img[tile]];
        const int tl = tile_local[tile];
        const int tx = (im.W + TW - 1) / TW;
        const int ty = tl / tx;
        const int y0 = ty * TR, x0 = (tl - ty * tx) * TW;
        __syncthreads();  // the previous tile's reads of the LDS tile are done
        commit();
        __syncthreads();
        if (tile + (int)gridDim.x < ntiles) fetch(tile + gridDim.x);
        f32x16 acc[RW];
#pragma unroll
        for (int q = 0; q < RW; ++q) acc[q] = f32x16{};
#pragma unroll
        for (int s = 0; s < KS; ++s) {
            const int k0 = 16 * s + 8 * half;
            const int t = k0 / CIN, ci0 = k0 - t * CIN;
            int oy = t / 3 - 1, ox = t % 3 - 1;
            if (NEG) {
                oy = -oy;
                ox = -ox;
            }
#pragma unroll
            for (int q = 0; q < RW; ++q) {
                const int r = wv * RW + q;
                bf16x8 a = {};
                if (t < 9) a = *reinterpret_cast<const bf16x8 *>(tile_in + ((r + 1 + oy) * (TW + 2) + (i + 1 + ox)) * LS + ci0);
                // A = weights (rows = output channels), B = pixels: D[co][px], so that
                // each lane ends with 16 channels of one pixel (4 runs of 4)
                acc[q] = __builtin_amdgcn_mfma_f32_32x32x16_bf16(b[s], a, acc[q], 0, 0, 0);
            }
        }
        // lane: pixel x0 + i, channels 8g + 4 half + 0..3 in acc[.][4g .. 4g+3]
        const int px = x0 + i;
#pragma unroll
        for (int q = 0; q < RW; ++q) {
            const int y = y0 + wv * RW + q;
            if (y >= im.H || px >= im.W) continue;
            const long long p = im.pix + (long long)y * im.W + px;
            if constexpr (EPI == EPI_OUT_F32) {
                if (half == 0) {
                    f32x2 o = {acc[q][0] + bco[0], acc[q][1] + bco[1]};
                    *reinterpret_cast<f32x2 *>(static_cast<float *>(out) + p * NOUT) = o;
                }
            } else {
#pragma unroll
                for (int g = 0; g < 4; ++g) {
                    const int co0 = 8 * g + 4 * half;
                    bf16x4 o;
                    if constexpr (EPI == EPI_RELU_BF16) {
#pragma unroll
                        for (int e = 0; e < 4; ++e) o[e] = (bf16)fmaxf(acc[q][4 * g + e] + bco[4 * g + e], 0.0f);
                    } else {
                        const bf16x4 ap = *reinterpret_cast<const bf16x4 *>(act_prev + p * CH + co0);
#pragma unroll
                        for (int e = 0; e < 4; ++e) o[e] = (bf16)((float)ap[e] > 0.0f ? acc[q][4 * g + e] : 0.0f);
                    }
                    *reinterpret_cast<bf16x4 *>(static_cast<bf16 *>(out) + p * CH + co0) = o;
                }
            }
        }
    }
}

// Layers 0 and 1 of the closure in one pass (inference).  For each tile of
// layer 1 (TR x 32 outputs) the block stages the members' normalised
// (u, v, eta, depth) on the (TR + 4) x 36 pixels the two layers reach,
// computes layer 0's activations on layer 1's (TR + 2) x 34 input pixels
// into LDS (zero outside the image, as layer 1's padding), then layer 1
// from LDS.  Same products in the same order and the same rounding as two
// separate launches, so the activations are bit-identical, without layer
// 0's launch and its activation round trip through memory.  (Fusing layers
// 2 and 3 the same way was measured slower: layer 2's 1.33x recompute on
// the halo costs more than the launch it saves.)
template <int FT>
__global__ void __launch_bounds__(256)
    k_conv01(const Img *__restrict__ imgs, const int *__restrict__ tile_img, const int *__restrict__ tile_local,
             int ntiles, const bf16 *__restrict__ w0, const float *__restrict__ bias0, const bf16 *__restrict__ w1,
             const float *__restrict__ bias1, bf16 *__restrict__ out, const ClosureMember *__restrict__ mem, Norm nm)
{
    using FT_T = typename std::conditional<FT == 2, double, float>::type;
    constexpr int LSS = lds_stride<CIN0>();
    constexpr int KS0 = (9 * CIN0 + 15) / 16, KS1 = 9 * CH / 16;
    constexpr int LS = lds_stride<CH>();
    constexpr int FH = TR + 4, FW = TW + 4;            // staged feature pixels
    constexpr int AH = TR + 2, AW = TW + 2;            // layer-0 outputs = layer-1 inputs
    constexpr int NA = AH * AW, NBLK0 = (NA + 31) / 32;
    __shared__ __attribute__((aligned(16))) bf16 feat[FH * FW * LSS];
    __shared__ __attribute__((aligned(16))) bf16 tile_in[AH * AW * LS];

    const int wv = threadIdx.x >> 6, lane = threadIdx.x & 63, i = lane & 31, half = lane >> 5;
    bf16x8 b0[KS0], b1[KS1];
#pragma unroll
    for (int s = 0; s < KS0; ++s)
        b0[s] = *reinterpret_cast<const bf16x8 *>(w0 + (size_t)i * (16 * KS0) + 16 * s + 8 * half);
#pragma unroll
    for (int s = 0; s < KS1; ++s)
        b1[s] = *reinterpret_cast<const bf16x8 *>(w1 + (size_t)i * (16 * KS1) + 16 * s + 8 * half);
    float bc0[16], bc1[16];
#pragma unroll
    for (int r = 0; r < 16; ++r) {
        bc0[r] = bias0[8 * (r >> 2) + 4 * half + (r & 3)];
        bc1[r] = bias1[8 * (r >> 2) + 4 * half + (r & 3)];
    }

    for (int tile = blockIdx.x; tile < ntiles; tile += gridDim.x) {
        const Img im = imgs[tile_img[tile]];
        const int tl = tile_local[tile];
        const int tx = (im.W + TW - 1) / TW;
        const int ty = tl / tx;
        const int y0 = ty * TR, x0 = (tl - ty * tx) * TW;
        __syncthreads();  // the previous tile's reads of both LDS tiles are done
        // features of pixels (y0 - 2 + fy, x0 - 2 + fx)
        const ClosureMember &m = mem[tile_img[tile]];
        for (int q = threadIdx.x; q < FH * FW; q += 256) {
            const int fy = q / FW, fx = q - fy * FW;
            const int y = y0 - 2 + fy, x = x0 - 2 + fx;
            bf16x8 v = {};
            if (y >= 0 && y < im.H && x >= 0 && x < im.W) {
                const long long a = (long long)y * m.nx + x;
                const FT_T f[NIN] = {static_cast<const FT_T *>(m.u)[a], static_cast<const FT_T *>(m.v)[a],
                                     static_cast<const FT_T *>(m.s)[a], static_cast<const FT_T *>(m.H)[a]};
#pragma unroll
                for (int c = 0; c < NIN; ++c) v[c] = (bf16)(((float)f[c] - nm.shift[c]) * nm.scale[c]);
            }
            *reinterpret_cast<bf16x8 *>(feat + q * LSS) = v;
        }
        __syncthreads();
        // layer 0 on the AH x AW pixels (y0 - 1 + ry, x0 - 1 + rx), 32 per block of products
        for (int j = wv; j < NBLK0; j += 4) {
            const int p = 32 * j + i;
            const int ry = p / AW, rx = p - ry * AW;
            f32x16 acc = f32x16{};
#pragma unroll
            for (int s = 0; s < KS0; ++s) {
                const int k0 = 16 * s + 8 * half;
                const int t = k0 / CIN0;
                bf16x8 a = {};
                if (t < 9 && p < NA) {
                    const int oy = t / 3 - 1, ox = t % 3 - 1;
                    a = *reinterpret_cast<const bf16x8 *>(feat + ((ry + 1 + oy) * FW + (rx + 1 + ox)) * LSS);
                }
                acc = __builtin_amdgcn_mfma_f32_32x32x16_bf16(b0[s], a, acc, 0, 0, 0);
            }
            if (p < NA) {
                const int y = y0 - 1 + ry, x = x0 - 1 + rx;
                const bool in_img = y >= 0 && y < im.H && x >= 0 && x < im.W;
#pragma unroll
                for (int g = 0; g < 4; ++g) {
                    bf16x4 o;
#pragma unroll
                    for (int e = 0; e < 4; ++e)
                        o[e] = in_img ? (bf16)fmaxf(acc[4 * g + e] + bc0[4 * g + e], 0.0f) : (bf16)0.0f;
                    *reinterpret_cast<bf16x4 *>(tile_in + p * LS + 8 * g + 4 * half) = o;
                }
            }
        }
        __syncthreads();
        // layer 1 from LDS, as k_conv
        f32x16 acc[RW];
#pragma unroll
        for (int q = 0; q < RW; ++q) acc[q] = f32x16{};
#pragma unroll
        for (int s = 0; s < KS1; ++s) {
            const int k0 = 16 * s + 8 * half;
            const int t = k0 / CH, ci0 = k0 - t * CH;
            const int oy = t / 3 - 1, ox = t % 3 - 1;
#pragma unroll
            for (int q = 0; q < RW; ++q) {
                const int r = wv * RW + q;
                const bf16x8 a =
                    *reinterpret_cast<const bf16x8 *>(tile_in + ((r + 1 + oy) * AW + (i + 1 + ox)) * LS + ci0);
                acc[q] = __builtin_amdgcn_mfma_f32_32x32x16_bf16(b1[s], a, acc[q], 0, 0, 0);
            }
        }
        const int px = x0 + i;
#pragma unroll
        for (int q = 0; q < RW; ++q) {
            const int y = y0 + wv * RW + q;
            if (y >= im.H || px >= im.W) continue;
            const long long pp = im.pix + (long long)y * im.W + px;
#pragma unroll
            for (int g = 0; g < 4; ++g) {
                const int co0 = 8 * g + 4 * half;
                bf16x4 o;
#pragma unroll
                for (int e = 0; e < 4; ++e) o[e] = (bf16)fmaxf(acc[q][4 * g + e] + bc1[4 * g + e], 0.0f);
                *reinterpret_cast<bf16x4 *>(out + pp * CH + co0) = o;
            }
        }
    }
}

// One layer of the fused inference on a region of Ho x Wo pixels whose
// top-left pixel is (oy0, ox0) of the image: its 3 x 3 inputs are the
// (Ho + 2) x (Wo + 2) region in LDS `in` (pixel stride LIN), its outputs
// go to LDS `outl` (pixel stride LS, ReLU, zero outside the image, as the
// next layer's padding) or, for the output layer, to memory.  The region
// is walked in blocks of 32 consecutive pixels of its row-major order (one
// MFMA column each, so a region need not be 32 wide); per pixel the
// products and their order are k_conv's, so the results are the same bits.
template <int CIN, int KS, int LIN, bool LAST>
__device__ __forceinline__ void fused_layer(const bf16 *__restrict__ in, int Ho, int Wo, const bf16x8 (&b)[KS],
                                            const float (&bc)[16], bf16 *__restrict__ outl,
                                            float *__restrict__ outg, const Img &im, int oy0, int ox0, int wv,
                                            int i, int half)
{
    constexpr int LS = lds_stride<CH>();
    const int np = Ho * Wo, nb = (np + 31) / 32, Wi = Wo + 2;
    for (int j = wv; j < nb; j += 4) {
        const int p = 32 * j + i;
        const int ry = p / Wo, rx = p - ry * Wo;
        f32x16 acc = f32x16{};
#pragma unroll
        for (int s = 0; s < KS; ++s) {
            const int k0 = 16 * s + 8 * half;
            const int t = k0 / CIN, ci0 = k0 - t * CIN;
            bf16x8 a = {};
            if (t < 9 && p < np) {
                const int oy = t / 3 - 1, ox = t % 3 - 1;
                a = *reinterpret_cast<const bf16x8 *>(in + ((ry + 1 + oy) * Wi + (rx + 1 + ox)) * LIN + ci0);
            }
            acc = __builtin_amdgcn_mfma_f32_32x32x16_bf16(b[s], a, acc, 0, 0, 0);
        }
        if (p >= np) continue;
        const int y = oy0 + ry, x = ox0 + rx;
        const bool in_img = y >= 0 && y < im.H && x >= 0 && x < im.W;
        if constexpr (LAST) {
            if (in_img && half == 0) {
                f32x2 o = {acc[0] + bc[0], acc[1] + bc[1]};
                *reinterpret_cast<f32x2 *>(outg + (im.pix + (long long)y * im.W + x) * NOUT) = o;
            }
        } else {
#pragma unroll
            for (int g = 0; g < 4; ++g) {
                bf16x4 o;
#pragma unroll
                for (int e = 0; e < 4; ++e)
                    o[e] = in_img ? (bf16)fmaxf(acc[4 * g + e] + bc[4 * g + e], 0.0f) : (bf16)0.0f;
                *reinterpret_cast<bf16x4 *>(outl + p * LS + 8 * g + 4 * half) = o;
            }
        }
    }
}

// The pointer, hidden from the optimiser: weight loads through it stay where
// they are written (inside the tile loop) instead of being hoisted out of
// it, where all four layers' weights would have to stay in registers.
template <class T>
__device__ __forceinline__ const T *opaque(const T *p)
{
    asm volatile("" : "+s"(p));
    return p;
}

template <int KS>
__device__ __forceinline__ void fused_weights(const bf16 *__restrict__ w, const float *__restrict__ bias,
                                              bf16x8 (&b)[KS], float (&bc)[16], int i, int half)
{
#pragma unroll
    for (int s = 0; s < KS; ++s) b[s] = *reinterpret_cast<const bf16x8 *>(w + (size_t)i * (16 * KS) + 16 * s + 8 * half);
#pragma unroll
    for (int r = 0; r < 16; ++r) bc[r] = bias[8 * (r >> 2) + 4 * half + (r & 3)];
}

// All four layers of the closure in one pass (inference), for batches too
// small to fill the chip with one layer at a time, where each launch costs
// more than its products.  For each TR x 32 tile of outputs the block
// stages the members' normalised (u, v, eta, depth) on the (TR + 8) x 40
// pixels the four layers reach, then forms layer 0 on (TR + 6) x 38
// pixels, layer 1 on (TR + 4) x 36, layer 2 on (TR + 2) x 34 in LDS and
// layer 3 on the tile, written as the tendency.  Region A holds layer 0's
// outputs and later layer 2's, region B the features and later layer 1's
// outputs: 77 KB, two blocks per CU.  About 1.44 x the products of the
// layer-by-layer launches (the halo rings), the same activations bit for
// bit.
template <int FT>
__global__ void __launch_bounds__(256)
    k_closure4(const Img *__restrict__ imgs, const int *__restrict__ tile_img, const int *__restrict__ tile_local,
               int ntiles, const bf16 *__restrict__ w0, const float *__restrict__ bias0, const bf16 *__restrict__ w1,
               const float *__restrict__ bias1, const bf16 *__restrict__ w2, const float *__restrict__ bias2,
               const bf16 *__restrict__ w3, const float *__restrict__ bias3, float *__restrict__ tend,
               const ClosureMember *__restrict__ mem, Norm nm)
{
    using FT_T = typename std::conditional<FT == 2, double, float>::type;
    constexpr int LSS = lds_stride<CIN0>(), LS = lds_stride<CH>();
    constexpr int KS0 = (9 * CIN0 + 15) / 16, KS1 = 9 * CH / 16;
    constexpr int FH = TR + 8, FW = TW + 8;
    constexpr int SZA = (TR + 6) * (TW + 6) * LS;
    constexpr int SZB = (TR + 4) * (TW + 4) * LS > FH * FW * LSS ? (TR + 4) * (TW + 4) * LS : FH * FW * LSS;
    __shared__ __attribute__((aligned(16))) bf16 lA[SZA];
    __shared__ __attribute__((aligned(16))) bf16 lB[SZB];

    const int wv = threadIdx.x >> 6, lane = threadIdx.x & 63, i = lane & 31, half = lane >> 5;
    bf16x8 b0[KS0], b[KS1];
    float bc0[16], bc[16];
    fused_weights<KS0>(w0, bias0, b0, bc0, i, half);

    for (int tile = blockIdx.x; tile < ntiles; tile += gridDim.x) {
        const Img im = imgs[tile_img[tile]];
        const int tl = tile_local[tile];
        const int tx = (im.W + TW - 1) / TW;
        const int ty = tl / tx;
        const int y0 = ty * TR, x0 = (tl - ty * tx) * TW;
        __syncthreads();  // the previous tile's reads of both regions are done
        // features of pixels (y0 - 4 + fy, x0 - 4 + fx) into region B
        const ClosureMember &m = mem[tile_img[tile]];
        for (int q = threadIdx.x; q < FH * FW; q += 256) {
            const int fy = q / FW, fx = q - fy * FW;
            const int y = y0 - 4 + fy, x = x0 - 4 + fx;
            bf16x8 v = {};
            if (y >= 0 && y < im.H && x >= 0 && x < im.W) {
                const long long a = (long long)y * m.nx + x;
                const FT_T f[NIN] = {static_cast<const FT_T *>(m.u)[a], static_cast<const FT_T *>(m.v)[a],
                                     static_cast<const FT_T *>(m.s)[a], static_cast<const FT_T *>(m.H)[a]};
#pragma unroll
                for (int c = 0; c < NIN; ++c) v[c] = (bf16)(((float)f[c] - nm.shift[c]) * nm.scale[c]);
            }
            *reinterpret_cast<bf16x8 *>(lB + q * LSS) = v;
        }
        fused_weights<KS1>(opaque(w1), opaque(bias1), b, bc, i, half);
        __syncthreads();
        fused_layer<CIN0, KS0, LSS, false>(lB, TR + 6, TW + 6, b0, bc0, lA, nullptr, im, y0 - 3, x0 - 3, wv, i,
                                           half);
        __syncthreads();
        fused_layer<CH, KS1, LS, false>(lA, TR + 4, TW + 4, b, bc, lB, nullptr, im, y0 - 2, x0 - 2, wv, i, half);
        fused_weights<KS1>(opaque(w2), opaque(bias2), b, bc, i, half);
        __syncthreads();
        fused_layer<CH, KS1, LS, false>(lB, TR + 2, TW + 2, b, bc, lA, nullptr, im, y0 - 1, x0 - 1, wv, i, half);
        fused_weights<KS1>(opaque(w3), opaque(bias3), b, bc, i, half);
        __syncthreads();
        fused_layer<CH, KS1, LS, true>(lA, TR, TW, b, bc, nullptr, tend, im, y0, x0, wv, i, half);
    }
}

// Partial weight and bias gradients (uniform batch: B images of H x W).
// dW[co][t][ci] = sum_p dy[p, co] x[p + o_t, ci] reduces over pixels, so
// the MFMA's k axis is pixels and each lane's operand is 8 pixels of one
// channel.  Rows of x and dy are staged in LDS as they lie in memory
// (pixel-major, 16-byte loads and stores) and transposed on the way into
// the operands, 2-byte LDS reads: operand element e of lane half h is
// pixel 2 e + h of the 16 (the same pairing in A and B, so the product
// still sums the 16 pixels), which puts the two halves' reads of one
// element 64 bytes apart, on different banks.
// A block of three waves walks work items of WG_R rows x WG_TW columns of
// one image; wave ky multiplies the row's dy (A = dy[px][co]) by row
// y + ky - 1 of x at the three column shifts kx (B = x[px + kx - 1][ci]):
// for CIN = 32 one accumulator per tap (its 32 x 32 block of dW), for
// CIN = 8 one accumulator holding the three taps of its row (column
// 8 kx + ci).  x rows are kept in a ring of four LDS slots (rows y - 1 ..
// y + 1 in use while row y + 2 is staged).  The next row's global loads are
// issued before the current row's matrix products and written to LDS after
// them; one barrier per row.  Blocks take consecutive runs of work items (a
// fixed assignment: the record of block k is always the same sum, so the
// gradient is deterministic).
constexpr int WG_TW = 64;            // columns of a work item
constexpr int WG_R = 16;             // rows of a work item
constexpr int WG_SLOTS = 4;          // x rows in the ring

template <int CIN>
struct WgradLds {
    bf16 xs[WG_SLOTS][WG_TW + 2][CIN];  // x row: [p][ci] = x[x0 - 1 + p][ci]
    bf16 dys[2][WG_TW][CH];             // dy row: [p][co] = dy[x0 + p][co]
};

template <int CIN>
__global__ void __launch_bounds__(3 * 64)
    k_wgrad(int H, int W, int nitems, int per, const bf16 *__restrict__ dy, const bf16 *__restrict__ x,
            float *__restrict__ part)
{
    constexpr int NACC = CIN == CH ? 3 : 1;
    constexpr int XCPP = CIN / 8, YCPP = CH / 8;                 // 16-byte chunks per pixel
    constexpr int XCH = (WG_TW + 2) * XCPP, YCH = WG_TW * YCPP;  // chunks per staged row
    constexpr int XPT = (XCH + 191) / 192, YPT = (YCH + 191) / 192;
    __shared__ __attribute__((aligned(16))) WgradLds<CIN> L;

    const int wv = __builtin_amdgcn_readfirstlane((int)(threadIdx.x >> 6));  // ky
    const int lane = threadIdx.x & 63, i = lane & 31, half = lane >> 5;
    const int nsx = (W + WG_TW - 1) / WG_TW, nry = (H + WG_R - 1) / WG_R;
    f32x16 acc[NACC];
#pragma unroll
    for (int a = 0; a < NACC; ++a) acc[a] = f32x16{};
    float bsum = 0.0f;

    bf16x8 rx[XPT], ry[YPT];
    // global loads of x row yy and dy row yd of image b, strip x0 (zero
    // outside the image; yd < 0: no dy row)
    auto fetch = [&](int b, int x0, int yy, int yd) {
#pragma unroll
        for (int j = 0; j < XPT; ++j) {
            const int c = threadIdx.x + 192 * j;
            const int px = c / XCPP, prt = c - px * XCPP, xx = x0 - 1 + px;
            rx[j] = bf16x8{};
            if (c < XCH && yy >= 0 && yy < H && xx >= 0 && xx < W)
                rx[j] = *reinterpret_cast<const bf16x8 *>(x + (((long long)b * H + yy) * W + xx) * CIN + prt * 8);
        }
#pragma unroll
        for (int j = 0; j < YPT; ++j) {
            const int c = threadIdx.x + 192 * j;
            const int px = c / YCPP, prt = c - px * YCPP, xx = x0 + px;
            ry[j] = bf16x8{};
            if (c < YCH && yd >= 0 && xx < W)
                ry[j] = *reinterpret_cast<const bf16x8 *>(dy + (((long long)b * H + yd) * W + xx) * CH + prt * 8);
        }
    };
    // the fetched rows into x slot xsl and dy buffer dsl (dsl < 0: none)
    auto commit = [&](int xsl, int dsl) {
#pragma unroll
        for (int j = 0; j < XPT; ++j) {
            const int c = threadIdx.x + 192 * j;
            const int px = c / XCPP, prt = c - px * XCPP;
            if (c < XCH) *reinterpret_cast<bf16x8 *>(&L.xs[xsl][px][prt * 8]) = rx[j];
        }
        if (dsl < 0) return;
#pragma unroll
        for (int j = 0; j < YPT; ++j) {
            const int c = threadIdx.x + 192 * j;
            const int px = c / YCPP, prt = c - px * YCPP;
            if (c < YCH) *reinterpret_cast<bf16x8 *>(&L.dys[dsl][px][prt * 8]) = ry[j];
        }
    };

    const int it0 = blockIdx.x * per, it1 = min(it0 + per, nitems);
    for (int it = it0; it < it1; ++it) {
        const int b = it / (nsx * nry), rem = it - b * nsx * nry;
        const int ry0 = rem / nsx, sx = rem - ry0 * nsx;
        const int x0 = sx * WG_TW, y0 = ry0 * WG_R, y1 = min(y0 + WG_R, H);
        // prime: x rows y0 - 1, y0, y0 + 1 and dy row y0
        fetch(b, x0, y0 - 1, -1);
        commit((y0 - 1 + WG_SLOTS) % WG_SLOTS, -1);
        fetch(b, x0, y0, -1);
        commit(y0 % WG_SLOTS, -1);
        fetch(b, x0, y0 + 1, y0);
        commit((y0 + 1) % WG_SLOTS, y0 & 1);
        __syncthreads();
        for (int y = y0; y < y1; ++y) {
            const bool more = y + 1 < y1;
            if (more) fetch(b, x0, y + 2, y + 1);
            const int xsl = (y + wv - 1 + WG_SLOTS) % WG_SLOTS, dsl = y & 1;
#pragma unroll
            for (int ks = 0; ks < WG_TW / 16; ++ks) {
                bf16x8 a;
#pragma unroll
                for (int e = 0; e < 8; ++e) a[e] = L.dys[dsl][16 * ks + 2 * e + half][i];
                if (wv == 0) {
#pragma unroll
                    for (int e = 0; e < 8; ++e) bsum += (float)a[e];
                }
                if constexpr (CIN == CH) {
#pragma unroll
                    for (int kx = 0; kx < 3; ++kx) {
                        bf16x8 bb;
#pragma unroll
                        for (int e = 0; e < 8; ++e) bb[e] = L.xs[xsl][16 * ks + 2 * e + half + kx][i];
                        acc[kx] = __builtin_amdgcn_mfma_f32_32x32x16_bf16(a, bb, acc[kx], 0, 0, 0);
                    }
                } else {
                    // column j = 8 kx + ci (j < 24)
                    const int kx = i >> 3, ci = i & 7;
                    bf16x8 bb = bf16x8{};
                    if (i < 24) {
#pragma unroll
                        for (int e = 0; e < 8; ++e) bb[e] = L.xs[xsl][16 * ks + 2 * e + half + kx][ci];
                    }
                    acc[0] = __builtin_amdgcn_mfma_f32_32x32x16_bf16(a, bb, acc[0], 0, 0, 0);
                }
            }
            if (more) commit((y + 2) % WG_SLOTS, (y + 1) & 1);
            __syncthreads();
        }
    }
    // record of this block: W partials [32][ncol] (column t CIN + ci), then
    // the 32 bias partials
    constexpr int ncol = CIN == CH ? 9 * CH : 96;
    float *rec = part + (size_t)blockIdx.x * (32 * ncol + 32);
#pragma unroll
    for (int a = 0; a < NACC; ++a)
#pragma unroll
        for (int r = 0; r < 16; ++r) {
            const int co = 8 * (r >> 2) + 4 * half + (r & 3);
            if constexpr (CIN == CH) {
                rec[co * ncol + (3 * wv + a) * CH + i] = acc[a][r];
            } else if (i < 24) {
                rec[co * ncol + 24 * wv + i] = acc[a][r];
                if (wv == 2) rec[co * ncol + 72 + i] = 0.0f;  // padding columns 72..95
            }
        }
    if (wv == 0) {
        bsum += __shfl_xor(bsum, 32);
        if (half == 0) rec[32 * ncol + i] = bsum;
    }
}

// Fixed-order reduction of the chunk records (rec floats each) in two
// passes: groups of RG consecutive chunks in parallel, then the groups.
constexpr int RG = 64;

__global__ void k_reduce_groups(int nchunks, int rec, const float *__restrict__ part, float *__restrict__ part2)
{
    const int e = blockIdx.x * blockDim.x + threadIdx.x;
    if (e >= rec) return;
    const int c0 = blockIdx.y * RG, c1 = min(c0 + RG, nchunks);
    float s = 0.0f;
    for (int c = c0; c < c1; ++c) s += part[(size_t)c * rec + e];
    part2[(size_t)blockIdx.y * rec + e] = s;
}

// groups -> gradient: W in [32][kpad] (zero past kvalid), bias [32]
__global__ void k_reduce_wgrad(int ngroups, int ncol, int kvalid, int kpad, const float *__restrict__ part2,
                               float *__restrict__ gw, float *__restrict__ gb)
{
    const int idx = blockIdx.x * blockDim.x + threadIdx.x;
    const size_t rec = (size_t)32 * ncol + 32;
    if (idx < 32 * kpad) {
        const int co = idx / kpad, k = idx - co * kpad;
        float s = 0.0f;
        if (k < kvalid)
            for (int g = 0; g < ngroups; ++g) s += part2[g * rec + (size_t)co * ncol + k];
        gw[idx] = s;
    }
    if (idx < 32) {
        float s = 0.0f;
        for (int g = 0; g < ngroups; ++g) s += part2[g * rec + (size_t)32 * ncol + idx];
        gb[idx] = s;
    }
}

// Masked mean-square loss over U (channel 0) and V (channel 1) points of
// a batch of H x W images: writes dL/dpred into dpred (32 channels, the
// unused ones zeroed) and per-block partial sums of the squared error.
// OUT_PSI: the tendencies are the stream-function differences (psi_uv) and
// dL/dpsi is their adjoint: psi(x, y) enters du(x, y), du(x, y+1),
// dv(x, y) and dv(x+1, y).
__global__ void k_loss(long long n, int H, int W, int kind, const float *__restrict__ pred,
                       const float *__restrict__ tgt, const uint8_t *__restrict__ mask, float inv_count,
                       bf16 *__restrict__ dpred, float *__restrict__ part)
{
    __shared__ float red[256];
    const long long p = (long long)blockIdx.x * blockDim.x + threadIdx.x;
    float e2 = 0.0f;
    if (p < n) {
        const uint8_t mk = mask[p];
        // the pixel's 32 channels, written as four 16-byte stores (the
        // channels past the two outputs are zero)
        bf16x8 g0 = bf16x8{};
        if (kind == OUT_UV) {
#pragma unroll
            for (int c = 0; c < NOUT; ++c) {
                const float on = (mk >> c) & 1 ? 1.0f : 0.0f;
                const float d = (pred[p * NOUT + c] - tgt[p * NOUT + c]) * on;
                e2 += d * d;
                g0[c] = (bf16)(2.0f * d * inv_count);
            }
        } else {
            const long long hw = (long long)H * W;
            const long long img = p / hw;
            const int r = (int)(p - img * hw), y = r / W, x = r - y * W;
            const float *ps = pred + img * hw * NOUT;
            const float *tg = tgt + img * hw * NOUT;
            const uint8_t *mi = mask + img * hw;
            // residual on U / V points of pixel (xx, yy) of this image
            auto res = [&](int xx, int yy, int c) -> float {
                const long long q = (long long)yy * W + xx;
                if (!((mi[q] >> c) & 1)) return 0.0f;
                float du, dv;
                psi_uv(ps, NOUT, H, W, xx, yy, &du, &dv);
                return (c == 0 ? du : dv) - tg[q * NOUT + c];
            };
            const float ru = res(x, y, 0), rv = res(x, y, 1);
            e2 = ru * ru + rv * rv;
            const float ru_n = y + 1 < H ? res(x, y + 1, 0) : 0.0f;
            const float rv_e = x + 1 < W ? res(x + 1, y, 1) : 0.0f;
            g0[0] = (bf16)(2.0f * inv_count * (-ru + ru_n + rv - rv_e));
        }
        bf16x8 *o = reinterpret_cast<bf16x8 *>(dpred + p * CH);
        o[0] = g0;
#pragma unroll
        for (int g = 1; g < CH / 8; ++g) o[g] = bf16x8{};
    }
    red[threadIdx.x] = e2;
    __syncthreads();
    for (int s = blockDim.x / 2; s > 0; s >>= 1) {
        if ((int)threadIdx.x < s) red[threadIdx.x] += red[threadIdx.x + s];
        __syncthreads();
    }
    if (threadIdx.x == 0) part[blockIdx.x] = red[0];
}

// Adam (Kingma & Ba 2015, Algorithm 1) in FP32; omb1 = 1 - b1, omb2 = 1 - b2
// and the bias corrections c1, c2 come from the host, computed in double.
__global__ void k_adam(int n, float *__restrict__ w, float *__restrict__ m, float *__restrict__ v,
                       const float *__restrict__ g, float lr, float b1, float omb1, float b2, float omb2,
                       float eps, float c1, float c2)
{
    const int i = blockIdx.x * blockDim.x + threadIdx.x;
    if (i >= n) return;
    const float gi = g[i];
    const float mi = b1 * m[i] + omb1 * gi;
    const float vi = b2 * v[i] + omb2 * gi * gi;
    m[i] = mi;
    v[i] = vi;
    w[i] -= lr * (mi / c1) / (sqrtf(vi / c2) + eps);
}

// FP32 master weights -> BF16 forward layout [co][kpad] and, for layers with
// 32 input channels, the transposed backward layout [ci][t*32 + co].
__global__ void k_pack(int kpad, int cin, const float *__restrict__ wm, bf16 *__restrict__ wf,
                       bf16 *__restrict__ wt)
{
    const int idx = blockIdx.x * blockDim.x + threadIdx.x;
    if (idx >= 32 * kpad) return;
    const int co = idx / kpad, k = idx - co * kpad;
    const bf16 v = (bf16)wm[idx];
    wf[idx] = v;
    if (wt && k < 9 * cin) {
        const int t = k / cin, ci = k - t * cin;
        wt[(size_t)ci * 288 + t * 32 + co] = v;
    }
}


// ---- closure inference over a set of basins (one block of 4 waves per basin row)

// Sums of (a, b) over a block of 4 waves in a fixed order: an xor butterfly
// in each wave, then the waves in order.  Every thread ends with the same
// bits (IEEE addition commutes).
__device__ inline void block_sum2(double *a, double *b)
{
    __shared__ double red[2][4];
#pragma unroll
    for (int off = 32; off > 0; off >>= 1) {
        *a += __shfl_xor(*a, off);
        *b += __shfl_xor(*b, off);
    }
    const int w = threadIdx.x >> 6;
    if ((threadIdx.x & 63) == 0) {
        red[0][w] = *a;
        red[1][w] = *b;
    }
    __syncthreads();
    *a = ((red[0][0] + red[0][1]) + red[0][2]) + red[0][3];
    *b = ((red[1][0] + red[1][1]) + red[1][2]) + red[1][3];
}

template <class T>
__device__ inline double ld(const void *p, long long i)
{
    return (double)static_cast<const T *>(p)[i];
}

// closure tendency [m/s^2] on the U and V points of member pixel (x, y)
__device__ inline void member_tend(const ClosureMember &m, const Norm &nm, const float *__restrict__ tend, int x,
                                   int y, double *cu, double *cv)
{
    if (nm.kind == OUT_UV) {
        const long long q = (m.pix + (long long)y * m.nx + x) * NOUT;
        *cu = (double)tend[q] * nm.oscale[0];
        *cv = (double)tend[q + 1] * nm.oscale[1];
    } else {
        float du, dv;
        psi_uv(tend + m.pix * NOUT, NOUT, m.ny, m.nx, x, y, &du, &dv);
        *cu = (double)du * nm.oscale[0];
        *cv = (double)dv * nm.oscale[0];
    }
}

// per-row squared norms of the closure tendency and of the plain viscous
// tendency nu * Laplacian(u, v) on U and V points
template <class T>
__global__ void __launch_bounds__(256)
    k_cl_norm(const ClosureMember *__restrict__ mem, const int *__restrict__ row_member,
              const int *__restrict__ row_y, Norm nm, const float *__restrict__ tend,
              double *__restrict__ part)
{
    const ClosureMember m = mem[row_member[blockIdx.x]];
    const int y = row_y[blockIdx.x];
    double sc = 0.0, sv = 0.0;
    if (y >= 1 && y < m.ny - 1) {
        const double k = m.nu / (m.d * m.d);
        for (int i = 1 + threadIdx.x; i < m.nx - 1; i += blockDim.x) {
            const long long a = (long long)y * m.nx + i;
            const uint8_t b = m.mb[a];
            double tu, tv;
            member_tend(m, nm, tend, i, y, &tu, &tv);
            if (b & MB_LCU) {
                const double c = tu;
                const double l = k * (ld<T>(m.u, a - 1) + ld<T>(m.u, a + 1) + ld<T>(m.u, a - m.nx) +
                                      ld<T>(m.u, a + m.nx) - 4.0 * ld<T>(m.u, a));
                sc += c * c;
                sv += l * l;
            }
            if (b & MB_LCV) {
                const double c = tv;
                const double l = k * (ld<T>(m.v, a - 1) + ld<T>(m.v, a + 1) + ld<T>(m.v, a - m.nx) +
                                      ld<T>(m.v, a + m.nx) - 4.0 * ld<T>(m.v, a));
                sc += c * c;
                sv += l * l;
            }
        }
    }
    block_sum2(&sc, &sv);
    if (threadIdx.x == 0) {
        part[2 * (m.row0 + y) + 0] = sc;
        part[2 * (m.row0 + y) + 1] = sv;
    }
}

// The fallback decision of this step and its application.  Every block of a
// member reduces the member's row partials in the same fixed order (so all
// reach the same decision); the block of row 0 records the norms and counts
// a fallback step.  Applied: u += dt * tendency on U points, v likewise.
template <class T>
__global__ void __launch_bounds__(256)
    k_cl_apply(ClosureMember *__restrict__ mem, const int *__restrict__ row_member,
               const int *__restrict__ row_y, Norm nm, const float *__restrict__ tend,
               const double *__restrict__ part, double ratio, double dt, double *__restrict__ norms,
               int *__restrict__ active, int *__restrict__ nfall)
{
    const int mi = row_member[blockIdx.x];
    const ClosureMember m = mem[mi];
    const int y = row_y[blockIdx.x];
    double sc = 0.0, sv = 0.0;
    for (int r = threadIdx.x; r < m.ny; r += blockDim.x) {
        sc += part[2 * (m.row0 + r) + 0];
        sv += part[2 * (m.row0 + r) + 1];
    }
    block_sum2(&sc, &sv);
    const double nc = sqrt(sc), nv = sqrt(sv);
    const int ok_s = nc <= ratio * nv ? 1 : 0;
    if (threadIdx.x == 0 && y == 0) {
        norms[2 * mi + 0] = nc;
        norms[2 * mi + 1] = nv;
        active[mi] = ok_s;
        nfall[mi] += 1 - ok_s;
    }
    if (!ok_s) return;
    T *u = static_cast<T *>(m.u), *v = static_cast<T *>(m.v);
    for (int i = threadIdx.x; i < m.nx; i += blockDim.x) {
        const long long a = (long long)y * m.nx + i;
        const uint8_t b = m.mb[a];
        double tu, tv;
        member_tend(m, nm, tend, i, y, &tu, &tv);
        if (b & MB_LCU) u[a] = (T)((double)u[a] + dt * tu);
        if (b & MB_LCV) v[a] = (T)((double)v[a] + dt * tv);
    }
}

}  // namespace cnn
}  // namespace ibt

using namespace ibt;
using namespace ibt::cnn;

struct inmsom_bt_cnn {
    int device = 0;
    hipStream_t stream = nullptr;
    int B = 0, H = 0, W = 0;  // training batch shape
    Layer L[NL];
    size_t nparam = 0;
    Norm norm;
    unsigned version = 0;  // bumped when norm or fuse changes (captured graphs hold them)
    int fuse = 0;          // inference layers in one pass: 0 = by batch size, 1 = never, 2 = always
    // ensembles using this network as their closure, told on destroy
    std::vector<std::pair<void *, void (*)(void *)>> holders;
    float *wm = nullptr, *am = nullptr, *av = nullptr, *grad = nullptr;  // master, Adam m/v, gradient
    bf16 *wf = nullptr, *wt = nullptr;                                   // BF16 forward / transposed
    size_t wf_off[NL] = {}, wt_off[NL] = {};
    int adam_t = 0;
    // training batch buffers
    bf16 *x = nullptr, *act[3] = {}, *dy = nullptr, *dx = nullptr;
    // inmsom_bt_cnn_keep_backprop: dL/dz of every layer in its own buffer
    // (otherwise the backward pass ping-pongs dy / dx)
    bf16 *gz[NL] = {};
    float *pred = nullptr, *tgt = nullptr, *loss_part = nullptr;
    float *part_w = nullptr, *part_b = nullptr;  // weight-gradient chunk records, group records
    uint8_t *mask = nullptr;
    Img *imgs = nullptr;
    int *tile_img = nullptr, *tile_local = nullptr;
    int ntiles = 0;
    int wg_chunk = 2048, wg_nchunks = 0;  // weight-gradient records: one per wg_chunk pixels of the batch
    // device-resident training set
    int nsamples = 0;
    bf16 *ds_x = nullptr;
    float *ds_t = nullptr;
    uint8_t *ds_m = nullptr;
    std::vector<long long> ds_count;
};

namespace {

int cfail(const std::string &m, int code)
{
    set_error(m);
    return code;
}

#define CHC(expr)                                                                             \
    do {                                                                                      \
        hipError_t e_ = (expr);                                                               \
        if (e_ != hipSuccess) return cfail(std::string(#expr) + ": " + hipGetErrorString(e_), \
                                           INMSOM_BT_EDEVICE);                                \
    } while (0)

#define CTRY(expr)                         \
    do {                                   \
        int r_ = (expr);                   \
        if (r_ != INMSOM_BT_OK) return r_; \
    } while (0)

template <class T>
int dmalloc(T **p, size_t count)
{
    CHC(hipMalloc(p, count * sizeof(T) + 64));
    CHC(hipMemset(*p, 0, count * sizeof(T) + 64));
    return INMSOM_BT_OK;
}

void layer_table(Layer *L, size_t *nparam)
{
    size_t off = 0;
    for (int l = 0; l < NL; ++l) {
        L[l].cin = l == 0 ? CIN0 : CH;
        L[l].kpad = (9 * L[l].cin + 15) / 16 * 16;
        L[l].woff = off;
        off += (size_t)32 * L[l].kpad;
        L[l].boff = off;
        off += 32;
    }
    *nparam = off;
}

// tile tables (one tile = TR rows x TW columns of one image) of a batch of images
int build_tiles(const std::vector<Img> &imgs, Img **d_imgs, int **d_ti, int **d_tl, int *ntiles)
{
    std::vector<int> ti, tl;
    for (size_t k = 0; k < imgs.size(); ++k) {
        const int n = ((imgs[k].H + TR - 1) / TR) * ((imgs[k].W + TW - 1) / TW);
        for (int t = 0; t < n; ++t) {
            ti.push_back((int)k);
            tl.push_back(t);
        }
    }
    CTRY(dmalloc(d_imgs, imgs.size()));
    CTRY(dmalloc(d_ti, ti.size()));
    CTRY(dmalloc(d_tl, tl.size()));
    CHC(hipMemcpy(*d_imgs, imgs.data(), imgs.size() * sizeof(Img), hipMemcpyHostToDevice));
    CHC(hipMemcpy(*d_ti, ti.data(), ti.size() * 4, hipMemcpyHostToDevice));
    CHC(hipMemcpy(*d_tl, tl.data(), tl.size() * 4, hipMemcpyHostToDevice));
    *ntiles = (int)ti.size();
    return INMSOM_BT_OK;
}

void conv_launch(const inmsom_bt_cnn *c, hipStream_t st, int l, bool bwd, const Img *imgs, const int *ti,
                 const int *tl, int ntiles, const bf16 *in, const bf16 *act_prev, void *out)
{
    // persistent blocks: a few per CU, each looping over tiles
    const dim3 grid((unsigned)(ntiles < 2048 ? ntiles : 2048)), blk(256);
    const Layer &Ly = c->L[l];
    const float *bias = c->wm + Ly.boff;
    if (!bwd) {
        const bf16 *w = c->wf + c->wf_off[l];
        if (l == 0)
            hipLaunchKernelGGL((k_conv<CIN0, false, EPI_RELU_BF16>), grid, blk, 0, st, imgs, ti, tl, ntiles, in,
                               w, bias, nullptr, out);
        else if (l < NL - 1)
            hipLaunchKernelGGL((k_conv<CH, false, EPI_RELU_BF16>), grid, blk, 0, st, imgs, ti, tl, ntiles, in,
                               w, bias, nullptr, out);
        else
            hipLaunchKernelGGL((k_conv<CH, false, EPI_OUT_F32>), grid, blk, 0, st, imgs, ti, tl, ntiles, in, w,
                               bias, nullptr, out);
    } else {
        // data gradient through layer l (l >= 1): weights as [ci][t*32 + co]
        const bf16 *w = c->wt + c->wt_off[l];
        hipLaunchKernelGGL((k_conv<CH, true, EPI_MASK_BF16>), grid, blk, 0, st, imgs, ti, tl, ntiles, in, w,
                           nullptr, act_prev, out);
    }
}

void wgrad_launch(inmsom_bt_cnn *c, int l, const bf16 *dy, const bf16 *x)
{
    // work items (image, 16-row band, 64-column strip) in runs of `per`
    // per block, at most one record per block of the record buffer
    const int nitems = c->B * ((c->H + WG_R - 1) / WG_R) * ((c->W + WG_TW - 1) / WG_TW);
    const int per = (nitems + c->wg_nchunks - 1) / c->wg_nchunks;
    const int nch = (nitems + per - 1) / per;
    const Layer &Ly = c->L[l];
    if (l == 0)
        hipLaunchKernelGGL(k_wgrad<CIN0>, dim3(nch), dim3(3 * 64), 0, c->stream, c->H, c->W, nitems, per, dy, x,
                           c->part_w);
    else
        hipLaunchKernelGGL(k_wgrad<CH>, dim3(nch), dim3(3 * 64), 0, c->stream, c->H, c->W, nitems, per, dy, x,
                           c->part_w);
    const int ncol = l == 0 ? 96 : 288;
    const int rec = 32 * ncol + 32;
    const int ngroups = (nch + RG - 1) / RG;
    hipLaunchKernelGGL(k_reduce_groups, dim3((rec + 255) / 256, ngroups), dim3(256), 0, c->stream, nch, rec,
                       c->part_w, c->part_b);
    const int n = 32 * Ly.kpad;
    hipLaunchKernelGGL(k_reduce_wgrad, dim3((n + 255) / 256), dim3(256), 0, c->stream, ngroups, ncol,
                       9 * Ly.cin, Ly.kpad, c->part_b, c->grad + Ly.woff, c->grad + Ly.boff);
}

void pack_all(inmsom_bt_cnn *c)
{
    for (int l = 0; l < NL; ++l) {
        const Layer &Ly = c->L[l];
        const int n = 32 * Ly.kpad;
        hipLaunchKernelGGL(k_pack, dim3((n + 255) / 256), dim3(256), 0, c->stream, Ly.kpad, Ly.cin,
                           c->wm + Ly.woff, c->wf + c->wf_off[l], l >= 1 ? c->wt + c->wt_off[l] : nullptr);
    }
}

// forward over a set of images: a0 <- L0(x), a1 <- L1(a0), a2 <- L2(a1), out <- L3(a2)
void forward_images(const inmsom_bt_cnn *c, hipStream_t st, const Img *imgs, const int *ti, const int *tl,
                    int ntiles, const bf16 *x, bf16 *a0, bf16 *a1, bf16 *a2, float *out)
{
    conv_launch(c, st, 0, false, imgs, ti, tl, ntiles, x, nullptr, a0);
    conv_launch(c, st, 1, false, imgs, ti, tl, ntiles, a0, nullptr, a1);
    conv_launch(c, st, 2, false, imgs, ti, tl, ntiles, a1, nullptr, a2);
    conv_launch(c, st, 3, false, imgs, ti, tl, ntiles, a2, nullptr, out);
}

void forward_batch(inmsom_bt_cnn *c)
{
    forward_images(c, c->stream, c->imgs, c->tile_img, c->tile_local, c->ntiles, c->x, c->act[0], c->act[1],
                   c->act[2], c->pred);
}

// loss + full backward of the current batch; gradient into c->grad
int backward(inmsom_bt_cnn *c, long long count, double *loss)
{
    const long long np = (long long)c->B * c->H * c->W;
    const float inv = count > 0 ? 1.0f / (float)count : 0.0f;
    const int nb = (int)((np + 255) / 256);
    // dL/dz_l of layer l (z_3: the outputs; z_0..2: before the ReLU)
    bf16 *g3 = c->gz[3] ? c->gz[3] : c->dy, *g2 = c->gz[2] ? c->gz[2] : c->dx;
    bf16 *g1 = c->gz[1] ? c->gz[1] : c->dy, *g0 = c->gz[0] ? c->gz[0] : c->dx;
    hipLaunchKernelGGL(k_loss, dim3(nb), dim3(256), 0, c->stream, np, c->H, c->W, c->norm.kind, c->pred, c->tgt,
                       c->mask, inv, g3, c->loss_part);
    wgrad_launch(c, 3, g3, c->act[2]);
    conv_launch(c, c->stream, 3, true, c->imgs, c->tile_img, c->tile_local, c->ntiles, g3, c->act[2], g2);
    wgrad_launch(c, 2, g2, c->act[1]);
    conv_launch(c, c->stream, 2, true, c->imgs, c->tile_img, c->tile_local, c->ntiles, g2, c->act[1], g1);
    wgrad_launch(c, 1, g1, c->act[0]);
    conv_launch(c, c->stream, 1, true, c->imgs, c->tile_img, c->tile_local, c->ntiles, g1, c->act[0], g0);
    wgrad_launch(c, 0, g0, c->x);
    CHC(hipGetLastError());
    if (loss) {
        std::vector<float> part(nb);
        CHC(hipMemcpyAsync(part.data(), c->loss_part, (size_t)nb * 4, hipMemcpyDeviceToHost, c->stream));
        CHC(hipStreamSynchronize(c->stream));
        double s = 0.0;
        for (float v : part) s += v;
        *loss = count > 0 ? s / (double)count : 0.0;
    }
    return INMSOM_BT_OK;
}

std::vector<bf16> to_bf16(const float *x, size_t n)
{
    std::vector<bf16> out(n);
    for (size_t i = 0; i < n; ++i) out[i] = (bf16)x[i];
    return out;
}

}  // namespace

// ---- batched closure inference (used by the ensemble step)
namespace ibt {
namespace cnn {

int infer_create(int device, int nmem, const int *nx, const int *ny, Infer *inf)
{
    CHC(hipSetDevice(device));
    *inf = Infer();
    inf->nmem = nmem;
    std::vector<Img> imgs(nmem);
    std::vector<int> rm, ry;
    long long pix = 0;
    for (int k = 0; k < nmem; ++k) {
        imgs[k] = Img{ny[k], nx[k], pix};
        pix += (long long)nx[k] * ny[k];
        for (int y = 0; y < ny[k]; ++y) {
            rm.push_back(k);
            ry.push_back(y);
        }
    }
    inf->npix = pix;
    inf->nrows = (int)rm.size();
    CTRY(build_tiles(imgs, &inf->imgs, &inf->tile_img, &inf->tile_local, &inf->ntiles));
    CTRY(dmalloc(&inf->row_member, rm.size()));
    CTRY(dmalloc(&inf->row_y, ry.size()));
    CHC(hipMemcpy(inf->row_member, rm.data(), rm.size() * 4, hipMemcpyHostToDevice));
    CHC(hipMemcpy(inf->row_y, ry.data(), ry.size() * 4, hipMemcpyHostToDevice));
    CTRY(dmalloc(&inf->a0, (size_t)pix * CH));
    CTRY(dmalloc(&inf->a1, (size_t)pix * CH));
    CTRY(dmalloc(&inf->tend, (size_t)pix * NOUT));
    CTRY(dmalloc(&inf->part, (size_t)inf->nrows * 2));
    CTRY(dmalloc(&inf->norms, (size_t)nmem * 2));
    CTRY(dmalloc(&inf->active, (size_t)nmem));
    CTRY(dmalloc(&inf->nfall, (size_t)nmem));
    return INMSOM_BT_OK;
}

void infer_destroy(Infer *inf)
{
    for (void *p : {(void *)inf->imgs, (void *)inf->tile_img, (void *)inf->tile_local, (void *)inf->row_member,
                    (void *)inf->row_y, (void *)inf->a0, (void *)inf->a1, (void *)inf->tend,
                    (void *)inf->part, (void *)inf->norms, (void *)inf->active, (void *)inf->nfall})
        if (p) hipFree(p);
    *inf = Infer();
}

bool closure_fused(const inmsom_bt_cnn *net, const Infer &inf)
{
    return net->fuse == 2 || (net->fuse == 0 && inf.ntiles <= FUSED_MAX_TILES);
}

void closure_launch(const inmsom_bt_cnn *net, hipStream_t st, bool f32, ClosureMember *members, Infer &inf,
                    double dt, double ratio)
{
    const dim3 rows((unsigned)inf.nrows), blk(256);
    const dim3 grid((unsigned)(inf.ntiles < 2048 ? inf.ntiles : 2048));
    const Layer &L0 = net->L[0];
    const Layer &L1 = net->L[1];
    const bf16 *w0 = net->wf + net->wf_off[0], *w1 = net->wf + net->wf_off[1];
    if (closure_fused(net, inf)) {
        // the four layers in one pass (-> tend)
        const bf16 *w2 = net->wf + net->wf_off[2], *w3 = net->wf + net->wf_off[3];
        const float *bs[NL] = {net->wm + L0.boff, net->wm + L1.boff, net->wm + net->L[2].boff,
                               net->wm + net->L[3].boff};
        if (f32)
            hipLaunchKernelGGL((k_closure4<1>), grid, blk, 0, st, inf.imgs, inf.tile_img, inf.tile_local, inf.ntiles,
                               w0, bs[0], w1, bs[1], w2, bs[2], w3, bs[3], inf.tend, members, net->norm);
        else
            hipLaunchKernelGGL((k_closure4<2>), grid, blk, 0, st, inf.imgs, inf.tile_img, inf.tile_local, inf.ntiles,
                               w0, bs[0], w1, bs[1], w2, bs[2], w3, bs[3], inf.tend, members, net->norm);
    } else {
        // layers 0 and 1 from the members' state in one pass (-> a1), then a1 -> a0 -> tend
        if (f32)
            hipLaunchKernelGGL((k_conv01<1>), grid, blk, 0, st, inf.imgs, inf.tile_img, inf.tile_local, inf.ntiles,
                               w0, net->wm + L0.boff, w1, net->wm + L1.boff, inf.a1, members, net->norm);
        else
            hipLaunchKernelGGL((k_conv01<2>), grid, blk, 0, st, inf.imgs, inf.tile_img, inf.tile_local, inf.ntiles,
                               w0, net->wm + L0.boff, w1, net->wm + L1.boff, inf.a1, members, net->norm);
        conv_launch(net, st, 2, false, inf.imgs, inf.tile_img, inf.tile_local, inf.ntiles, inf.a1, nullptr, inf.a0);
        conv_launch(net, st, 3, false, inf.imgs, inf.tile_img, inf.tile_local, inf.ntiles, inf.a0, nullptr,
                    inf.tend);
    }
    if (f32) {
        hipLaunchKernelGGL(k_cl_norm<float>, rows, blk, 0, st, members, inf.row_member, inf.row_y, net->norm,
                           inf.tend, inf.part);
        hipLaunchKernelGGL(k_cl_apply<float>, rows, blk, 0, st, members, inf.row_member, inf.row_y, net->norm,
                           inf.tend, inf.part, ratio, dt, inf.norms, inf.active, inf.nfall);
    } else {
        hipLaunchKernelGGL(k_cl_norm<double>, rows, blk, 0, st, members, inf.row_member, inf.row_y, net->norm,
                           inf.tend, inf.part);
        hipLaunchKernelGGL(k_cl_apply<double>, rows, blk, 0, st, members, inf.row_member, inf.row_y, net->norm,
                           inf.tend, inf.part, ratio, dt, inf.norms, inf.active, inf.nfall);
    }
}

int net_device(const inmsom_bt_cnn *net) { return net->device; }
unsigned net_version(const inmsom_bt_cnn *net) { return net->version; }
void net_attach(inmsom_bt_cnn *net, void *owner, void (*detach)(void *owner))
{
    net->holders.emplace_back(owner, detach);
}
void net_detach(inmsom_bt_cnn *net, void *owner)
{
    auto &h = net->holders;
    for (size_t i = 0; i < h.size();)
        if (h[i].first == owner)
            h.erase(h.begin() + (long)i);
        else
            ++i;
}

}  // namespace cnn
}  // namespace ibt

extern "C" {

int inmsom_bt_cnn_create(int batch, int ny, int nx, int device, inmsom_bt_cnn **out)
{
    if (!out || batch < 1 || ny < 3 || nx < 3) return cfail("bad arguments", INMSOM_BT_EINVAL);
    *out = nullptr;
    CHC(hipSetDevice(device));
    auto *c = new inmsom_bt_cnn();
    c->device = device;
    c->B = batch;
    c->H = ny;
    c->W = nx;
    layer_table(c->L, &c->nparam);
    for (int k = 0; k < NIN; ++k) {
        c->norm.shift[k] = 0.0f;
        c->norm.scale[k] = 1.0f;
    }
    c->norm.oscale[0] = c->norm.oscale[1] = 1.0f;
    c->norm.kind = OUT_UV;
    int rc = [&]() -> int {
        CHC(hipStreamCreateWithFlags(&c->stream, hipStreamNonBlocking));
        CTRY(dmalloc(&c->wm, c->nparam));
        CTRY(dmalloc(&c->am, c->nparam));
        CTRY(dmalloc(&c->av, c->nparam));
        CTRY(dmalloc(&c->grad, c->nparam));
        size_t fo = 0, to = 0;
        for (int l = 0; l < NL; ++l) {
            c->wf_off[l] = fo;
            fo += (size_t)32 * c->L[l].kpad;
            c->wt_off[l] = to;
            if (l >= 1) to += (size_t)32 * 9 * CH;
        }
        CTRY(dmalloc(&c->wf, fo));
        CTRY(dmalloc(&c->wt, to));
        const size_t np = (size_t)batch * ny * nx;
        CTRY(dmalloc(&c->x, np * CIN0));
        for (int k = 0; k < 3; ++k) CTRY(dmalloc(&c->act[k], np * CH));
        CTRY(dmalloc(&c->dy, np * CH));
        CTRY(dmalloc(&c->dx, np * CH));
        CTRY(dmalloc(&c->pred, np * NOUT));
        CTRY(dmalloc(&c->tgt, np * NOUT));
        CTRY(dmalloc(&c->mask, np));
        CTRY(dmalloc(&c->loss_part, (np + 255) / 256));
        c->wg_nchunks = (int)((np + c->wg_chunk - 1) / c->wg_chunk);
        // chunk records, then group records (32 x 288 weights + 32 biases each)
        CTRY(dmalloc(&c->part_w, (size_t)c->wg_nchunks * (32 * 9 * CH + 32)));
        CTRY(dmalloc(&c->part_b, (size_t)((c->wg_nchunks + RG - 1) / RG) * (32 * 9 * CH + 32)));
        std::vector<Img> imgs(batch);
        for (int b = 0; b < batch; ++b) imgs[b] = Img{ny, nx, (long long)b * ny * nx};
        CTRY(build_tiles(imgs, &c->imgs, &c->tile_img, &c->tile_local, &c->ntiles));
        return INMSOM_BT_OK;
    }();
    if (rc) {
        inmsom_bt_cnn_destroy(c);
        return rc;
    }
    *out = c;
    return INMSOM_BT_OK;
}

int inmsom_bt_cnn_destroy(inmsom_bt_cnn *c)
{
    if (!c) return INMSOM_BT_OK;
    // ensembles still holding the network drop it (and their graphs) first
    while (!c->holders.empty()) {
        auto h = c->holders.back();
        c->holders.pop_back();
        h.second(h.first);
    }
    hipSetDevice(c->device);
    if (c->stream) hipStreamSynchronize(c->stream);
    for (void *p : {(void *)c->wm, (void *)c->am, (void *)c->av, (void *)c->grad, (void *)c->wf,
                    (void *)c->wt, (void *)c->x, (void *)c->act[0], (void *)c->act[1], (void *)c->act[2],
                    (void *)c->dy, (void *)c->dx, (void *)c->pred, (void *)c->tgt, (void *)c->mask,
                    (void *)c->loss_part, (void *)c->part_w, (void *)c->part_b, (void *)c->imgs,
                    (void *)c->tile_img, (void *)c->tile_local, (void *)c->ds_x, (void *)c->ds_t,
                    (void *)c->ds_m, (void *)c->gz[0], (void *)c->gz[1], (void *)c->gz[2], (void *)c->gz[3]})
        if (p) hipFree(p);
    if (c->stream) hipStreamDestroy(c->stream);
    delete c;
    return INMSOM_BT_OK;
}

int inmsom_bt_cnn_num_params(inmsom_bt_cnn *c, long long *n)
{
    if (!c || !n) return cfail("null pointer", INMSOM_BT_EINVAL);
    *n = (long long)c->nparam;
    return INMSOM_BT_OK;
}

int inmsom_bt_cnn_set_norm(inmsom_bt_cnn *c, const float *in_shift, const float *in_scale,
                           const float *out_scale)
{
    if (!c || !in_shift || !in_scale || !out_scale) return cfail("null pointer", INMSOM_BT_EINVAL);
    for (int k = 0; k < NIN; ++k) {
        c->norm.shift[k] = in_shift[k];
        c->norm.scale[k] = in_scale[k];
    }
    c->norm.oscale[0] = out_scale[0];
    c->norm.oscale[1] = out_scale[1];
    ++c->version;
    return INMSOM_BT_OK;
}

int inmsom_bt_cnn_set_fused(inmsom_bt_cnn *c, int mode)
{
    if (!c) return cfail("null pointer", INMSOM_BT_EINVAL);
    if (mode < 0 || mode > 2) return cfail("fused mode must be 0 (by batch size), 1 (never) or 2 (always)",
                                           INMSOM_BT_EINVAL);
    c->fuse = mode;
    ++c->version;
    return INMSOM_BT_OK;
}

int inmsom_bt_cnn_set_output(inmsom_bt_cnn *c, int kind)
{
    if (!c) return cfail("null pointer", INMSOM_BT_EINVAL);
    if (kind != OUT_UV && kind != OUT_PSI) return cfail("unknown output kind", INMSOM_BT_EINVAL);
    c->norm.kind = kind;
    ++c->version;
    return INMSOM_BT_OK;
}

int inmsom_bt_cnn_set_params(inmsom_bt_cnn *c, const float *params)
{
    if (!c || !params) return cfail("null pointer", INMSOM_BT_EINVAL);
    CHC(hipSetDevice(c->device));
    CHC(hipStreamSynchronize(c->stream));
    // padded entries of W (k >= 9 cin) must stay zero
    std::vector<float> p(params, params + c->nparam);
    for (int l = 0; l < NL; ++l)
        for (int co = 0; co < 32; ++co)
            for (int k = 9 * c->L[l].cin; k < c->L[l].kpad; ++k) p[c->L[l].woff + (size_t)co * c->L[l].kpad + k] = 0.f;
    CHC(hipMemcpy(c->wm, p.data(), c->nparam * 4, hipMemcpyHostToDevice));
    CHC(hipMemset(c->am, 0, c->nparam * 4));
    CHC(hipMemset(c->av, 0, c->nparam * 4));
    c->adam_t = 0;
    pack_all(c);
    CHC(hipGetLastError());
    CHC(hipStreamSynchronize(c->stream));
    return INMSOM_BT_OK;
}

int inmsom_bt_cnn_get_params(inmsom_bt_cnn *c, float *params)
{
    if (!c || !params) return cfail("null pointer", INMSOM_BT_EINVAL);
    CHC(hipSetDevice(c->device));
    CHC(hipStreamSynchronize(c->stream));
    CHC(hipMemcpy(params, c->wm, c->nparam * 4, hipMemcpyDeviceToHost));
    return INMSOM_BT_OK;
}

int inmsom_bt_cnn_get_grad(inmsom_bt_cnn *c, float *grad)
{
    if (!c || !grad) return cfail("null pointer", INMSOM_BT_EINVAL);
    CHC(hipSetDevice(c->device));
    CHC(hipStreamSynchronize(c->stream));
    CHC(hipMemcpy(grad, c->grad, c->nparam * 4, hipMemcpyDeviceToHost));
    return INMSOM_BT_OK;
}

int inmsom_bt_cnn_set_batch(inmsom_bt_cnn *c, const float *x, const float *target, const uint8_t *mask)
{
    if (!c || !x) return cfail("null pointer", INMSOM_BT_EINVAL);
    CHC(hipSetDevice(c->device));
    CHC(hipStreamSynchronize(c->stream));
    const size_t np = (size_t)c->B * c->H * c->W;
    const std::vector<bf16> xb = to_bf16(x, np * CIN0);
    CHC(hipMemcpy(c->x, xb.data(), xb.size() * 2, hipMemcpyHostToDevice));
    if (target) CHC(hipMemcpy(c->tgt, target, np * NOUT * 4, hipMemcpyHostToDevice));
    if (mask) CHC(hipMemcpy(c->mask, mask, np, hipMemcpyHostToDevice));
    return INMSOM_BT_OK;
}

int inmsom_bt_cnn_forward(inmsom_bt_cnn *c, float *out)
{
    if (!c) return cfail("null pointer", INMSOM_BT_EINVAL);
    CHC(hipSetDevice(c->device));
    forward_batch(c);
    CHC(hipGetLastError());
    if (out)
        CHC(hipMemcpyAsync(out, c->pred, (size_t)c->B * c->H * c->W * NOUT * 4, hipMemcpyDeviceToHost,
                           c->stream));
    CHC(hipStreamSynchronize(c->stream));
    return INMSOM_BT_OK;
}

int inmsom_bt_cnn_get_activation(inmsom_bt_cnn *c, int layer, float *out)
{
    if (!c || !out || layer < 0 || layer > 2) return cfail("bad arguments", INMSOM_BT_EINVAL);
    CHC(hipSetDevice(c->device));
    CHC(hipStreamSynchronize(c->stream));
    const size_t n = (size_t)c->B * c->H * c->W * CH;
    std::vector<bf16> a(n);
    CHC(hipMemcpy(a.data(), c->act[layer], n * 2, hipMemcpyDeviceToHost));
    for (size_t i = 0; i < n; ++i) out[i] = (float)a[i];
    return INMSOM_BT_OK;
}

int inmsom_bt_cnn_keep_backprop(inmsom_bt_cnn *c, int on)
{
    if (!c) return cfail("null pointer", INMSOM_BT_EINVAL);
    CHC(hipSetDevice(c->device));
    CHC(hipStreamSynchronize(c->stream));
    const size_t n = (size_t)c->B * c->H * c->W * CH;
    for (int l = 0; l < NL; ++l) {
        if (on && !c->gz[l]) {
            CTRY(dmalloc(&c->gz[l], n));
            CHC(hipMemset(c->gz[l], 0, n * sizeof(bf16)));
        } else if (!on && c->gz[l]) {
            CHC(hipFree(c->gz[l]));
            c->gz[l] = nullptr;
        }
    }
    return INMSOM_BT_OK;
}

int inmsom_bt_cnn_get_backprop(inmsom_bt_cnn *c, int layer, float *out)
{
    if (!c || !out || layer < 0 || layer >= NL) return cfail("bad arguments", INMSOM_BT_EINVAL);
    if (!c->gz[layer]) return cfail("back-propagated values are kept only after inmsom_bt_cnn_keep_backprop",
                                    INMSOM_BT_ESTATE);
    CHC(hipSetDevice(c->device));
    CHC(hipStreamSynchronize(c->stream));
    const size_t n = (size_t)c->B * c->H * c->W * CH;
    std::vector<bf16> a(n);
    CHC(hipMemcpy(a.data(), c->gz[layer], n * 2, hipMemcpyDeviceToHost));
    for (size_t i = 0; i < n; ++i) out[i] = (float)a[i];
    return INMSOM_BT_OK;
}

int inmsom_bt_cnn_loss_grad(inmsom_bt_cnn *c, double *loss)
{
    if (!c) return cfail("null pointer", INMSOM_BT_EINVAL);
    CHC(hipSetDevice(c->device));
    CHC(hipStreamSynchronize(c->stream));
    const size_t np = (size_t)c->B * c->H * c->W;
    std::vector<uint8_t> mk(np);
    CHC(hipMemcpy(mk.data(), c->mask, np, hipMemcpyDeviceToHost));
    long long count = 0;
    for (uint8_t m : mk) count += (m & 1) + ((m >> 1) & 1);
    forward_batch(c);
    CTRY(backward(c, count, loss));
    CHC(hipStreamSynchronize(c->stream));
    return INMSOM_BT_OK;
}

int inmsom_bt_cnn_dataset_alloc(inmsom_bt_cnn *c, int nsamples)
{
    if (!c || nsamples < 1) return cfail("bad arguments", INMSOM_BT_EINVAL);
    CHC(hipSetDevice(c->device));
    CHC(hipStreamSynchronize(c->stream));
    for (void **p : {(void **)&c->ds_x, (void **)&c->ds_t, (void **)&c->ds_m})
        if (*p) {
            CHC(hipFree(*p));
            *p = nullptr;
        }
    c->nsamples = 0;
    const size_t hw = (size_t)c->H * c->W;
    CTRY(dmalloc(&c->ds_x, (size_t)nsamples * hw * CIN0));
    CTRY(dmalloc(&c->ds_t, (size_t)nsamples * hw * NOUT));
    CTRY(dmalloc(&c->ds_m, (size_t)nsamples * hw));
    c->nsamples = nsamples;
    c->ds_count.assign(nsamples, 0);
    return INMSOM_BT_OK;
}

int inmsom_bt_cnn_dataset_set(inmsom_bt_cnn *c, int i, const float *x, const float *target, const uint8_t *mask)
{
    if (!c || i < 0 || i >= c->nsamples || !x || !target || !mask) return cfail("bad arguments", INMSOM_BT_EINVAL);
    CHC(hipSetDevice(c->device));
    const size_t hw = (size_t)c->H * c->W;
    const std::vector<bf16> xb = to_bf16(x, hw * CIN0);
    CHC(hipMemcpy(c->ds_x + (size_t)i * hw * CIN0, xb.data(), xb.size() * 2, hipMemcpyHostToDevice));
    CHC(hipMemcpy(c->ds_t + (size_t)i * hw * NOUT, target, hw * NOUT * 4, hipMemcpyHostToDevice));
    CHC(hipMemcpy(c->ds_m + (size_t)i * hw, mask, hw, hipMemcpyHostToDevice));
    long long cnt = 0;
    for (size_t k = 0; k < hw; ++k) cnt += (mask[k] & 1) + ((mask[k] >> 1) & 1);
    c->ds_count[i] = cnt;
    return INMSOM_BT_OK;
}

}  // extern "C"

namespace {
// gathers samples idx[0..B-1] of the training set into the batch buffers
int gather(inmsom_bt_cnn *c, const int *idx, long long *count)
{
    if (!c || !idx) return cfail("null pointer", INMSOM_BT_EINVAL);
    if (!c->ds_x) return cfail("no training set", INMSOM_BT_ESTATE);
    for (int b = 0; b < c->B; ++b)
        if (idx[b] < 0 || idx[b] >= c->nsamples) return cfail("sample index out of range", INMSOM_BT_EINVAL);
    CHC(hipSetDevice(c->device));
    const size_t hw = (size_t)c->H * c->W;
    *count = 0;
    for (int b = 0; b < c->B; ++b) {
        const size_t i = (size_t)idx[b];
        CHC(hipMemcpyAsync(c->x + b * hw * CIN0, c->ds_x + i * hw * CIN0, hw * CIN0 * 2, hipMemcpyDeviceToDevice,
                           c->stream));
        CHC(hipMemcpyAsync(c->tgt + b * hw * NOUT, c->ds_t + i * hw * NOUT, hw * NOUT * 4,
                           hipMemcpyDeviceToDevice, c->stream));
        CHC(hipMemcpyAsync(c->mask + b * hw, c->ds_m + i * hw, hw, hipMemcpyDeviceToDevice, c->stream));
        *count += c->ds_count[i];
    }
    return INMSOM_BT_OK;
}
}  // namespace

extern "C" {

int inmsom_bt_cnn_eval(inmsom_bt_cnn *c, const int *idx, double *loss)
{
    long long count = 0;
    CTRY(gather(c, idx, &count));
    if (!loss) return cfail("null pointer", INMSOM_BT_EINVAL);
    forward_batch(c);
    const long long np = (long long)c->B * c->H * c->W;
    const int nb = (int)((np + 255) / 256);
    hipLaunchKernelGGL(k_loss, dim3(nb), dim3(256), 0, c->stream, np, c->H, c->W, c->norm.kind, c->pred, c->tgt,
                       c->mask,
                       count > 0 ? 1.0f / (float)count : 0.0f, c->dy, c->loss_part);
    CHC(hipGetLastError());
    std::vector<float> part(nb);
    CHC(hipMemcpyAsync(part.data(), c->loss_part, (size_t)nb * 4, hipMemcpyDeviceToHost, c->stream));
    CHC(hipStreamSynchronize(c->stream));
    double s = 0.0;
    for (float v : part) s += v;
    *loss = count > 0 ? s / (double)count : 0.0;
    return INMSOM_BT_OK;
}

int inmsom_bt_cnn_train_step(inmsom_bt_cnn *c, const int *idx, float lr, double *loss)
{
    long long count = 0;
    CTRY(gather(c, idx, &count));
    forward_batch(c);
    CTRY(backward(c, count, loss));
    c->adam_t += 1;
    const double b1 = 0.9, b2 = 0.999;
    const float c1 = (float)(1.0 - pow(b1, c->adam_t)), c2 = (float)(1.0 - pow(b2, c->adam_t));
    const int n = (int)c->nparam;
    hipLaunchKernelGGL(k_adam, dim3((n + 255) / 256), dim3(256), 0, c->stream, n, c->wm, c->am, c->av, c->grad, lr,
                       (float)b1, (float)(1.0 - b1), (float)b2, (float)(1.0 - b2), 1e-8f, c1, c2);
    pack_all(c);
    CHC(hipGetLastError());
    CHC(hipStreamSynchronize(c->stream));
    return INMSOM_BT_OK;
}

int inmsom_bt_cnn_synchronize(inmsom_bt_cnn *c)
{
    if (!c) return cfail("null pointer", INMSOM_BT_EINVAL);
    CHC(hipSetDevice(c->device));
    CHC(hipStreamSynchronize(c->stream));
    return INMSOM_BT_OK;
}

}  // extern "C"
