// Neural-CA surrogate kernels for gfx950: coarse-graining encoder, forward
// and backward of a 3x3 depthwise perception + 1x1 MLP (16 -> 64 -> 16).
//
// Layouts (all device memory):
//   fine field  uint8 [batch][fine_h][fine_w], value = byte & value_mask
//   x, y, dy    fp32  [batch][h][w][16]   (coarse grid, channels last)
//   pred        fp32  [batch][h][w]       (mean of the 16 output channels)
//   dw          fp32  [16][9]   tap t = (dy + 1) * 3 + (dx + 1)
//   bdw         fp32  [16]
//   w1          bf16  [64][16]  (hidden x input), b1 fp32 [64]
//   w2          bf16  [16][64]  (output x hidden), b2 fp32 [16]
//   grads       fp32  [kParams] = dw | bdw | w1 | b1 | w2 | b2
// The coarse grid is a torus per sample.  Coarse cell (y, x) covers fine
// rows 16y..16y+15 and columns 16x..16x+15; input channel 4*sy + sx is the
// mean of the 4x4 fine sub-block (sy, sx) of it.
//
// Forward of one coarse cell:
//   p   = bdw + sum_t dw[:, t] * x[neighbour t]        (fp32, taps in order)
//   z1  = b1 + W1 . bf16(p)                              (MFMA, fp32 acc)
//   y   = b2 + W2 . bf16(relu(z1))                       (MFMA, fp32 acc)
// The two layers run on 16 cells at a time in the transposed form
// Z1^T = W1 . P^T and Y^T = W2 . H^T.  Layer 1 contracts over the 16
// input channels: v_mfma_f32_16x16x16_bf16, A[m][k] in lane (m, k/4),
// B[k][n] and D[m][n] in lane (n, m/4), one instruction per block of 16
// hidden units.  The D tile of a hidden block is then already laid out as
// a B operand: lane (n, g) holds hidden units 16 b + 4g .. +3.  Layer 2
// contracts over the 64 hidden units with gfx950's
// v_mfma_f32_16x16x32_bf16 (A[m][k] in lane (m, k/8), B[k][n] in lane
// (n, k/8)), two hidden blocks per instruction: the order of k inside a
// contraction is free, so lane g's eight k values are taken as the four
// units of block 2j followed by the four of block 2j + 1 -- exactly what
// the lane holds -- and W2's A operand is loaded in the same order.  No
// data moves between lanes.  Each lane owns one cell (lane % 16) and four
// channels (4 * (lane / 16)).
#include <hip/hip_runtime.h>
#include <stdint.h>

#include <mutex>

#include "xentica_hip.h"
#include "xentica_nca.h"
#include "common.h"

namespace {

constexpr int kC = 16;
constexpr int kHid = 64;
constexpr int kTaps = 9;
constexpr int kBlock = 256;
constexpr int kWaves = kBlock / 64;

// parameter vector offsets
constexpr int kOffDw = 0;
constexpr int kOffBdw = kOffDw + kC * kTaps;
constexpr int kOffW1 = kOffBdw + kC;
constexpr int kOffB1 = kOffW1 + kHid * kC;
constexpr int kOffW2 = kOffB1 + kHid;
constexpr int kOffB2 = kOffW2 + kC * kHid;
constexpr int kParams = kOffB2 + kC;
static_assert(kParams == XEN_NCA_PARAMS, "parameter layout");

typedef short bf16x4 __attribute__((ext_vector_type(4)));
typedef short bf16x8 __attribute__((ext_vector_type(8)));
typedef __bf16 mfma_bf16x8 __attribute__((ext_vector_type(8)));
typedef float f32x4 __attribute__((ext_vector_type(4)));

// fp32 -> bf16 bits, round to nearest even (inputs are finite)
__device__ __forceinline__ short to_bf16(float f) {
    uint32_t u = __float_as_uint(f);
    u += 0x7FFFu + ((u >> 16) & 1u);
    return static_cast<short>(u >> 16);
}

__device__ __forceinline__ f32x4 mfma(bf16x4 a, bf16x4 b, f32x4 c) {
    return __builtin_amdgcn_mfma_f32_16x16x16bf16_1k(a, b, c, 0, 0, 0);
}

// D += A . B over K = 32 (v_mfma_f32_16x16x32_bf16): a lane's eight k
// values are the four of a_lo / b_lo followed by the four of a_hi / b_hi.
__device__ __forceinline__ f32x4 mfma32(bf16x4 a_lo, bf16x4 a_hi, bf16x4 b_lo, bf16x4 b_hi,
                                        f32x4 c) {
    const bf16x8 a = __builtin_shufflevector(a_lo, a_hi, 0, 1, 2, 3, 4, 5, 6, 7);
    const bf16x8 b = __builtin_shufflevector(b_lo, b_hi, 0, 1, 2, 3, 4, 5, 6, 7);
    return __builtin_amdgcn_mfma_f32_16x16x32_bf16(__builtin_bit_cast(mfma_bf16x8, a),
                                                   __builtin_bit_cast(mfma_bf16x8, b), c, 0,
                                                   0, 0);
}

__device__ __forceinline__ void wave_sync() {
    __builtin_amdgcn_fence(__ATOMIC_RELEASE, "wavefront");
    __builtin_amdgcn_wave_barrier();
    __builtin_amdgcn_fence(__ATOMIC_ACQUIRE, "wavefront");
}

__device__ __forceinline__ uint32_t byte_sum(uint32_t w) {
    uint32_t s = (w & 0x00FF00FFu) + ((w >> 8) & 0x00FF00FFu);
    return (s & 0xFFFFu) + (s >> 16);
}

// One thread per coarse cell: 16 fine rows of 16 bytes.
__global__ __launch_bounds__(kBlock) void encode_kernel(
        const uint8_t *__restrict__ fine, long long cells, int h, int w,
        uint32_t mask, float *__restrict__ x, float *__restrict__ mean) {
    const long long cell = blockIdx.x * (long long)kBlock + threadIdx.x;
    if (cell >= cells) return;
    const long long hw = (long long)h * w;
    const long long b = cell / hw;
    const int yy = static_cast<int>((cell % hw) / w), xx = static_cast<int>(cell % w);
    const long long fine_w = 16LL * w;
    const uint8_t *base = fine + (b * 16 * h + 16LL * yy) * fine_w + 16LL * xx;
    const uint32_t mask4 = mask * 0x01010101u;
    uint32_t sums[16];
#pragma unroll
    for (int c = 0; c < 16; ++c) sums[c] = 0;
#pragma unroll
    for (int r = 0; r < 16; ++r) {
        const uint4 v = *reinterpret_cast<const uint4 *>(base + r * fine_w);
        const int sy = r >> 2;
        sums[4 * sy + 0] += byte_sum(v.x & mask4);
        sums[4 * sy + 1] += byte_sum(v.y & mask4);
        sums[4 * sy + 2] += byte_sum(v.z & mask4);
        sums[4 * sy + 3] += byte_sum(v.w & mask4);
    }
    float4 *out = reinterpret_cast<float4 *>(x + cell * kC);
    uint32_t total = 0;
#pragma unroll
    for (int q = 0; q < 4; ++q) {
        out[q] = make_float4(sums[4 * q] * (1.0f / 16), sums[4 * q + 1] * (1.0f / 16),
                             sums[4 * q + 2] * (1.0f / 16), sums[4 * q + 3] * (1.0f / 16));
        total += sums[4 * q] + sums[4 * q + 1] + sums[4 * q + 2] + sums[4 * q + 3];
    }
    if (mean) mean[cell] = total * (1.0f / 256);
}

struct Cell {
    long long index;  // flat coarse cell of this lane
    long long b;
    int y, x;
};

__device__ __forceinline__ Cell lane_cell(long long chunk, int lane, int h, int w) {
    Cell c;
    c.index = chunk * 16 + (lane & 15);
    const long long hw = (long long)h * w;
    c.b = c.index / hw;
    const long long rem = c.index % hw;
    c.y = static_cast<int>(rem / w);
    c.x = static_cast<int>(rem % w);
    return c;
}

// Depthwise 3x3 of the lane's four channels; keeps the neighbour values.
__device__ __forceinline__ void perceive(const float *__restrict__ x, const Cell &c, int g,
                                         int h, int w, const float *__restrict__ dw,
                                         const float *__restrict__ bdw, float p[4],
                                         float4 nb[kTaps]) {
    // the nine torus-wrapped neighbour addresses first, then all nine loads,
    // then the taps in the order dy-major, dx-minor (so that the loads are
    // in flight together rather than one or two at a time)
    long long row[3];
    int col[3];
#pragma unroll
    for (int d = 0; d < 3; ++d) {
        int yn = c.y + d - 1;
        yn = yn < 0 ? yn + h : (yn >= h ? yn - h : yn);
        row[d] = (c.b * h + yn) * (long long)w;
        int xn = c.x + d - 1;
        col[d] = xn < 0 ? xn + w : (xn >= w ? xn - w : xn);
    }
#pragma unroll
    for (int t = 0; t < kTaps; ++t)
        nb[t] = *reinterpret_cast<const float4 *>(x + (row[t / 3] + col[t % 3]) * kC + 4 * g);
#pragma unroll
    for (int i = 0; i < 4; ++i) p[i] = bdw[4 * g + i];
#pragma unroll
    for (int t = 0; t < kTaps; ++t) {
        const float4 v = nb[t];
        p[0] = fmaf(dw[(4 * g + 0) * kTaps + t], v.x, p[0]);
        p[1] = fmaf(dw[(4 * g + 1) * kTaps + t], v.y, p[1]);
        p[2] = fmaf(dw[(4 * g + 2) * kTaps + t], v.z, p[2]);
        p[3] = fmaf(dw[(4 * g + 3) * kTaps + t], v.w, p[3]);
    }
}

__device__ __forceinline__ bf16x4 pack4(float a, float b, float c, float d) {
    bf16x4 r;
    r[0] = to_bf16(a);
    r[1] = to_bf16(b);
    r[2] = to_bf16(c);
    r[3] = to_bf16(d);
    return r;
}

__device__ __forceinline__ f32x4 load_bias(const float *__restrict__ b, int off) {
    f32x4 r;
    r[0] = b[off];
    r[1] = b[off + 1];
    r[2] = b[off + 2];
    r[3] = b[off + 3];
    return r;
}

// The 1x1 MLP of one 16-cell chunk from its bf16 perception (lane: cell m,
// channels 4g..4g+3), and the stores of y and pred (either may be null).
__device__ __forceinline__ f32x4 mlp(bf16x4 pb, int g, int m, const short *__restrict__ w1,
                                     const float *__restrict__ b1, const short *__restrict__ w2,
                                     const float *__restrict__ b2) {
    f32x4 out = load_bias(b2, 4 * g);
#pragma unroll
    for (int pair = 0; pair < 2; ++pair) {
        bf16x4 hb[2], a2[2];
#pragma unroll
        for (int half = 0; half < 2; ++half) {
            const int blk = 2 * pair + half;
            const bf16x4 a1 =
                *reinterpret_cast<const bf16x4 *>(w1 + (16 * blk + m) * kC + 4 * g);
            const f32x4 z = mfma(a1, pb, load_bias(b1, 16 * blk + 4 * g));
            hb[half] = pack4(fmaxf(z[0], 0.f), fmaxf(z[1], 0.f), fmaxf(z[2], 0.f),
                             fmaxf(z[3], 0.f));
            a2[half] = *reinterpret_cast<const bf16x4 *>(w2 + m * kHid + 16 * blk + 4 * g);
        }
        out = mfma32(a2[0], a2[1], hb[0], hb[1], out);
    }
    return out;
}

__device__ __forceinline__ void store_out(f32x4 out, long long index, int g,
                                          float *__restrict__ y, float *__restrict__ pred) {
    if (y)
        *reinterpret_cast<float4 *>(y + index * kC + 4 * g) =
            make_float4(out[0], out[1], out[2], out[3]);
    if (pred) {
        float s = (out[0] + out[1]) + (out[2] + out[3]);
        s += __shfl_xor(s, 16);
        s += __shfl_xor(s, 32);
        if (g == 0) pred[index] = s * (1.0f / 16);
    }
}

// One wavefront per chunk of 16 coarse cells, grid-stride over chunks.
__global__ __launch_bounds__(kBlock) void forward_kernel(
        const float *__restrict__ x, long long chunks, int h, int w,
        const float *__restrict__ dw, const float *__restrict__ bdw,
        const short *__restrict__ w1, const float *__restrict__ b1,
        const short *__restrict__ w2, const float *__restrict__ b2,
        float *__restrict__ y, float *__restrict__ pred) {
    const int lane = threadIdx.x & 63;
    const int g = lane >> 4, m = lane & 15;
    const long long stride = (long long)gridDim.x * kWaves;
    for (long long chunk = blockIdx.x * (long long)kWaves + (threadIdx.x >> 6);
         chunk < chunks; chunk += stride) {
        const Cell c = lane_cell(chunk, lane, h, w);
        float p[4];
        float4 nb[kTaps];
        perceive(x, c, g, h, w, dw, bdw, p, nb);
        const bf16x4 pb = pack4(p[0], p[1], p[2], p[3]);
        store_out(mlp(pb, g, m, w1, b1, w2, b2), c.index, g, y, pred);
    }
}

// Fine field to prediction in one launch (the time-to-coarse-field path).
// A workgroup owns kPredH x 16 coarse cells: it encodes them and their
// one-cell torus halo into LDS (work item = coarse cell x sub-block row,
// four 16-byte fine rows each), then every wave runs the forward of whole
// tile rows (16 cells = one MFMA chunk) from LDS.  The sums, the tap order
// of the perception and the MLP are those of encode_kernel and
// forward_kernel, so y and pred equal theirs bit for bit.
constexpr int kPredH = 8;
constexpr int kPredW = 16;

__global__ __launch_bounds__(kBlock) void predict_kernel(
        const uint8_t *__restrict__ fine, long long batch, int h, int w, uint32_t mask,
        const float *__restrict__ dw, const float *__restrict__ bdw,
        const short *__restrict__ w1, const float *__restrict__ b1,
        const short *__restrict__ w2, const float *__restrict__ b2,
        float *__restrict__ y, float *__restrict__ pred) {
    constexpr int RW = kPredW + 2, RH = kPredH + 2, RC = RW * RH;
    __shared__ float4 xs[RC * 4];  // [region cell][channel group]
    const int tiles_x = w / kPredW;
    const long long tiles = static_cast<long long>(tiles_x) * (h / kPredH);
    const long long fine_w = 16LL * w;
    const uint32_t mask4 = mask * 0x01010101u;
    const int lane = threadIdx.x & 63, wave = threadIdx.x >> 6;
    const int g = lane >> 4, m = lane & 15;
    for (long long t = blockIdx.x; t < batch * tiles; t += gridDim.x) {
        const long long b = t / tiles;
        const int tt = static_cast<int>(t % tiles);
        const int ty0 = (tt / tiles_x) * kPredH, tx0 = (tt % tiles_x) * kPredW;
        for (int it = threadIdx.x; it < RC * 4; it += kBlock) {
            const int rc = it >> 2, sy = it & 3;
            const int ry = rc / RW, rx = rc - ry * RW;
            int cy = ty0 - 1 + ry, cx = tx0 - 1 + rx;
            cy = cy < 0 ? cy + h : (cy >= h ? cy - h : cy);
            cx = cx < 0 ? cx + w : (cx >= w ? cx - w : cx);
            const uint8_t *base = fine + ((b * h + cy) * 16 + 4 * sy) * fine_w + 16LL * cx;
            uint32_t s0 = 0, s1 = 0, s2 = 0, s3 = 0;
#pragma unroll
            for (int r = 0; r < 4; ++r) {
                const uint4 v = *reinterpret_cast<const uint4 *>(base + r * fine_w);
                s0 += byte_sum(v.x & mask4);
                s1 += byte_sum(v.y & mask4);
                s2 += byte_sum(v.z & mask4);
                s3 += byte_sum(v.w & mask4);
            }
            xs[rc * 4 + sy] = make_float4(s0 * (1.0f / 16), s1 * (1.0f / 16),
                                          s2 * (1.0f / 16), s3 * (1.0f / 16));
        }
        __syncthreads();
        for (int r = wave; r < kPredH; r += kWaves) {
            float p[4];
#pragma unroll
            for (int i = 0; i < 4; ++i) p[i] = bdw[4 * g + i];
#pragma unroll
            for (int dy = -1; dy <= 1; ++dy)
#pragma unroll
                for (int dx = -1; dx <= 1; ++dx) {
                    const int tap = (dy + 1) * 3 + (dx + 1);
                    const float4 v = xs[((r + 1 + dy) * RW + (m + 1 + dx)) * 4 + g];
                    p[0] = fmaf(dw[(4 * g + 0) * kTaps + tap], v.x, p[0]);
                    p[1] = fmaf(dw[(4 * g + 1) * kTaps + tap], v.y, p[1]);
                    p[2] = fmaf(dw[(4 * g + 2) * kTaps + tap], v.z, p[2]);
                    p[3] = fmaf(dw[(4 * g + 3) * kTaps + tap], v.w, p[3]);
                }
            const bf16x4 pb = pack4(p[0], p[1], p[2], p[3]);
            const long long index = (b * h + ty0 + r) * static_cast<long long>(w) + tx0 + m;
            store_out(mlp(pb, g, m, w1, b1, w2, b2), index, g, y, pred);
        }
        __syncthreads();
    }
}

// Per-wave LDS tiles of the backward pass, feature-major ([feature][cell]).
struct WaveTiles {
    short dot[kC * 16];     // dY^T  (output x cell)
    short ht[kHid * 16];    // relu(z1)^T (hidden x cell)
    short dzt[kHid * 16];   // dZ1^T (hidden x cell)
    short pt[kC * 16];      // P^T   (input x cell)
};

__device__ __forceinline__ float wave_sum16(float v) {
    // sum over the 16 lanes that share lane / 16
    v += __shfl_xor(v, 1);
    v += __shfl_xor(v, 2);
    v += __shfl_xor(v, 4);
    v += __shfl_xor(v, 8);
    return v;
}

// Backward: every wave accumulates the gradient of its chunks in registers;
// the workgroup adds its waves in LDS and writes one row of `partials`
// (kParams floats); reduce_kernel sums the rows in a fixed order.
__global__ __launch_bounds__(kBlock) void backward_kernel(
        const float *__restrict__ x, const float *__restrict__ dy, long long chunks, int h,
        int w, const float *__restrict__ dw, const float *__restrict__ bdw,
        const short *__restrict__ w1, const float *__restrict__ b1,
        const short *__restrict__ w2, const float *__restrict__ b2,
        float *__restrict__ partials) {
    __shared__ WaveTiles tiles[kWaves];
    const int wave = threadIdx.x >> 6;
    const int lane = threadIdx.x & 63;
    const int g = lane >> 4, m = lane & 15;
    WaveTiles &t = tiles[wave];

    float g_dw[4][kTaps], g_bdw[4], g_b1[4][4], g_b2[4];
    f32x4 g_w1[4], g_w2[4];
#pragma unroll
    for (int i = 0; i < 4; ++i) {
        g_bdw[i] = g_b2[i] = 0.f;
#pragma unroll
        for (int k = 0; k < kTaps; ++k) g_dw[i][k] = 0.f;
#pragma unroll
        for (int blk = 0; blk < 4; ++blk) g_b1[blk][i] = 0.f;
        g_w1[i] = f32x4{0.f, 0.f, 0.f, 0.f};
        g_w2[i] = f32x4{0.f, 0.f, 0.f, 0.f};
    }

    const long long stride = (long long)gridDim.x * kWaves;
    for (long long chunk = blockIdx.x * (long long)kWaves + wave; chunk < chunks;
         chunk += stride) {
        const Cell c = lane_cell(chunk, lane, h, w);
        float p[4];
        float4 nb[kTaps];
        perceive(x, c, g, h, w, dw, bdw, p, nb);
        const bf16x4 pb = pack4(p[0], p[1], p[2], p[3]);
        const float4 d4 = *reinterpret_cast<const float4 *>(dy + c.index * kC + 4 * g);
        const bf16x4 db = pack4(d4.x, d4.y, d4.z, d4.w);
        g_b2[0] += d4.x;
        g_b2[1] += d4.y;
        g_b2[2] += d4.z;
        g_b2[3] += d4.w;
        f32x4 dp = {0.f, 0.f, 0.f, 0.f};
        bf16x4 a1t_lo, dzb_lo;  // the even block's operands, until the odd one
#pragma unroll
        for (int i = 0; i < 4; ++i) {
            t.dot[(4 * g + i) * 16 + m] = db[i];
            t.pt[(4 * g + i) * 16 + m] = pb[i];
        }
#pragma unroll
        for (int blk = 0; blk < 4; ++blk) {
            const bf16x4 a1 = *reinterpret_cast<const bf16x4 *>(w1 + (16 * blk + m) * kC + 4 * g);
            const f32x4 z = mfma(a1, pb, load_bias(b1, 16 * blk + 4 * g));
            const bf16x4 hb = pack4(fmaxf(z[0], 0.f), fmaxf(z[1], 0.f), fmaxf(z[2], 0.f),
                                    fmaxf(z[3], 0.f));
            // dH^T block = W2^T[block] . dY^T: A[m=hidden][k=out] = W2[k][16 blk + m]
            bf16x4 a2t;
#pragma unroll
            for (int i = 0; i < 4; ++i) a2t[i] = w2[(4 * g + i) * kHid + 16 * blk + m];
            const f32x4 dh = mfma(a2t, db, f32x4{0.f, 0.f, 0.f, 0.f});
            float dz[4];
#pragma unroll
            for (int i = 0; i < 4; ++i) {
                dz[i] = z[i] > 0.f ? dh[i] : 0.f;
                g_b1[blk][i] += dz[i];
            }
            const bf16x4 dzb = pack4(dz[0], dz[1], dz[2], dz[3]);
            // dP^T += W1[blocks]^T . dZ1^T over two hidden blocks at a time
            // (K = 32): A[m=in][k=hidden] = W1[16 blk + k][m]
            bf16x4 a1t;
#pragma unroll
            for (int i = 0; i < 4; ++i) a1t[i] = w1[(16 * blk + 4 * g + i) * kC + m];
            if (blk & 1)
                dp = mfma32(a1t_lo, a1t, dzb_lo, dzb, dp);
            a1t_lo = a1t;
            dzb_lo = dzb;
#pragma unroll
            for (int i = 0; i < 4; ++i) {
                t.ht[(16 * blk + 4 * g + i) * 16 + m] = hb[i];
                t.dzt[(16 * blk + 4 * g + i) * 16 + m] = dzb[i];
            }
        }
        // depthwise weights and bias: dP (channel 4g+i of this lane's cell)
#pragma unroll
        for (int k = 0; k < kTaps; ++k) {
            g_dw[0][k] = fmaf(dp[0], nb[k].x, g_dw[0][k]);
            g_dw[1][k] = fmaf(dp[1], nb[k].y, g_dw[1][k]);
            g_dw[2][k] = fmaf(dp[2], nb[k].z, g_dw[2][k]);
            g_dw[3][k] = fmaf(dp[3], nb[k].w, g_dw[3][k]);
        }
#pragma unroll
        for (int i = 0; i < 4; ++i) g_bdw[i] += dp[i];
        wave_sync();
        // dW2[out][hid] += dY^T . H over the 16 cells (K = cells)
        const bf16x4 a_do = *reinterpret_cast<const bf16x4 *>(&t.dot[m * 16 + 4 * g]);
        const bf16x4 b_p = *reinterpret_cast<const bf16x4 *>(&t.pt[m * 16 + 4 * g]);
#pragma unroll
        for (int blk = 0; blk < 4; ++blk) {
            const bf16x4 b_h =
                *reinterpret_cast<const bf16x4 *>(&t.ht[(16 * blk + m) * 16 + 4 * g]);
            g_w2[blk] = mfma(a_do, b_h, g_w2[blk]);
            // dW1[hid][in] += dZ1^T . P over the cells
            const bf16x4 a_dz =
                *reinterpret_cast<const bf16x4 *>(&t.dzt[(16 * blk + m) * 16 + 4 * g]);
            g_w1[blk] = mfma(a_dz, b_p, g_w1[blk]);
        }
        wave_sync();
    }

    // sums over the 16 cells of a lane group, in every wave at once
#pragma unroll
    for (int i = 0; i < 4; ++i) {
#pragma unroll
        for (int k = 0; k < kTaps; ++k) g_dw[i][k] = wave_sum16(g_dw[i][k]);
        g_bdw[i] = wave_sum16(g_bdw[i]);
        g_b2[i] = wave_sum16(g_b2[i]);
#pragma unroll
        for (int blk = 0; blk < 4; ++blk) g_b1[blk][i] = wave_sum16(g_b1[blk][i]);
    }
    // combine the waves of the workgroup in LDS in wave order, then write
    // one row of partial sums per workgroup
    __shared__ float wg_grad[kParams];
    for (int turn = 0; turn < kWaves; ++turn) {
        if (wave == turn) {
            auto put = [&](int idx, float v) {
                if (turn == 0)
                    wg_grad[idx] = v;
                else
                    wg_grad[idx] += v;
            };
#pragma unroll
            for (int i = 0; i < 4; ++i) {
                if (m == 0) {
#pragma unroll
                    for (int k = 0; k < kTaps; ++k) put(kOffDw + (4 * g + i) * kTaps + k, g_dw[i][k]);
                    put(kOffBdw + 4 * g + i, g_bdw[i]);
                    put(kOffB2 + 4 * g + i, g_b2[i]);
#pragma unroll
                    for (int blk = 0; blk < 4; ++blk) put(kOffB1 + 16 * blk + 4 * g + i, g_b1[blk][i]);
                }
#pragma unroll
                for (int blk = 0; blk < 4; ++blk) {
                    // D of dW2 tile: [out 4g+i][hid 16 blk + m]; of dW1:
                    // [hid 16 blk + 4g+i][in m]
                    put(kOffW2 + (4 * g + i) * kHid + 16 * blk + m, g_w2[blk][i]);
                    put(kOffW1 + (16 * blk + 4 * g + i) * kC + m, g_w1[blk][i]);
                }
            }
        }
        __syncthreads();
    }
    float *row = partials + (long long)blockIdx.x * kParams;
    for (int q = threadIdx.x; q < kParams; q += kBlock) row[q] = wg_grad[q];
}

// grads[p] = sum of partials[r][p] over the rows: 16 parameters x 16 row
// slices per workgroup, the slices added in a fixed tree order.
constexpr int kRedParams = 16;
constexpr int kRedSlices = kBlock / kRedParams;

__global__ __launch_bounds__(kBlock) void reduce_kernel(const float *__restrict__ partials,
                                                        long long rows,
                                                        float *__restrict__ grads) {
    __shared__ float part[kRedSlices][kRedParams];
    const int lp = threadIdx.x % kRedParams, slice = threadIdx.x / kRedParams;
    const int p = blockIdx.x * kRedParams + lp;
    float s = 0.f;
    if (p < kParams)
        for (long long r = slice; r < rows; r += kRedSlices) s += partials[r * kParams + p];
    part[slice][lp] = s;
    __syncthreads();
    for (int half = kRedSlices / 2; half > 0; half /= 2) {
        if (slice < half) part[slice][lp] += part[slice + half][lp];
        __syncthreads();
    }
    if (slice == 0 && p < kParams) grads[p] = part[0][lp];
}

// Workspace of xen_nca_backward_noalloc: the partial-gradient rows of
// backward_kernel (one per workgroup; launches never use more than
// kScratchRows workgroups).  Every GPU the library runs on holds its own
// copy (2.3 MB) until the process ends.  Host code must not pass the
// array's name to a kernel: it stands for the host-side shadow, not a
// device address; the device address comes from hipGetSymbolAddress.
constexpr int kScratchRows = 256;
constexpr int kMaxDevices = 64;
static __device__ __align__(256) float scratch[kScratchRows * kParams];

// Device address of `scratch` on the current device, looked up once per
// device (the call also makes the module's globals resident there).
int scratch_address(float **out) {
    static std::mutex lock;
    static float *addr[kMaxDevices] = {};
    int dev = 0;
    if (int rc = xen::check(hipGetDevice(&dev), "hipGetDevice")) return rc;
    if (dev < 0 || dev >= kMaxDevices)
        return xen::fail("xen_nca_backward_noalloc: device index out of range", 22);
    std::lock_guard<std::mutex> guard(lock);
    if (!addr[dev]) {
        void *ptr = nullptr;
        if (int rc = xen::check(hipGetSymbolAddress(&ptr, HIP_SYMBOL(scratch)),
                                "hipGetSymbolAddress(scratch)"))
            return rc;
        addr[dev] = static_cast<float *>(ptr);
    }
    *out = addr[dev];
    return 0;
}

int check_grid(long long batch, int h, int w, long long *chunks) {
    if (batch < 1 || h < 1 || w < 1)
        return xen::fail("xen_nca: batch, h and w must be positive", 22);
    const long long cells = batch * h * (long long)w;
    if (cells % 16) return xen::fail("xen_nca: batch * h * w must be a multiple of 16", 22);
    if (cells > (1LL << 40)) return xen::fail("xen_nca: too many coarse cells", 22);
    *chunks = cells / 16;
    return 0;
}

int wave_groups(long long chunks, long long cap) {
    long long groups = (chunks + kWaves - 1) / kWaves;
    if (groups > cap) groups = cap;
    return static_cast<int>(groups < 1 ? 1 : groups);
}

}  // namespace

extern "C" {

int xen_nca_encode(const uint8_t *fine, long long batch, int h, int w, uint32_t value_mask,
                   float *x, float *block_mean, xen_stream stream) {
    long long chunks;
    if (int rc = check_grid(batch, h, w, &chunks)) return rc;
    if (value_mask > 255) return xen::fail("xen_nca_encode: value_mask must fit a byte", 22);
    const long long cells = chunks * 16;
    hipLaunchKernelGGL(encode_kernel, dim3(static_cast<unsigned>((cells + kBlock - 1) / kBlock)),
                       dim3(kBlock), 0, reinterpret_cast<hipStream_t>(stream), fine, cells, h, w,
                       value_mask, x, block_mean);
    return xen::check(hipGetLastError(), "encode_kernel launch");
}

int xen_nca_forward(const float *x, long long batch, int h, int w, const float *dw,
                    const float *bdw, const uint16_t *w1, const float *b1, const uint16_t *w2,
                    const float *b2, float *y, float *pred, xen_stream stream) {
    long long chunks;
    if (int rc = check_grid(batch, h, w, &chunks)) return rc;
    hipLaunchKernelGGL(forward_kernel, dim3(wave_groups(chunks, 256LL * 16)), dim3(kBlock), 0,
                       reinterpret_cast<hipStream_t>(stream), x, chunks, h, w, dw, bdw,
                       reinterpret_cast<const short *>(w1), b1,
                       reinterpret_cast<const short *>(w2), b2, y, pred);
    return xen::check(hipGetLastError(), "forward_kernel launch");
}

int xen_nca_predict(const uint8_t *fine, long long batch, int h, int w, uint32_t value_mask,
                    const float *dw, const float *bdw, const uint16_t *w1, const float *b1,
                    const uint16_t *w2, const float *b2, float *y, float *pred,
                    xen_stream stream) {
    if (batch < 1 || h < 1 || w < 1) return xen::fail("xen_nca_predict: empty grid", 22);
    if (h % kPredH || w % kPredW)
        return xen::fail("xen_nca_predict: h must be a multiple of 8 and w of 16", 22);
    if (value_mask > 255) return xen::fail("xen_nca_predict: value_mask must fit a byte", 22);
    if (batch * h * static_cast<long long>(w) > (1LL << 40))
        return xen::fail("xen_nca_predict: too many coarse cells", 22);
    const long long tiles = batch * (h / kPredH) * static_cast<long long>(w / kPredW);
    const unsigned groups = static_cast<unsigned>(tiles < 256LL * 8 ? tiles : 256LL * 8);
    hipLaunchKernelGGL(predict_kernel, dim3(groups), dim3(kBlock), 0,
                       reinterpret_cast<hipStream_t>(stream), fine, batch, h, w, value_mask, dw,
                       bdw, reinterpret_cast<const short *>(w1), b1,
                       reinterpret_cast<const short *>(w2), b2, y, pred);
    return xen::check(hipGetLastError(), "predict_kernel launch");
}

long long xen_nca_partial_rows(long long batch, int h, int w) {
    long long chunks;
    if (check_grid(batch, h, w, &chunks)) return -1;
    return wave_groups(chunks, 256);
}

int xen_nca_backward(const float *x, const float *dy, long long batch, int h, int w,
                     const float *dw, const float *bdw, const uint16_t *w1, const float *b1,
                     const uint16_t *w2, const float *b2, float *partials,
                     long long partial_rows, float *grads, xen_stream stream) {
    long long chunks;
    if (int rc = check_grid(batch, h, w, &chunks)) return rc;
    const int groups = wave_groups(chunks, 256);
    if (partial_rows < groups)
        return xen::fail("xen_nca_backward: partials holds fewer rows than "
                         "xen_nca_partial_rows() asks for", 22);
    hipStream_t s = reinterpret_cast<hipStream_t>(stream);
    hipLaunchKernelGGL(backward_kernel, dim3(groups), dim3(kBlock), 0, s, x, dy, chunks, h, w,
                       dw, bdw, reinterpret_cast<const short *>(w1), b1,
                       reinterpret_cast<const short *>(w2), b2, partials);
    if (int rc = xen::check(hipGetLastError(), "backward_kernel launch")) return rc;
    hipLaunchKernelGGL(reduce_kernel, dim3((kParams + kRedParams - 1) / kRedParams),
                       dim3(kBlock), 0, s, partials, static_cast<long long>(groups), grads);
    return xen::check(hipGetLastError(), "reduce_kernel launch");
}

int xen_nca_backward_noalloc(const float *x, const float *dy, long long batch, int h, int w,
                             const float *dw, const float *bdw, const uint16_t *w1,
                             const float *b1, const uint16_t *w2, const float *b2,
                             float *grads, xen_stream stream) {
    long long chunks;
    if (int rc = check_grid(batch, h, w, &chunks)) return rc;
    const long long rows = xen_nca_partial_rows(batch, h, w);
    if (rows < 1 || rows > kScratchRows)
        return xen::fail("xen_nca_backward_noalloc: the grid needs " + std::to_string(rows) +
                             " workspace rows, more than the " +
                             std::to_string(kScratchRows) + " the static workspace holds",
                         22);
    float *partials = nullptr;
    if (int rc = scratch_address(&partials)) return rc;
    // every row the reduction reads is written in full by backward_kernel
    // first, so the workspace needs no clearing between calls
    return xen_nca_backward(x, dy, batch, h, w, dw, bdw, w1, b1, w2, b2, partials, rows, grads,
                            stream);
}

}  // extern "C"
