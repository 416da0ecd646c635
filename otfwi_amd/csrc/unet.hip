// Kernels of the learned preconditioner (libotfwi_unet.so): a 4-level U-Net
// with BF16 weights and activations and FP32 accumulation on the gfx950
// matrix cores, forward and backward, plus Adam.
//
// Tensors are NHWC ([n][y][x][c], channels innermost, BF16).  A 3x3 "same"
// convolution is an implicit GEMM: M = 256 output pixels of a TY x TX tile,
// N = TN output channels, K = 9 taps x Cin, on v_mfma_f32_32x32x16_bf16
// (A = 32 pixels x 16 channels, B = 16 channels x 32 output channels).  The
// input patch (tile + 1-pixel halo) and the weights of one CK-channel chunk
// are staged in LDS once and read by all 9 taps.  The loader reads the
// decoder's concatenation [skip, nearest-upsample(coarse)] directly, so
// neither the concatenated nor the upsampled tensor is ever written.
// The data gradient of a convolution is the same kernel with the weights
// flipped and transposed; the weight gradient is a second MFMA kernel with
// K = pixels.
#include <hip/hip_runtime.h>
#include <stdint.h>
#include <algorithm>
#include <atomic>
#include <cmath>
#include <cstdio>
#include <string>

#include "otfwi_unet.h"

typedef __bf16 bf16;
typedef __bf16 bf16x8 __attribute__((ext_vector_type(8)));
typedef float f32x16 __attribute__((ext_vector_type(16)));
typedef uint32_t u32x4 __attribute__((ext_vector_type(4)));
typedef uint32_t u32x2 __attribute__((ext_vector_type(2)));

namespace {

thread_local std::string g_uerr;

int ufail(const char *msg)
{
    g_uerr = msg;
    return -1;
}

int hipck(hipError_t e)
{
    if (e == hipSuccess) return 0;
    g_uerr = hipGetErrorString(e);
    return -2;
}

inline unsigned blocks(size_t n, unsigned b = 256) { return (unsigned)((n + b - 1) / b); }

// ---------------------------------------------------------------------------
// input of a convolution: [A | B] along channels; B at half resolution when
// b_half (nearest upsample), zero outside the image and beyond Cin
// ---------------------------------------------------------------------------
struct Src {
    const bf16 *A; int CA;
    const bf16 *B; int CB; int b_half;
    int H, W;
};

__device__ inline u32x4 load8(const Src &s, int n, int y, int x, int ch)
{
    const u32x4 z = {0u, 0u, 0u, 0u};
    if (y < 0 || y >= s.H || x < 0 || x >= s.W) return z;
    if (ch < s.CA)
        return *(const u32x4 *)(s.A + (((size_t)n * s.H + y) * s.W + x) * s.CA + ch);
    const int c = ch - s.CA;
    if (c >= s.CB) return z;
    if (s.b_half) {
        const int Hb = s.H >> 1, Wb = s.W >> 1;
        return *(const u32x4 *)(s.B + (((size_t)n * Hb + (y >> 1)) * Wb + (x >> 1)) * s.CB + c);
    }
    return *(const u32x4 *)(s.B + (((size_t)n * s.H + y) * s.W + x) * s.CB + c);
}

struct ConvArgs {
    Src in;
    const bf16 *Wt;      // [Cout][9][Cin], tap = 3*dy + dx
    const float *bias;   // [Cout] or null
    const bf16 *mask;    // [N][H][W][Cout] or null: zero the output where mask <= 0
    bf16 *Y;             // [N][H][W][Cout]
    int Cout, relu;
};

template <int TX, int TN, int CK>
struct ConvShape {
    static constexpr int TY = 256 / TX;
    static constexpr int PW = TX + 2, PH = TY + 2;
    static constexpr int PS = CK + 8;   // LDS stride of one patch pixel (bf16)
    static constexpr int WS = CK + 8;   // LDS stride of one weight row
    static constexpr size_t in_bytes = ((size_t)PH * PW * PS + (size_t)9 * TN * WS) * 2;
    static constexpr size_t out_bytes = (size_t)TN * 512;   // output tile, channel-major
    static constexpr size_t lds_bytes = in_bytes > out_bytes ? in_bytes : out_bytes;
};

// Epilogue of a 256-pixel x TN-channel output tile.  The tile goes to LDS
// channel-major (row c: 256 BF16 pixels, 512 bytes), four consecutive
// pixels of one channel per 8-byte store, slot qi (pixels 4 qi .. 4 qi + 3)
// of row c at qi ^ ot_key(c): the 32 lanes of a 64-bit store (32 channels,
// one slot) and of a 64-bit load in conv_tile_out (8 channel rows' parts x 4
// slots, or 4 x 8 for 32 channels) then hit 32 different 8-byte bank slots.
__device__ __forceinline__ int ot_key(int c) { return (c & 31) ^ ((c >> 5) << 2); }

__device__ __forceinline__ unsigned pack_bf16(float lo, float hi)
{
    typedef __bf16 bf16x2 __attribute__((ext_vector_type(2)));
    const bf16x2 v = {(bf16)lo, (bf16)hi};
    return __builtin_bit_cast(unsigned, v);
}

// Accumulators + bias -> ReLU -> BF16 tile (the per-element arithmetic of a
// plain [pixel][channel] epilogue, so results are unchanged).  C layout of the
// 32x32 MFMA: lane holds column (lane % 32) and rows 8*(e/4) + 4*(lane/32) +
// e%4, so e = 4g .. 4g+3 are four consecutive pixels of one channel.
template <int NJ>
__device__ __forceinline__ void conv_tile_stage(const f32x16 (&acc)[2][NJ], const float *bias,
                                                bool relu, bf16 *ot, int wave, int lane)
{
    const int r = lane & 31, h = lane >> 5;
#pragma unroll
    for (int j = 0; j < NJ; j++) {
        const int c = j * 32 + r;
        const float b = bias ? bias[c] : 0.0f;
#pragma unroll
        for (int i = 0; i < 2; i++)
#pragma unroll
            for (int g = 0; g < 4; g++) {
                const int qi = wave * 16 + i * 8 + 2 * g + h;
                float v[4];
#pragma unroll
                for (int e = 0; e < 4; e++) {
                    v[e] = acc[i][j][4 * g + e] + b;
                    if (relu) v[e] = fmaxf(v[e], 0.0f);
                }
                *(u32x2 *)(ot + c * 256 + (qi ^ ot_key(c)) * 4) =
                    u32x2{pack_bf16(v[0], v[1]), pack_bf16(v[2], v[3])};
            }
    }
}

// The staged tile out to Y in 16-byte pieces of 8 channels (coalesced: the
// lanes of a unit group cover a pixel's TN channels), zeroed where the ReLU
// mask of a data-gradient convolution is not > 0.  A unit (4 pixels x 8
// channels) loads 8 channel rows x 8 bytes and transposes them in
// registers (v_perm) into the 4 pixel pieces.
template <int TX, int TN>
__device__ __forceinline__ void conv_tile_out(const bf16 *ot, const ConvArgs &a, int n, int tx0,
                                              int ty0, int co0, int tid)
{
    constexpr int PV8 = TN / 8;
    const int H = a.in.H, W = a.in.W;
#pragma unroll
    for (int u0 = 0; u0 < 64 * PV8; u0 += 256) {
        const int u = u0 + tid;
        const int part = u % PV8, qi = u / PV8;
        u32x2 rw[8];
#pragma unroll
        for (int k = 0; k < 8; k++) {
            const int c = part * 8 + k;
            rw[k] = *(const u32x2 *)(ot + c * 256 + (qi ^ ot_key(c)) * 4);
        }
#pragma unroll
        for (int p = 0; p < 4; p++) {
            const int q = 4 * qi + p;
            const int y = ty0 + q / TX, x = tx0 + q % TX;
            const size_t o = (((size_t)n * H + y) * W + x) * a.Cout + co0 + part * 8;
            const unsigned sel = (p & 1) ? 0x07060302u : 0x05040100u;
            u32x4 val;
#pragma unroll
            for (int d = 0; d < 4; d++)
                val[d] = __builtin_amdgcn_perm(p < 2 ? rw[2 * d + 1].x : rw[2 * d + 1].y,
                                               p < 2 ? rw[2 * d].x : rw[2 * d].y, sel);
            if (a.mask) {
                // zero where the mask is not > 0 (zero, negative or NaN)
                const bf16x8 mk = __builtin_bit_cast(bf16x8, *(const u32x4 *)(a.mask + o));
#pragma unroll
                for (int k = 0; k < 4; k++) {
                    const unsigned lo = (float)mk[2 * k] > 0.0f ? 0x0000FFFFu : 0u;
                    const unsigned hi = (float)mk[2 * k + 1] > 0.0f ? 0xFFFF0000u : 0u;
                    val[k] &= lo | hi;
                }
            }
            *(u32x4 *)(a.Y + o) = val;
        }
    }
}

// Tap loop of conv3x3_kernel unrolled by 3 (one kernel row per unrolled
// body: the next tap's fragments load during this tap's MFMAs).  Measured at
// 8 x 1024^2: training step 17.60 / 17.43 / 17.64 ms and forward 4.17 / 4.09 /
// 4.19 ms for 1 / 3 / 9.
#ifndef CONV_MAIN_TAP_UNROLL
#define CONV_MAIN_TAP_UNROLL 3
#endif
template <int TX, int TN, int CK>
__global__ __launch_bounds__(256) void conv3x3_kernel(ConvArgs a)
{
    using S = ConvShape<TX, TN, CK>;
    constexpr int TY = S::TY, PW = S::PW, PH = S::PH, PS = S::PS, WS = S::WS;
    constexpr int NJ = TN / 32, PV = CK / 8;
    extern __shared__ __attribute__((aligned(16))) bf16 lds[];
    bf16 *patch = lds;
    bf16 *wl = lds + PH * PW * PS;

    const int tid = threadIdx.x, lane = tid & 63, wave = tid >> 6;
    const int W = a.in.W;
    const int tiles_x = W / TX;
    const int tx0 = (blockIdx.x % tiles_x) * TX, ty0 = (blockIdx.x / tiles_x) * TY;
    const int co0 = blockIdx.y * TN, n = blockIdx.z;
    const int Cin = a.in.CA + a.in.CB;
    const int r = lane & 31, kh8 = (lane >> 5) * 8;

    f32x16 acc[2][NJ];
#pragma unroll
    for (int i = 0; i < 2; i++)
#pragma unroll
        for (int j = 0; j < NJ; j++)
#pragma unroll
            for (int e = 0; e < 16; e++) acc[i][j][e] = 0.0f;

    int pbase[2];
#pragma unroll
    for (int i = 0; i < 2; i++) {
        const int q = wave * 64 + i * 32 + r;
        pbase[i] = (q / TX) * PW + (q % TX);
    }

    // The next chunk's patch and weights are loaded into registers while the
    // current chunk's MFMAs run, then written to LDS after the barrier.
    constexpr int NPV = (PH * PW * PV + 255) / 256, NWV = (9 * TN * PV + 255) / 256;
    u32x4 rp[NPV], rw[NWV];
    auto fetch = [&](int c0) {
#pragma unroll
        for (int k = 0; k < NPV; k++) {
            const int v = tid + 256 * k;
            if (v < PH * PW * PV) {
                const int pp = v / PV, part = v - pp * PV;
                const int py = pp / PW, px = pp - py * PW;
                rp[k] = load8(a.in, n, ty0 + py - 1, tx0 + px - 1, c0 + part * 8);
            }
        }
#pragma unroll
        for (int k = 0; k < NWV; k++) {
            const int v = tid + 256 * k;
            if (v < 9 * TN * PV) {
                const int row = v / PV, part = v - row * PV;   // row = tap*TN + co
                const int tap = row / TN, co = row - tap * TN;
                rw[k] = *(const u32x4 *)(a.Wt + ((size_t)(co0 + co) * 9 + tap) * Cin + c0 +
                                         part * 8);
            }
        }
    };
    fetch(0);
    for (int c0 = 0; c0 < Cin; c0 += CK) {
#pragma unroll
        for (int k = 0; k < NPV; k++) {
            const int v = tid + 256 * k;
            if (v < PH * PW * PV) {
                const int pp = v / PV, part = v - pp * PV;
                *(u32x4 *)(patch + pp * PS + part * 8) = rp[k];
            }
        }
#pragma unroll
        for (int k = 0; k < NWV; k++) {
            const int v = tid + 256 * k;
            if (v < 9 * TN * PV) {
                const int row = v / PV, part = v - row * PV;
                *(u32x4 *)(wl + row * WS + part * 8) = rw[k];
            }
        }
        __syncthreads();
        if (c0 + CK < Cin) fetch(c0 + CK);
#pragma unroll CONV_MAIN_TAP_UNROLL
        for (int tap = 0; tap < 9; tap++) {
            const int dy = tap / 3, dx = tap - 3 * (tap / 3);
#pragma unroll
            for (int kk = 0; kk < CK; kk += 16) {
                bf16x8 af[2], bfm[NJ];
#pragma unroll
                for (int i = 0; i < 2; i++)
                    af[i] = *(const bf16x8 *)(patch + (pbase[i] + dy * PW + dx) * PS + kk + kh8);
#pragma unroll
                for (int j = 0; j < NJ; j++)
                    bfm[j] = *(const bf16x8 *)(wl + (tap * TN + j * 32 + r) * WS + kk + kh8);
#pragma unroll
                for (int i = 0; i < 2; i++)
#pragma unroll
                    for (int j = 0; j < NJ; j++)
                        acc[i][j] = __builtin_amdgcn_mfma_f32_32x32x16_bf16(af[i], bfm[j],
                                                                            acc[i][j], 0, 0, 0);
            }
        }
        __syncthreads();
    }

    // the output tile reuses the input staging (the loop ended with a barrier)
    bf16 *ot = lds;
    conv_tile_stage<NJ>(acc, a.bias ? a.bias + co0 : nullptr, a.relu != 0, ot, wave, lane);
    __syncthreads();
    conv_tile_out<TX, TN>(ot, a, n, tx0, ty0, co0, tid);
}

template <int TX, int TN, int CK>
int launch_conv(const ConvArgs &a, int N, hipStream_t st)
{
    using S = ConvShape<TX, TN, CK>;
    const void *fn = (const void *)conv3x3_kernel<TX, TN, CK>;
    if (int rc = hipck(hipFuncSetAttribute(fn, hipFuncAttributeMaxDynamicSharedMemorySize,
                                           (int)S::lds_bytes)))
        return rc;
    dim3 grid((a.in.W / TX) * (a.in.H / S::TY), a.Cout / TN, N);
    hipLaunchKernelGGL((conv3x3_kernel<TX, TN, CK>), grid, dim3(256), S::lds_bytes, st, a);
    return hipck(hipGetLastError());
}

// Convolutions whose input channels fit one chunk (Cin == CK: the level-0
// layers of the network and their data gradients).  conv3x3_kernel can only
// load, wait and compute there; here a work-group stages the weights once and
// runs NT consecutive pixel tiles, with the patches of the next CONV_PF tiles
// in flight into registers while the current tile's MFMAs and epilogue run
// (measured at 8 x 1024^2, 32 -> 32 channels: one tile per work-group
// 420-503 us; 4 tiles with 1 patch ahead 341-357 us; 4 with 2 ahead 340-348;
// 8 with 2 ahead 324-332).  The output tile
// has its own LDS region, so two barriers per tile suffice.  Every output is
// the same MFMA sequence as in conv3x3_kernel (bitwise equal results).
template <int TX, int TN, int CK>
struct ConvTiles {
    using S = ConvShape<TX, TN, CK>;
    static constexpr size_t patch_elems = (size_t)S::PH * S::PW * S::PS;
    static constexpr size_t w_elems = (size_t)9 * TN * S::WS;
    static constexpr size_t lds_bytes = (patch_elems + w_elems) * 2 + S::out_bytes;
};

#ifndef CONV_TAP_UNROLL
#define CONV_TAP_UNROLL 9
#endif
#ifndef CONV_PF
#define CONV_PF 2
#endif
template <int TX, int TN, int CK, int NT>
__global__ __launch_bounds__(256) void conv3x3_tiles_kernel(ConvArgs a)
{
    using S = ConvShape<TX, TN, CK>;
    using L = ConvTiles<TX, TN, CK>;
    constexpr int TY = S::TY, PW = S::PW, PH = S::PH, PS = S::PS, WS = S::WS;
    constexpr int NJ = TN / 32, PV = CK / 8;
    extern __shared__ __attribute__((aligned(16))) bf16 lds[];
    bf16 *patch = lds;
    bf16 *wl = lds + L::patch_elems;
    bf16 *ot = wl + L::w_elems;

    const int tid = threadIdx.x, lane = tid & 63, wave = tid >> 6;
    const int W = a.in.W;
    const int tiles_x = W / TX;
    const int co0 = blockIdx.y * TN, n = blockIdx.z;
    const int r = lane & 31, kh8 = (lane >> 5) * 8;
    int pbase[2];
#pragma unroll
    for (int i = 0; i < 2; i++) {
        const int q = wave * 64 + i * 32 + r;
        pbase[i] = (q / TX) * PW + (q % TX);
    }
    constexpr int NPV = (PH * PW * PV + 255) / 256, NWV = (9 * TN * PV + 255) / 256;
#pragma unroll
    for (int k = 0; k < NWV; k++) {   // the weights, once (Cin == CK)
        const int v = tid + 256 * k;
        if (v < 9 * TN * PV) {
            const int row = v / PV, part = v - row * PV;   // row = tap*TN + co
            const int tap = row / TN, co = row - tap * TN;
            *(u32x4 *)(wl + row * WS + part * 8) =
                *(const u32x4 *)(a.Wt + ((size_t)(co0 + co) * 9 + tap) * CK + part * 8);
        }
    }
    // CONV_PF patches in registers: tile t + CONV_PF is fetched while tile t runs
    u32x4 rp[CONV_PF][NPV];
    auto fetch = [&](int tile, u32x4 (&dst)[NPV]) {
        const int tx0 = (tile % tiles_x) * TX, ty0 = (tile / tiles_x) * TY;
#pragma unroll
        for (int k = 0; k < NPV; k++) {
            const int v = tid + 256 * k;
            if (v < PH * PW * PV) {
                const int pp = v / PV, part = v - pp * PV;
                const int py = pp / PW, px = pp - py * PW;
                dst[k] = load8(a.in, n, ty0 + py - 1, tx0 + px - 1, part * 8);
            }
        }
    };
    const int tile0 = blockIdx.x * NT;
#pragma unroll
    for (int u = 0; u < CONV_PF; u++)
        if (u < NT) fetch(tile0 + u, rp[u]);
#pragma unroll
    for (int t = 0; t < NT; t++) {
        const int tile = tile0 + t;
        const int tx0 = (tile % tiles_x) * TX, ty0 = (tile / tiles_x) * TY;
#pragma unroll
        for (int k = 0; k < NPV; k++) {
            const int v = tid + 256 * k;
            if (v < PH * PW * PV) {
                const int pp = v / PV, part = v - pp * PV;
                *(u32x4 *)(patch + pp * PS + part * 8) = rp[t % CONV_PF][k];
            }
        }
        __syncthreads();   // patch (and, the first time, weights) staged; last tile's output read
        if (t + CONV_PF < NT) fetch(tile + CONV_PF, rp[t % CONV_PF]);
        f32x16 acc[2][NJ];
#pragma unroll
        for (int i = 0; i < 2; i++)
#pragma unroll
            for (int j = 0; j < NJ; j++)
#pragma unroll
                for (int e = 0; e < 16; e++) acc[i][j][e] = 0.0f;
#pragma unroll CONV_TAP_UNROLL
        for (int tap = 0; tap < 9; tap++) {
            const int dy = tap / 3, dx = tap - 3 * (tap / 3);
#pragma unroll
            for (int kk = 0; kk < CK; kk += 16) {
                bf16x8 af[2], bfm[NJ];
#pragma unroll
                for (int i = 0; i < 2; i++)
                    af[i] = *(const bf16x8 *)(patch + (pbase[i] + dy * PW + dx) * PS + kk + kh8);
#pragma unroll
                for (int j = 0; j < NJ; j++)
                    bfm[j] = *(const bf16x8 *)(wl + (tap * TN + j * 32 + r) * WS + kk + kh8);
#pragma unroll
                for (int i = 0; i < 2; i++)
#pragma unroll
                    for (int j = 0; j < NJ; j++)
                        acc[i][j] = __builtin_amdgcn_mfma_f32_32x32x16_bf16(af[i], bfm[j],
                                                                            acc[i][j], 0, 0, 0);
            }
        }
        // epilogue as conv3x3_kernel's, through the output region
        conv_tile_stage<NJ>(acc, a.bias ? a.bias + co0 : nullptr, a.relu != 0, ot, wave, lane);
        __syncthreads();   // output tile complete; every read of this patch done
        conv_tile_out<TX, TN>(ot, a, n, tx0, ty0, co0, tid);
    }
}

#ifndef CONV_TILES
#define CONV_TILES 8
#endif
constexpr int CONV_NT = CONV_TILES;   // pixel tiles per work-group of conv3x3_tiles_kernel

template <int TX, int TN, int CK>
int launch_conv_tiles(const ConvArgs &a, int N, hipStream_t st)
{
    using L = ConvTiles<TX, TN, CK>;
    const void *fn = (const void *)conv3x3_tiles_kernel<TX, TN, CK, CONV_NT>;
    if (int rc = hipck(hipFuncSetAttribute(fn, hipFuncAttributeMaxDynamicSharedMemorySize,
                                           (int)L::lds_bytes)))
        return rc;
    const int ntiles = (a.in.W / TX) * (a.in.H / (256 / TX));
    dim3 grid(ntiles / CONV_NT, a.Cout / TN, N);
    hipLaunchKernelGGL((conv3x3_tiles_kernel<TX, TN, CK, CONV_NT>), grid, dim3(256), L::lds_bytes,
                       st, a);
    return hipck(hipGetLastError());
}

// 0: by shape (the tile loop where Cin fits one chunk, the output block is 32
// channels and the launch still has >= 1024 work-groups), 1: one tile per
// work-group, CONV_NT: the tile loop wherever Cin fits one chunk, the block is
// 32 channels and the tiles divide evenly.  Read once per launch; setting it
// while another thread launches convolutions picks either value.
std::atomic<int> g_conv_tiles{0};

// ---------------------------------------------------------------------------
// weight gradient: dW[co][tap][ci] += sum_p G[p][co] * X[p + tap][ci]
// M = 32*MJ output channels, N = 32 input channels, K = pixels.  Wave w
// (of 3) takes the taps of row dy = w.  Each work-group runs tiles_per_wg
// pixel tiles and then either adds its partial sums with FP32 atomics or
// stores them to its row of a workspace, which wgrad_reduce_kernel sums in a
// fixed order (bitwise reproducible).
// ---------------------------------------------------------------------------
struct WgradArgs {
    Src in;
    const bf16 *G;   // [N][H][W][Cout]
    float *dW;       // [Cout][9][Cin]
    float *dB;       // [Cout] bias gradient (added by the ci-block 0 work-groups) or null
    // null: atomics; else row blockIdx.x of part (row width Cout*9*Cin + Cout:
    // the weight sums, then the bias sums) takes this work-group's sums
    float *part;
    int N, Cout, tiles_per_wg;
};

// The transposed tiles' channel rows are padded by 8 (gradient) or 16
// (patch) elements per group of 8 rows, and every 8-element fragment read
// stays 16-byte aligned.  Staging the tiles in units of 4 pixels x 8
// channels transposed in registers (ds_write_b64 per channel row) instead of
// per 16-byte piece (eight ds_write_b16) also frees registers: the 64-channel
// kernel fits 2 waves per SIMD (158 VGPRs instead of 198) and ran 15-27 %
// faster on MI355X at batch 8 (Cout 64-256); with one 32-channel block the
// patch keeps the per-piece staging (evenly spread over the 192 threads,
// where 4-pixel units leave a third round to a quarter of them: 12 % slower),
// the gradient tile takes the units (4-10 % faster than before).
template <int TX, int MJ>
struct WgShape {
    static constexpr int TY = 256 / TX;
    static constexpr int PW = TX + 8, PH = TY + 2;   // transposed patch row (px contiguous)
    static constexpr int GS = 256 + 8;               // transposed gradient row
    static constexpr int CO = 32 * MJ;
    static constexpr int GBYTES = (CO * GS + CO) * 2;
    static constexpr size_t lds_bytes = (size_t)GBYTES + ((size_t)32 * PH * PW + 64) * 2;
    static __device__ int gofs(int c) { return c * GS + (c >> 3) * 8; }
    static __device__ int xofs(int c) { return c * PH * PW + (c >> 3) * 16; }
};

#ifndef WG_XUNITS_MJ
#define WG_XUNITS_MJ 2
#endif
template <int TX, int MJ>
__global__ __launch_bounds__(192) void wgrad_kernel(WgradArgs a)
{
    using S = WgShape<TX, MJ>;
    constexpr int TY = S::TY, PW = S::PW, PH = S::PH;
    constexpr int CO = 32 * MJ;
    extern __shared__ __attribute__((aligned(16))) bf16 lds[];
    bf16 *gT = lds;                       // [CO][GS]: gradient tile, pixels contiguous
    bf16 *xT = lds + S::GBYTES / 2;       // [32][PH][PW]: input patch, pixels contiguous

    const int tid = threadIdx.x, lane = tid & 63, wave = tid >> 6;
    const int H = a.in.H, W = a.in.W;
    const int Cin = a.in.CA + a.in.CB;
    const int co0 = blockIdx.y * CO, ci0 = blockIdx.z * 32;
    const int r = lane & 31, kh8 = (lane >> 5) * 8;
    const int tiles_x = W / TX, tiles_img = tiles_x * (H / TY);
    const int tiles_all = tiles_img * a.N;

    f32x16 acc[MJ][3];
#pragma unroll
    for (int m = 0; m < MJ; m++)
#pragma unroll
        for (int d = 0; d < 3; d++)
#pragma unroll
            for (int e = 0; e < 16; e++) acc[m][d][e] = 0.0f;

    const bool do_bias = a.dB && blockIdx.z == 0 && tid < CO;
    float bsum = 0.0f;
    const int t_begin = blockIdx.x * a.tiles_per_wg;
    const int t_end = min(t_begin + a.tiles_per_wg, tiles_all);
    // The next tile's gradient and input patch are loaded into registers
    // while the current tile's MFMAs run.  Staging unit: 4 consecutive pixels
    // x 8 channels, loaded as four 16-byte pixel pieces and transposed in
    // registers (v_perm) into eight 8-byte channel rows, one ds_write_b64 each
    // (instead of eight 2-byte writes per piece).
    constexpr int NBX = PW / 4;                          // 4-pixel blocks of a patch row
    constexpr int NUG = 64 * (CO / 8), NUX = PH * NBX * 4;
    constexpr int UG = (NUG + 191) / 192, UX = (NUX + 191) / 192;
    // with one 32-channel output block the patch is staged per 16-byte piece
    // (TX + 2 pixels a row, spread evenly over the 192 threads)
    constexpr bool XU = MJ >= WG_XUNITS_MJ;
    constexpr int NX = PH * (TX + 2) * 4, KX = (NX + 191) / 192;
    u32x4 rg[UG][4], rx[XU ? UX : 1][4], rxp[XU ? 1 : KX];
    auto fetch = [&](int t) {
        const int n = t / tiles_img, tt = t - n * tiles_img;
        const int tx0 = (tt % tiles_x) * TX, ty0 = (tt / tiles_x) * TY;
#pragma unroll
        for (int k = 0; k < UG; k++) {
            const int u = tid + 192 * k;
            if (u < NUG) {
                const int part = u % (CO / 8), q0 = 4 * (u / (CO / 8));
                const int y = ty0 + q0 / TX, x = tx0 + q0 % TX;
                const bf16 *g = a.G + (((size_t)n * H + y) * W + x) * a.Cout + co0 + part * 8;
#pragma unroll
                for (int p = 0; p < 4; p++) rg[k][p] = *(const u32x4 *)(g + (size_t)p * a.Cout);
            }
        }
        if constexpr (XU) {
#pragma unroll
            for (int k = 0; k < UX; k++) {
                const int u = tid + 192 * k;
                if (u < NUX) {
                    const int part = u & 3, pb = u >> 2;
                    const int py = pb / NBX, px0 = (pb - py * NBX) * 4;
#pragma unroll
                    for (int p = 0; p < 4; p++)
                        rx[k][p] = load8(a.in, n, ty0 + py - 1, tx0 + px0 + p - 1, ci0 + part * 8);
                }
            }
        } else {
#pragma unroll
            for (int k = 0; k < KX; k++) {
                const int v = tid + 192 * k;
                if (v < NX) {
                    const int pp = v >> 2, part = v & 3;
                    const int py = pp / (TX + 2), px = pp - py * (TX + 2);
                    rxp[k] = load8(a.in, n, ty0 + py - 1, tx0 + px - 1, ci0 + part * 8);
                }
            }
        }
    };
    // channel c of four pixel pieces v[0..3] (8 channels each) as 4 bf16
    auto column = [](const u32x4 (&v)[4], int c) {
        const unsigned sel = (c & 1) ? 0x07060302u : 0x05040100u;
        const int d = c >> 1;
        return u32x2{__builtin_amdgcn_perm(v[1][d], v[0][d], sel),
                     __builtin_amdgcn_perm(v[3][d], v[2][d], sel)};
    };
    if (t_begin < t_end) fetch(t_begin);
    for (int t = t_begin; t < t_end; t++) {
        // gradient tile -> gT[co][q], input patch -> xT[ci][py][px]
#pragma unroll
        for (int k = 0; k < UG; k++) {
            const int u = tid + 192 * k;
            if (u < NUG) {
                const int part = u % (CO / 8), q0 = 4 * (u / (CO / 8));
#pragma unroll
                for (int c = 0; c < 8; c++)
                    *(u32x2 *)(gT + S::gofs(part * 8 + c) + q0) = column(rg[k], c);
            }
        }
        if constexpr (XU) {
#pragma unroll
            for (int k = 0; k < UX; k++) {
                const int u = tid + 192 * k;
                if (u < NUX) {
                    const int part = u & 3, pb = u >> 2;
                    const int py = pb / NBX, px0 = (pb - py * NBX) * 4;
#pragma unroll
                    for (int c = 0; c < 8; c++)
                        *(u32x2 *)(xT + S::xofs(part * 8 + c) + py * PW + px0) = column(rx[k], c);
                }
            }
        } else {
#pragma unroll
            for (int k = 0; k < KX; k++) {
                const int v = tid + 192 * k;
                if (v < NX) {
                    const int pp = v >> 2, part = v & 3;
                    const int py = pp / (TX + 2), px = pp - py * (TX + 2);
                    const bf16 *xs = (const bf16 *)&rxp[k];
#pragma unroll
                    for (int e = 0; e < 8; e++) xT[S::xofs(part * 8 + e) + py * PW + px] = xs[e];
                }
            }
        }
        __syncthreads();
        if (t + 1 < t_end) fetch(t + 1);
        if (do_bias)
            for (int q = 0; q < 256; q++) bsum += (float)gT[S::gofs(tid) + q];
        const int dy = wave;
#pragma unroll 2
        for (int ks = 0; ks < 16; ks++) {
            const int q0 = ks * 16 + kh8;             // 8 consecutive pixels of one tile row
            const int ty = q0 / TX, tx = q0 % TX;
            bf16x8 af[MJ];
#pragma unroll
            for (int m = 0; m < MJ; m++)
                af[m] = *(const bf16x8 *)(gT + S::gofs(m * 32 + r) + q0);
            // pixels tx+dx .. tx+dx+7 of channel r for dx = 0, 1, 2: two aligned
            // 16-byte reads (elements tx .. tx+15) and 2-byte funnel shifts
            const bf16 *xrow = xT + S::xofs(r) + (ty + dy) * PW + tx;
            const u32x4 lo = *(const u32x4 *)xrow, hi = *(const u32x4 *)(xrow + 8);
            u32x4 bw[3];
            bw[0] = lo;
            bw[1] = u32x4{__builtin_amdgcn_alignbyte(lo.y, lo.x, 2),
                          __builtin_amdgcn_alignbyte(lo.z, lo.y, 2),
                          __builtin_amdgcn_alignbyte(lo.w, lo.z, 2),
                          __builtin_amdgcn_alignbyte(hi.x, lo.w, 2)};
            bw[2] = u32x4{lo.y, lo.z, lo.w, hi.x};
#pragma unroll
            for (int dx = 0; dx < 3; dx++) {
                const bf16x8 bv = __builtin_bit_cast(bf16x8, bw[dx]);
#pragma unroll
                for (int m = 0; m < MJ; m++)
                    acc[m][dx] =
                        __builtin_amdgcn_mfma_f32_32x32x16_bf16(af[m], bv, acc[m][dx], 0, 0, 0);
            }
        }
        __syncthreads();
    }
    const size_t nw = (size_t)a.Cout * 9 * Cin;
    float *row = a.part ? a.part + (size_t)blockIdx.x * (nw + a.Cout) : nullptr;
    if (do_bias) {
        if (row) row[nw + co0 + tid] = bsum;
        else atomicAdd(a.dB + co0 + tid, bsum);
    }
    const int ci = ci0 + r;
    if (ci >= Cin) return;
#pragma unroll
    for (int m = 0; m < MJ; m++)
#pragma unroll
        for (int dx = 0; dx < 3; dx++)
#pragma unroll
            for (int e = 0; e < 16; e++) {
                const int co = co0 + m * 32 + 8 * (e >> 2) + 4 * (lane >> 5) + (e & 3);
                const size_t o = ((size_t)co * 9 + 3 * wave + dx) * Cin + ci;
                if (row) row[o] = acc[m][dx][e];
                else atomicAdd(a.dW + o, acc[m][dx][e]);
            }
}

// One pass of the fixed-order reduction of the weight-gradient workspace:
// out[j][i] = sum over g in [16 j, min(16 j + 16, rows)) of in[g][i], summed
// in order; the last pass (rows <= 16) adds the sum to dW (i < nw) and dB
// (nw <= i < nw + Cout, when dB is given) instead.
constexpr int WG_RED = 16;
__global__ void wgrad_reduce_kernel(const float *in, int rows, size_t width, size_t nw, float *out,
                                    float *dW, float *dB)
{
    const size_t i = (size_t)blockIdx.x * blockDim.x + threadIdx.x;
    const int j = blockIdx.y;
    if (i >= width) return;
    const int g0 = j * WG_RED, g1 = min(g0 + WG_RED, rows);
    float s = 0.0f;
    for (int g = g0; g < g1; g++) s += in[(size_t)g * width + i];
    if (rows > WG_RED) {
        out[(size_t)j * width + i] = s;
    } else if (i < nw) {
        dW[i] += s;
    } else if (dB) {
        dB[i - nw] += s;
    }
}

// Pixel groups (work-group rows along x) of a weight-gradient launch.
inline int wgrad_groups(int N, int H, int W, int tiles_per_wg)
{
    const long long tiles = (long long)N * H * W / 256;
    return (int)((tiles + tiles_per_wg - 1) / tiles_per_wg);
}

// Workspace floats of the fixed-order weight gradient: the rows of every
// pixel group, then the rows of the first reduction pass.
inline size_t wgrad_work_elems(int N, int H, int W, int Cin, int Cout, int tiles_per_wg)
{
    const size_t g = (size_t)wgrad_groups(N, H, W, tiles_per_wg);
    const size_t width = (size_t)Cout * 9 * Cin + Cout;
    return (g + (g + WG_RED - 1) / WG_RED) * width;
}

template <int TX, int MJ>
int launch_wgrad(const WgradArgs &a, hipStream_t st)
{
    using S = WgShape<TX, MJ>;
    const void *fn = (const void *)wgrad_kernel<TX, MJ>;
    if (int rc = hipck(hipFuncSetAttribute(fn, hipFuncAttributeMaxDynamicSharedMemorySize,
                                           (int)S::lds_bytes)))
        return rc;
    const int tiles = (a.in.W / TX) * (a.in.H / S::TY) * a.N;
    const int Cin = a.in.CA + a.in.CB;
    dim3 grid((tiles + a.tiles_per_wg - 1) / a.tiles_per_wg, a.Cout / (32 * MJ), (Cin + 31) / 32);
    hipLaunchKernelGGL((wgrad_kernel<TX, MJ>), grid, dim3(192), S::lds_bytes, st, a);
    if (int rc = hipck(hipGetLastError())) return rc;
    if (!a.part) return 0;
    // fixed-order reduction over the pixel groups, 16 rows per sum, ping-pong
    // between the group rows and the first pass's rows
    const size_t nw = (size_t)a.Cout * 9 * Cin, width = nw + a.Cout;
    float *bufs[2] = {a.part, a.part + (size_t)grid.x * width};
    int rows = (int)grid.x, k = 0;
    for (;;) {
        const int out_rows = rows > WG_RED ? (rows + WG_RED - 1) / WG_RED : 1;
        hipLaunchKernelGGL(wgrad_reduce_kernel, dim3(blocks(width), out_rows), dim3(256), 0, st,
                           bufs[k], rows, width, nw, bufs[k ^ 1], a.dW, a.dB);
        if (int rc = hipck(hipGetLastError())) return rc;
        if (rows <= WG_RED) return 0;
        rows = out_rows;
        k ^= 1;
    }
}

// ---------------------------------------------------------------------------
// small kernels
// ---------------------------------------------------------------------------
__global__ void flip_weights_kernel(const bf16 *W, int Cout, int Cin, bf16 *Wf)
{
    // Wf[ci][tap][co] = W[co][8 - tap][ci]
    const size_t i = (size_t)blockIdx.x * blockDim.x + threadIdx.x;
    const size_t n = (size_t)Cout * 9 * Cin;
    if (i >= n) return;
    const int co = (int)(i % Cout);
    const int tap = (int)((i / Cout) % 9);
    const int ci = (int)(i / ((size_t)Cout * 9));
    Wf[i] = W[((size_t)co * 9 + (8 - tap)) * Cin + ci];
}

__global__ void pool2_kernel(const bf16 *X, int N, int H, int W, int C, bf16 *Y)
{
    // Y[n][y][x][c] = mean of X over the 2x2 block; 8 channels per thread
    const int Ho = H / 2, Wo = W / 2, C8 = C / 8;
    const size_t i = (size_t)blockIdx.x * blockDim.x + threadIdx.x;
    if (i >= (size_t)N * Ho * Wo * C8) return;
    const int c = (int)(i % C8) * 8;
    size_t p = i / C8;
    const int x = (int)(p % Wo); p /= Wo;
    const int y = (int)(p % Ho);
    const int n = (int)(p / Ho);
    float s[8] = {};
    for (int dy = 0; dy < 2; dy++)
        for (int dx = 0; dx < 2; dx++) {
            const u32x4 v =
                *(const u32x4 *)(X + (((size_t)n * H + 2 * y + dy) * W + 2 * x + dx) * C + c);
            const bf16 *b = (const bf16 *)&v;
            for (int k = 0; k < 8; k++) s[k] += (float)b[k];
        }
    u32x4 o;
    bf16 *ob = (bf16 *)&o;
    for (int k = 0; k < 8; k++) ob[k] = (bf16)(0.25f * s[k]);
    *(u32x4 *)(Y + (((size_t)n * Ho + y) * Wo + x) * C + c) = o;
}

// out[n][y][x][c] (coarse, C channels) = mask > 0 ? sum over the 2x2 fine
// pixels of src[.][.][.][off + c] : 0   (backward of nearest upsample + ReLU);
// 8 channels per thread
__global__ void upsum_mask_kernel(const bf16 *src, int Cs, int off, const bf16 *mask, int N,
                                  int Ho, int Wo, int C, bf16 *out)
{
    const int C8 = C / 8;
    const size_t i = (size_t)blockIdx.x * blockDim.x + threadIdx.x;
    if (i >= (size_t)N * Ho * Wo * C8) return;
    const int c = (int)(i % C8) * 8;
    const size_t p = i / C8;
    const int x = (int)(p % Wo);
    const int y = (int)((p / Wo) % Ho);
    const int n = (int)(p / ((size_t)Wo * Ho));
    const int H = 2 * Ho, W = 2 * Wo;
    const u32x4 mv = *(const u32x4 *)(mask + p * C + c);
    const bf16 *mb = (const bf16 *)&mv;
    float v[8] = {};
    for (int dy = 0; dy < 2; dy++)
        for (int dx = 0; dx < 2; dx++) {
            const u32x4 sv = *(const u32x4 *)(src + (((size_t)n * H + 2 * y + dy) * W + 2 * x + dx) *
                                                       Cs + off + c);
            const bf16 *sb = (const bf16 *)&sv;
#pragma unroll
            for (int k = 0; k < 8; k++) v[k] += (float)sb[k];
        }
    u32x4 o;
    bf16 *ob = (bf16 *)&o;
#pragma unroll
    for (int k = 0; k < 8; k++) ob[k] = (bf16)((float)mb[k] > 0.0f ? v[k] : 0.0f);
    *(u32x4 *)(out + p * C + c) = o;
}

// out[n][y][x][c] = mask > 0 ? src[.][off + c] + below[n][y/2][x/2][c] / 4 : 0
// (gradient of an encoder output: skip path + average-pool path, then ReLU);
// 8 channels per thread
__global__ void skip_combine_kernel(const bf16 *src, int Cs, int off, const bf16 *below,
                                    const bf16 *mask, int N, int H, int W, int C, bf16 *out)
{
    const int C8 = C / 8;
    const size_t i = (size_t)blockIdx.x * blockDim.x + threadIdx.x;
    if (i >= (size_t)N * H * W * C8) return;
    const int c = (int)(i % C8) * 8;
    const size_t p = i / C8;
    const int x = (int)(p % W);
    const int y = (int)((p / W) % H);
    const int n = (int)(p / ((size_t)W * H));
    const u32x4 mv = *(const u32x4 *)(mask + p * C + c);
    const u32x4 sv = *(const u32x4 *)(src + p * Cs + off + c);
    const u32x4 bv =
        *(const u32x4 *)(below + (((size_t)n * (H / 2) + y / 2) * (W / 2) + x / 2) * C + c);
    const bf16 *mb = (const bf16 *)&mv, *sb = (const bf16 *)&sv, *bb = (const bf16 *)&bv;
    u32x4 o;
    bf16 *ob = (bf16 *)&o;
#pragma unroll
    for (int k = 0; k < 8; k++) {
        const float v = (float)sb[k] + 0.25f * (float)bb[k];
        ob[k] = (bf16)((float)mb[k] > 0.0f ? v : 0.0f);
    }
    *(u32x4 *)(out + p * C + c) = o;
}

// x0[n][y][x][0..15] = (in[n][0][y][x], in[n][1][y][x], 0, ...) in BF16
__global__ void pack_input_kernel(const float *in, int N, int H, int W, bf16 *x0)
{
    const size_t i = (size_t)blockIdx.x * blockDim.x + threadIdx.x;
    const size_t hw = (size_t)H * W;
    if (i >= (size_t)N * hw) return;
    const size_t n = i / hw, p = i - n * hw;
    u32x4 v0 = {0u, 0u, 0u, 0u}, v1 = {0u, 0u, 0u, 0u};
    bf16 *b = (bf16 *)&v0;
    b[0] = (bf16)in[(n * 2) * hw + p];
    b[1] = (bf16)in[(n * 2 + 1) * hw + p];
    *(u32x4 *)(x0 + i * 16) = v0;
    *(u32x4 *)(x0 + i * 16 + 8) = v1;
}

// out[p] = b + sum_c X[p][c] * w[c]  (final 1x1 convolution, C = 32)
__global__ void head_fwd_kernel(const bf16 *X, const float *w, const float *b, size_t np,
                                float *out)
{
    const size_t p = (size_t)blockIdx.x * blockDim.x + threadIdx.x;
    if (p >= np) return;
    float s = b[0];
    for (int c8 = 0; c8 < 32; c8 += 8) {
        const u32x4 v = *(const u32x4 *)(X + p * 32 + c8);
        const bf16 *xb = (const bf16 *)&v;
        for (int k = 0; k < 8; k++) s = fmaf((float)xb[k], w[c8 + k], s);
    }
    out[p] = s;
}

// MSE head backward: g = scale*(out - target); loss += sum (out-target)^2;
// gX[p][c] = X[p][c] > 0 ? g*w[c] : 0 (BF16); dw[c] += sum_p g*X[p][c];
// db += sum_p g.  256 threads per block; a thread takes elements e = (pixel
// e / 4, channels 8 (e % 4) .. + 7), so the lanes of a wave read and write 16
// consecutive pixels' 64 bytes each (coalesced), grid-stride with four
// elements' loads issued together.
// bsums (may be null): block b stores its 34 sums (dw[0..31], db, loss) to
// bsums[34 b ..] instead of adding them atomically; head_bwd_sum_kernel adds
// them in block order.
__global__ __launch_bounds__(256) void head_bwd_kernel(const bf16 *X, const float *w,
                                                       const float *out, const float *target,
                                                       size_t np, float scale, bf16 *gX,
                                                       float *dw, float *db, double *loss,
                                                       float *bsums)
{
    __shared__ float red[4][34];
    const int k = threadIdx.x & 3;               // channel group of every element of this thread
    float wk[8];
#pragma unroll
    for (int j = 0; j < 8; j++) wk[j] = w[8 * k + j];
    float part[8] = {}, pg = 0.0f, pl = 0.0f;
    const size_t ne = np * 4, stride = (size_t)gridDim.x * blockDim.x;
    for (size_t e0 = (size_t)blockIdx.x * blockDim.x + threadIdx.x; e0 < ne; e0 += 4 * stride) {
        u32x4 xv[4];
        float ov[4], tv[4];
#pragma unroll
        for (int u = 0; u < 4; u++) {
            const size_t e = e0 + u * stride;
            if (e < ne) {
                xv[u] = *(const u32x4 *)(X + e * 8);
                ov[u] = out[e >> 2];
                tv[u] = target[e >> 2];
            }
        }
#pragma unroll
        for (int u = 0; u < 4; u++) {
            const size_t e = e0 + u * stride;
            if (e >= ne) break;
            const float d = ov[u] - tv[u];
            const float g = scale * d;
            if (k == 0) {
                pg += g;
                pl += d * d;
            }
            const bf16 *xb = (const bf16 *)&xv[u];
            u32x4 o;
            bf16 *ob = (bf16 *)&o;
#pragma unroll
            for (int j = 0; j < 8; j++) {
                const float x = (float)xb[j];
                part[j] += g * x;
                ob[j] = (bf16)(x > 0.0f ? g * wk[j] : 0.0f);
            }
            *(u32x4 *)(gX + e * 8) = o;
        }
    }
    // lanes of one channel group (lane % 4) meet in a fixed butterfly
    const int lane = threadIdx.x & 63, wave = threadIdx.x >> 6;
#pragma unroll
    for (int o = 4; o < 64; o <<= 1) {
#pragma unroll
        for (int j = 0; j < 8; j++) part[j] += __shfl_xor(part[j], o, 64);
        pg += __shfl_xor(pg, o, 64);
        pl += __shfl_xor(pl, o, 64);
    }
    if (lane < 4) {
#pragma unroll
        for (int j = 0; j < 8; j++) red[wave][8 * lane + j] = part[j];
        if (lane == 0) {
            red[wave][32] = pg;
            red[wave][33] = pl;
        }
    }
    __syncthreads();
    if (threadIdx.x < 34) {
        const int c = threadIdx.x;
        const float s = red[0][c] + red[1][c] + red[2][c] + red[3][c];
        if (bsums) bsums[34 * blockIdx.x + c] = s;
        else if (c < 32) atomicAdd(dw + c, s);
        else if (c == 32) atomicAdd(db, s);
        else atomicAdd(loss, (double)s);
    }
}

// The head's block sums added in a fixed order: work-group k (one wave)
// takes sum k; lane l adds blocks l, l + 64, ... in order, then the lanes'
// sums meet in a fixed butterfly.  dw, db in FP32, the loss in FP64.
__global__ __launch_bounds__(64) void head_bwd_sum_kernel(const float *part, int nb, float *dw,
                                                          float *db, double *loss)
{
    const int k = blockIdx.x, l = threadIdx.x;
    double s = 0.0;
    float f = 0.0f;
    for (int b = l; b < nb; b += 64) {
        if (k == 33) s += (double)part[34 * b + k];
        else f += part[34 * b + k];
    }
    for (int o = 32; o > 0; o >>= 1) {
        s += __shfl_xor(s, o, 64);
        f += __shfl_xor(f, o, 64);
    }
    if (l != 0) return;
    if (k == 33) *loss += s;
    else if (k == 32) *db += f;
    else dw[k] += f;
}

// dB[c] += sum over pixels of G[p][c]; each block sums 1024 pixels
__global__ __launch_bounds__(256) void bias_grad_kernel(const bf16 *G, size_t np, int C, float *dB)
{
    __shared__ float red[256];
    const int per = 256 / C;                  // pixel lanes per channel (C divides 256)
    const int c = threadIdx.x % C, lane = threadIdx.x / C;
    const size_t p0 = (size_t)blockIdx.x * 1024;
    float s = 0.0f;
    if (lane < per)
        for (size_t p = p0 + lane; p < p0 + 1024 && p < np; p += per) s += (float)G[p * C + c];
    red[threadIdx.x] = s;
    __syncthreads();
    if (threadIdx.x < C) {
        float t = 0.0f;
        for (int l = 0; l < per; l++) t += red[l * C + threadIdx.x];
        atomicAdd(dB + threadIdx.x, t);
    }
}

__global__ void adam_kernel(float *p, const float *g, float *m, float *v, bf16 *pb, size_t n,
                            float lr, float b1, float b2, float eps, float bc1, float bc2)
{
    const size_t i = (size_t)blockIdx.x * blockDim.x + threadIdx.x;
    if (i >= n) return;
    const float gi = g[i];
    const float mi = b1 * m[i] + (1.0f - b1) * gi;
    const float vi = b2 * v[i] + (1.0f - b2) * gi * gi;
    m[i] = mi;
    v[i] = vi;
    const float pi = p[i] - lr * (mi / bc1) / (sqrtf(vi / bc2) + eps);
    p[i] = pi;
    if (pb) pb[i] = (bf16)pi;
}

__global__ void to_bf16_kernel(const float *x, size_t n, bf16 *y)
{
    const size_t i = (size_t)blockIdx.x * blockDim.x + threadIdx.x;
    if (i < n) y[i] = (bf16)x[i];
}


bool al16(const void *p) { return ((uintptr_t)p & 15) == 0; }

int check_src(const Src &s, int N)
{
    if (N <= 0 || s.H <= 0 || s.W <= 0) return ufail("bad tensor shape");
    if (s.CA <= 0 || s.CA % 8 || s.CB < 0 || s.CB % 8)
        return ufail("channel counts must be positive multiples of 8");
    if (!s.A || !al16(s.A) || (s.CB && (!s.B || !al16(s.B))))
        return ufail("input tensors must be non-null and 16-byte aligned");
    if (s.CB && s.b_half && (s.H % 2 || s.W % 2)) return ufail("upsampled input needs even H, W");
    return 0;
}

}  // namespace

extern "C" {

const char *otfwi_unet_last_error(void) { return g_uerr.c_str(); }

#define BF(p) reinterpret_cast<const bf16 *>(p)
#define BFW(p) reinterpret_cast<bf16 *>(p)
#define ST(s) reinterpret_cast<hipStream_t>(s)

int otfwi_unet_conv3x3(const otfwi_bf16 *A, int CA, const otfwi_bf16 *B, int CB, int b_half,
                       const otfwi_bf16 *Wt, const float *bias, const otfwi_bf16 *mask,
                       otfwi_bf16 *Y, int N, int H, int W, int Cout, int relu, void *stream)
{
    ConvArgs a{};
    a.in = Src{BF(A), CA, BF(B), CB, b_half, H, W};
    a.Wt = BF(Wt); a.bias = bias; a.mask = BF(mask); a.Y = BFW(Y); a.Cout = Cout; a.relu = relu;
    if (int rc = check_src(a.in, N)) return rc;
    if (!Wt || !Y || !al16(Wt) || Cout <= 0 || Cout % 32)
        return ufail("bad weights/output (Cout must be a multiple of 32)");
    const int Cin = CA + CB;
    const int CK = Cin % 32 == 0 ? 32 : (Cin == 16 ? 16 : 0);
    if (!CK || (CB && CA % 32)) return ufail("Cin must be 16 or a multiple of 32 (CA too)");
    const int TX = W % 64 == 0 ? 64 : (W % 32 == 0 ? 32 : 0);
    if (!TX || H % (256 / TX)) return ufail("W must be a multiple of 32 and H of the tile height");
    const int TN = Cout % 64 == 0 ? 64 : 32;
    hipStream_t st = ST(stream);
    const long long ntiles = (long long)(W / TX) * (H / (256 / TX));
    // the tile loop for 32-channel output blocks (75 KB of LDS, two work-groups
    // per CU as conv3x3_kernel; 64-channel blocks would need 112 KB)
    const int nt = g_conv_tiles.load(std::memory_order_relaxed);
    const bool tiles = Cin == CK && TN == 32 && ntiles % CONV_NT == 0 && nt != 1 &&
                       (nt == CONV_NT || ntiles / CONV_NT * (Cout / TN) * (long long)N >= 1024);
    if (tiles) {
        if (CK == 16)
            return TX == 64 ? launch_conv_tiles<64, 32, 16>(a, N, st)
                            : launch_conv_tiles<32, 32, 16>(a, N, st);
        return TX == 64 ? launch_conv_tiles<64, 32, 32>(a, N, st)
                        : launch_conv_tiles<32, 32, 32>(a, N, st);
    }
    if (CK == 16) {
        if (TX == 64)
            return TN == 64 ? launch_conv<64, 64, 16>(a, N, st) : launch_conv<64, 32, 16>(a, N, st);
        return TN == 64 ? launch_conv<32, 64, 16>(a, N, st) : launch_conv<32, 32, 16>(a, N, st);
    }
    if (TX == 64)
        return TN == 64 ? launch_conv<64, 64, 32>(a, N, st) : launch_conv<64, 32, 32>(a, N, st);
    return TN == 64 ? launch_conv<32, 64, 32>(a, N, st) : launch_conv<32, 32, 32>(a, N, st);
}

int otfwi_unet_set_conv_tiles(int nt)
{
    if (nt != 0 && nt != 1 && nt != CONV_NT) return ufail("conv tiles per work-group: 0, 1 or 8");
    g_conv_tiles.store(nt, std::memory_order_relaxed);
    return 0;
}

int otfwi_unet_wgrad(const otfwi_bf16 *A, int CA, const otfwi_bf16 *B, int CB, int b_half,
                     const otfwi_bf16 *G, float *dW, float *dB, int N, int H, int W, int Cout,
                     int tiles_per_wg, void *stream)
{
    return otfwi_unet_wgrad_work(A, CA, B, CB, b_half, G, dW, dB, N, H, W, Cout, tiles_per_wg,
                                 nullptr, 0, stream);
}

long long otfwi_unet_wgrad_work_elems(int N, int H, int W, int Cin, int Cout, int tiles_per_wg)
{
    if (N <= 0 || H <= 0 || W <= 0 || Cin <= 0 || Cout <= 0 || tiles_per_wg <= 0) return -1;
    return (long long)wgrad_work_elems(N, H, W, Cin, Cout, tiles_per_wg);
}

int otfwi_unet_wgrad_work(const otfwi_bf16 *A, int CA, const otfwi_bf16 *B, int CB, int b_half,
                          const otfwi_bf16 *G, float *dW, float *dB, int N, int H, int W, int Cout,
                          int tiles_per_wg, float *work, size_t work_elems, void *stream)
{
    WgradArgs a{};
    a.in = Src{BF(A), CA, BF(B), CB, b_half, H, W};
    a.G = BF(G); a.dW = dW; a.dB = dB; a.N = N; a.Cout = Cout; a.tiles_per_wg = tiles_per_wg;
    a.part = work;
    if (int rc = check_src(a.in, N)) return rc;
    if (!G || !dW || !al16(G) || Cout <= 0 || Cout % 32)
        return ufail("bad gradient/output (Cout must be a multiple of 32)");
    if (tiles_per_wg <= 0) return ufail("tiles_per_wg must be positive");
    if (work && work_elems < wgrad_work_elems(N, H, W, CA + CB, Cout, tiles_per_wg))
        return ufail("weight-gradient workspace too small (otfwi_unet_wgrad_work_elems)");
    if (CB && CA % 32) return ufail("CA must be a multiple of 32 with a second input");
    const int TX = W % 64 == 0 ? 64 : (W % 32 == 0 ? 32 : 0);
    if (!TX || H % (256 / TX)) return ufail("W must be a multiple of 32 and H of the tile height");
    hipStream_t st = ST(stream);
    if (Cout % 64 == 0)
        return TX == 64 ? launch_wgrad<64, 2>(a, st) : launch_wgrad<32, 2>(a, st);
    return TX == 64 ? launch_wgrad<64, 1>(a, st) : launch_wgrad<32, 1>(a, st);
}

int otfwi_unet_flip_weights(const otfwi_bf16 *W, int Cout, int Cin, otfwi_bf16 *Wf, void *stream)
{
    if (!W || !Wf || Cout <= 0 || Cin <= 0) return ufail("bad weights");
    const size_t n = (size_t)Cout * 9 * Cin;
    hipLaunchKernelGGL(flip_weights_kernel, dim3(blocks(n)), dim3(256), 0, ST(stream), BF(W), Cout,
                       Cin, BFW(Wf));
    return hipck(hipGetLastError());
}

int otfwi_unet_pool2(const otfwi_bf16 *X, int N, int H, int W, int C, otfwi_bf16 *Y, void *stream)
{
    if (!X || !Y || C % 8 || H % 2 || W % 2 || !al16(X) || !al16(Y)) return ufail("bad pool args");
    const size_t n = (size_t)N * (H / 2) * (W / 2) * (C / 8);
    hipLaunchKernelGGL(pool2_kernel, dim3(blocks(n)), dim3(256), 0, ST(stream), BF(X), N, H, W, C,
                       BFW(Y));
    return hipck(hipGetLastError());
}

int otfwi_unet_upsum_mask(const otfwi_bf16 *src, int Cs, int off, const otfwi_bf16 *mask, int N,
                          int Ho, int Wo, int C, otfwi_bf16 *out, void *stream)
{
    if (!src || !mask || !out || off < 0 || off + C > Cs) return ufail("bad upsum args");
    if (C % 8 || Cs % 8 || off % 8 || !al16(src) || !al16(mask) || !al16(out))
        return ufail("upsum: channel counts must be multiples of 8, tensors 16-byte aligned");
    const size_t n = (size_t)N * Ho * Wo * (C / 8);
    hipLaunchKernelGGL(upsum_mask_kernel, dim3(blocks(n)), dim3(256), 0, ST(stream), BF(src), Cs,
                       off, BF(mask), N, Ho, Wo, C, BFW(out));
    return hipck(hipGetLastError());
}

int otfwi_unet_skip_combine(const otfwi_bf16 *src, int Cs, int off, const otfwi_bf16 *below,
                            const otfwi_bf16 *mask, int N, int H, int W, int C, otfwi_bf16 *out,
                            void *stream)
{
    if (!src || !below || !mask || !out || off < 0 || off + C > Cs || H % 2 || W % 2)
        return ufail("bad skip-combine args");
    if (C % 8 || Cs % 8 || off % 8 || !al16(src) || !al16(below) || !al16(mask) || !al16(out))
        return ufail("skip-combine: channel counts must be multiples of 8, tensors 16-byte aligned");
    const size_t n = (size_t)N * H * W * (C / 8);
    hipLaunchKernelGGL(skip_combine_kernel, dim3(blocks(n)), dim3(256), 0, ST(stream), BF(src), Cs,
                       off, BF(below), BF(mask), N, H, W, C, BFW(out));
    return hipck(hipGetLastError());
}

int otfwi_unet_pack_input(const float *in, int N, int H, int W, otfwi_bf16 *x0, void *stream)
{
    if (!in || !x0 || !al16(x0)) return ufail("bad pack args");
    const size_t n = (size_t)N * H * W;
    hipLaunchKernelGGL(pack_input_kernel, dim3(blocks(n)), dim3(256), 0, ST(stream), in, N, H, W,
                       BFW(x0));
    return hipck(hipGetLastError());
}

int otfwi_unet_head_fwd(const otfwi_bf16 *X, const float *w, const float *b, long long np,
                        float *out, void *stream)
{
    if (!X || !w || !b || !out || np <= 0 || !al16(X)) return ufail("bad head args");
    hipLaunchKernelGGL(head_fwd_kernel, dim3(blocks((size_t)np)), dim3(256), 0, ST(stream), BF(X),
                       w, b, (size_t)np, out);
    return hipck(hipGetLastError());
}

int otfwi_unet_head_bwd(const otfwi_bf16 *X, const float *w, const float *out,
                        const float *target, long long np, float scale, otfwi_bf16 *gX, float *dw,
                        float *db, double *loss, void *stream)
{
    return otfwi_unet_head_bwd_work(X, w, out, target, np, scale, gX, dw, db, loss, nullptr, 0,
                                    stream);
}

int otfwi_unet_head_bwd_work(const otfwi_bf16 *X, const float *w, const float *out,
                             const float *target, long long np, float scale, otfwi_bf16 *gX,
                             float *dw, float *db, double *loss, float *work, size_t work_elems,
                             void *stream)
{
    if (!X || !w || !out || !target || !gX || !dw || !db || !loss || np <= 0 || !al16(X) ||
        !al16(gX))
        return ufail("bad head-backward args");
    if (work && work_elems < (size_t)OTFWI_UNET_HEAD_WORK)
        return ufail("head-backward workspace too small (OTFWI_UNET_HEAD_WORK floats)");
    const unsigned nb = std::min(blocks((size_t)np * 4), 1024u);
    hipLaunchKernelGGL(head_bwd_kernel, dim3(nb), dim3(256), 0, ST(stream), BF(X),
                       w, out, target, (size_t)np, scale, BFW(gX), dw, db, loss, work);
    if (int rc = hipck(hipGetLastError())) return rc;
    if (!work) return 0;
    hipLaunchKernelGGL(head_bwd_sum_kernel, dim3(34), dim3(64), 0, ST(stream), work, (int)nb, dw, db,
                       loss);
    return hipck(hipGetLastError());
}

int otfwi_unet_bias_grad(const otfwi_bf16 *G, long long np, int C, float *dB, void *stream)
{
    if (!G || !dB || np <= 0 || C <= 0 || C > 256 || 256 % C) return ufail("bad bias-grad args");
    hipLaunchKernelGGL(bias_grad_kernel, dim3(blocks((size_t)np, 1024)), dim3(256), 0, ST(stream),
                       BF(G), (size_t)np, C, dB);
    return hipck(hipGetLastError());
}

int otfwi_unet_adam(float *p, const float *g, float *m, float *v, otfwi_bf16 *pb, long long n,
                    float lr, float b1, float b2, float eps, int step, void *stream)
{
    if (!p || !g || !m || !v || n <= 0 || step <= 0) return ufail("bad adam args");
    const float bc1 = 1.0f - powf(b1, (float)step), bc2 = 1.0f - powf(b2, (float)step);
    hipLaunchKernelGGL(adam_kernel, dim3(blocks((size_t)n)), dim3(256), 0, ST(stream), p, g, m, v,
                       BFW(pb), (size_t)n, lr, b1, b2, eps, bc1, bc2);
    return hipck(hipGetLastError());
}

int otfwi_unet_to_bf16(const float *x, long long n, otfwi_bf16 *y, void *stream)
{
    if (!x || !y || n <= 0) return ufail("bad conversion args");
    hipLaunchKernelGGL(to_bf16_kernel, dim3(blocks((size_t)n)), dim3(256), 0, ST(stream), x,
                       (size_t)n, BFW(y));
    return hipck(hipGetLastError());
}

}  // extern "C"
