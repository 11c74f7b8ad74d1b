// Learned warm start for WaveHoltz: a four-level convolutional
// encoder-decoder (U-Net) on element images, trained and evaluated with the
// kernels in this file.  The reference has no learned component; the
// network only proposes the initial iterate W_1, and the WaveHoltz
// iteration decides the final residual.
//
// Images.  On an M x M grid of quads numbered e = ex + M ey (the uniform
// rectangle numbering) element e is pixel (ey, ex); its channels are the
// element's nodal values, so the DG block of element e (3 n^2 doubles,
// (3, n, n) column-major) is the pixel's channel vector.
//
// Numerics.  Activations and weights are bf16 (round to nearest even),
// products run on v_mfma_f32_16x16x32_bf16 (v_mfma_f32_16x16x16bf16_1k in
// the kernels for 16-channel layers) with fp32 accumulation, the
// master weights, Adam moments, gradients, bias terms, loss and the
// output layer are fp32.  Every reduction runs in a fixed order (per-chunk
// partials, then a sequential sum over chunks): results are bitwise
// reproducible.
//
// Network (C = base width, channel counts x1, x2, x4, x8):
//   e1 = conv(conv(x))            M     C       p1 = avgpool(e1)
//   e2 = conv(conv(p1))           M/2   2C      p2 = avgpool(e2)
//   e3 = conv(conv(p2))           M/4   4C      p3 = avgpool(e3)
//   b  = conv(conv(p3))           M/8   8C
//   d3 = conv(conv([up(b),  e3])) M/4   4C
//   d2 = conv(conv([up(d3), e2])) M/2   2C
//   d1 = conv(conv([up(d2), e1])) M     C
//   y  = conv1x1(d1)                    out channels, no activation
// conv = 3x3, zero padding, bias, ReLU; up = nearest; [a, b] = channel concat.
#include "../../include/wavedg_hip.h"

#include <hip/hip_runtime.h>

#include <algorithm>
#include <cmath>
#include <cstdint>
#include <cstring>
#include <memory>
#include <string>
#include <vector>

#include "host_basis.hpp"
#include "host_mesh.hpp"

namespace {

typedef short bf16x4 __attribute__((ext_vector_type(4)));
typedef short i16x8 __attribute__((ext_vector_type(8)));
typedef __bf16 bf16x8 __attribute__((ext_vector_type(8)));
typedef float f32x4 __attribute__((ext_vector_type(4)));

constexpr int kWaves = 4;     // waves per workgroup of the conv kernels
constexpr int NLAYER_MAX = 15;  // layers of the network (the Layer enum below)
constexpr int kChunk = 1024;  // pixels per weight-gradient partial

__device__ __forceinline__ uint16_t f2bf(float f)
{
    uint32_t u = __float_as_uint(f);
    if ((u & 0x7f800000u) == 0x7f800000u) return (uint16_t)((u >> 16) | ((u & 0xffffu) ? 0x40u : 0u));
    u += 0x7fffu + ((u >> 16) & 1u);
    return (uint16_t)(u >> 16);
}

__device__ __forceinline__ float bf2f(uint16_t h) { return __uint_as_float((uint32_t)h << 16); }

// ------------------------------------------------------------------ conv
// out[p, co] = act(bias[co] + sum_{tap, ci} in[p + off(tap), ci] w[tap][co][ci])
// as a GEMM with M = output channels, N = pixels, K = taps x input channels.
// A wave owns TA x TB tiles of 16 channels x 16 pixels.  MFMA operands:
// A lane l holds channel co0 + l%16 and k = 4(l/16)..4(l/16)+3, i.e. four
// consecutive input channels of one weight row (one 8-byte load); B lane l
// holds pixel l%16 and the same four input channels (one 8-byte load from
// the NHWC image).  The accumulator lane l holds pixel l%16 and output
// channels 4(l/16)..+3, stored as one 8-byte (bf16) or 16-byte (fp32) write.
struct ConvArgs {
    const uint16_t* in;
    const uint16_t* w;   // [ks*ks][cout][cin] bf16
    const float* bias;   // [cout] or null
    void* out;
    int in_ld, in_off, cin;
    int out_ld, out_off, cout;
    int B, H, W, ks, relu;
    // conv3_k32_kernel with bf16 output: when set, outputs whose value at
    // mask[p, mask_off + c] (leading dimension mask_ld) is not positive are
    // stored as 0 (the ReLU backward of the layer below, fused)
    const uint16_t* mask;
    int mask_ld, mask_off;
};

template <int TA, int TB, bool F32OUT>
__global__ __launch_bounds__(64 * kWaves) void conv_kernel(ConvArgs a)
{
    const int lane = threadIdx.x & 63, wave = threadIdx.x >> 6;
    const int ln = lane & 15, kq = lane >> 4;
    const long long npix = (long long)a.B * a.H * a.W;
    const long long p0 = ((long long)blockIdx.x * kWaves + wave) * (16 * TB);
    const int co0 = blockIdx.y * (16 * TA);
    if (p0 >= npix) return;  // no barriers below

    int pb[TB], py[TB], pxx[TB];
    bool pok[TB];
#pragma unroll
    for (int j = 0; j < TB; ++j) {
        const long long p = p0 + 16 * j + ln;
        pok[j] = p < npix;
        const long long pc = pok[j] ? p : 0;
        pxx[j] = (int)(pc % a.W);
        const long long t = pc / a.W;
        py[j] = (int)(t % a.H);
        pb[j] = (int)(t / a.H);
    }
    f32x4 acc[TA][TB];
#pragma unroll
    for (int i = 0; i < TA; ++i)
#pragma unroll
        for (int j = 0; j < TB; ++j) acc[i][j] = f32x4{0.f, 0.f, 0.f, 0.f};

    const int r = a.ks >> 1;
    for (int tap = 0; tap < a.ks * a.ks; ++tap) {
        const int dy = tap / a.ks - r, dx = tap % a.ks - r;
        const uint16_t* src[TB];
        bool ok[TB];
#pragma unroll
        for (int j = 0; j < TB; ++j) {
            const int ny = py[j] + dy, nx = pxx[j] + dx;
            ok[j] = pok[j] && ny >= 0 && ny < a.H && nx >= 0 && nx < a.W;
            const long long q = ok[j] ? ((long long)pb[j] * a.H + ny) * a.W + nx : 0;
            src[j] = a.in + q * a.in_ld + a.in_off + 4 * kq;
        }
        const uint16_t* wsrc = a.w + ((long long)tap * a.cout + co0 + ln) * a.cin + 4 * kq;
        for (int c0 = 0; c0 < a.cin; c0 += 16) {
            bf16x4 av[TA], bv[TB];
#pragma unroll
            for (int j = 0; j < TB; ++j) {
                uint2 v = *reinterpret_cast<const uint2*>(src[j] + c0);
                if (!ok[j]) v = make_uint2(0u, 0u);
                bv[j] = __builtin_bit_cast(bf16x4, v);
            }
#pragma unroll
            for (int i = 0; i < TA; ++i) {
                const uint2 v = *reinterpret_cast<const uint2*>(wsrc + (long long)16 * i * a.cin + c0);
                av[i] = __builtin_bit_cast(bf16x4, v);
            }
#pragma unroll
            for (int i = 0; i < TA; ++i)
#pragma unroll
                for (int j = 0; j < TB; ++j)
                    acc[i][j] = __builtin_amdgcn_mfma_f32_16x16x16bf16_1k(av[i], bv[j], acc[i][j], 0, 0, 0);
        }
    }
#pragma unroll
    for (int i = 0; i < TA; ++i) {
        const int co = co0 + 16 * i + 4 * kq;
        float bias[4] = {0.f, 0.f, 0.f, 0.f};
        if (a.bias) {
            const float4 b4 = *reinterpret_cast<const float4*>(a.bias + co);
            bias[0] = b4.x, bias[1] = b4.y, bias[2] = b4.z, bias[3] = b4.w;
        }
#pragma unroll
        for (int j = 0; j < TB; ++j) {
            if (!pok[j]) continue;
            const long long p = p0 + 16 * j + ln;
            float v[4];
#pragma unroll
            for (int q = 0; q < 4; ++q) {
                v[q] = acc[i][j][q] + bias[q];
                if (a.relu) v[q] = fmaxf(v[q], 0.f);
            }
            const long long o = p * a.out_ld + a.out_off + co;
            if constexpr (F32OUT) {
                *reinterpret_cast<float4*>(static_cast<float*>(a.out) + o) = make_float4(v[0], v[1], v[2], v[3]);
            } else {
                const uint32_t lo = (uint32_t)f2bf(v[0]) | ((uint32_t)f2bf(v[1]) << 16);
                const uint32_t hi = (uint32_t)f2bf(v[2]) | ((uint32_t)f2bf(v[3]) << 16);
                *reinterpret_cast<uint2*>(static_cast<uint16_t*>(a.out) + o) = make_uint2(lo, hi);
            }
        }
    }
}

// LDS-tiled 3x3 convolution (same operands and accumulator layout as
// conv_kernel).  A workgroup owns CO_T = 32 TA output channels of one tile of
// 64 pixels: TH rows x TW columns of one image, TW = min(W, 64), W a power of
// two.  For each slice of 16 input channels it stages the tile with its
// one-pixel halo (zero outside the image, at most 3 x 66 pixels) and the
// slice's 9 x CO_T x 16 weights in LDS with 16-byte loads, so every input
// value is fetched once per slice instead of once per tap; then each wave
// runs 9 taps x TA x 2 MFMAs on 16 TA channels x 32 pixels.
constexpr int kHaloMax = 198;  // (TH + 2)(TW + 2) for TW in {1, ..., 64}, TH = 64 / TW

template <int TA, bool F32OUT>
__global__ __launch_bounds__(256) void conv3_tiled_kernel(ConvArgs a)
{
    constexpr int CO_T = 32 * TA;
    __shared__ __align__(16) uint16_t s_in[kHaloMax * 16];
    __shared__ __align__(16) uint16_t s_w[9 * CO_T * 16];
    const int tid = threadIdx.x, lane = tid & 63, wave = tid >> 6;
    const int ln = lane & 15, kq = lane >> 4;
    const int TW = min(a.W, 64), TH = 64 / TW;
    const int tiles_x = a.W / TW, tiles_y = (a.H + TH - 1) / TH;
    const int t = blockIdx.x;
    const int b = t / (tiles_x * tiles_y), ty = (t / tiles_x) % tiles_y, tx = t % tiles_x;
    const int y0 = ty * TH, x0 = tx * TW;
    const int co0 = blockIdx.y * CO_T;
    const int cw = (wave & 1) * (16 * TA);  // the wave's channels within the tile
    const int pw = (wave >> 1) * 32;        // the wave's first pixel within the tile
    const int HW2 = TW + 2, NH = (TH + 2) * HW2;
    int py[2], px[2];
#pragma unroll
    for (int j = 0; j < 2; ++j) {
        const int p = pw + 16 * j + ln;
        py[j] = p / TW;
        px[j] = p - py[j] * TW;
    }
    f32x4 acc[TA][2];
#pragma unroll
    for (int i = 0; i < TA; ++i)
#pragma unroll
        for (int j = 0; j < 2; ++j) acc[i][j] = f32x4{0.f, 0.f, 0.f, 0.f};

    for (int c0 = 0; c0 < a.cin; c0 += 16) {
        __syncthreads();  // the previous slice is consumed
        for (int q = tid; q < NH * 2; q += 256) {
            const int hp = q >> 1, half = q & 1;
            const int hy = hp / HW2, hx = hp - hy * HW2;
            const int y = y0 - 1 + hy, x = x0 - 1 + hx;
            uint4 v = make_uint4(0u, 0u, 0u, 0u);
            if (y >= 0 && y < a.H && x >= 0 && x < a.W)
                v = *reinterpret_cast<const uint4*>(a.in + ((size_t)(b * a.H + y) * a.W + x) * a.in_ld + a.in_off +
                                                    c0 + 8 * half);
            *reinterpret_cast<uint4*>(s_in + hp * 16 + 8 * half) = v;
        }
        for (int q = tid; q < 9 * CO_T * 2; q += 256) {
            const int row = q >> 1, half = q & 1;
            const int tap = row / CO_T, co = row - tap * CO_T;
            *reinterpret_cast<uint4*>(s_w + row * 16 + 8 * half) = *reinterpret_cast<const uint4*>(
                a.w + ((size_t)tap * a.cout + co0 + co) * a.cin + c0 + 8 * half);
        }
        __syncthreads();
#pragma unroll
        for (int tap = 0; tap < 9; ++tap) {
            const int hy = tap / 3, hx = tap % 3;  // halo offsets dy + 1, dx + 1
            bf16x4 av[TA], bv[2];
#pragma unroll
            for (int i = 0; i < TA; ++i)
                av[i] = __builtin_bit_cast(
                    bf16x4, *reinterpret_cast<const uint2*>(s_w + (tap * CO_T + cw + 16 * i + ln) * 16 + 4 * kq));
#pragma unroll
            for (int j = 0; j < 2; ++j)
                bv[j] = __builtin_bit_cast(
                    bf16x4, *reinterpret_cast<const uint2*>(s_in + ((py[j] + hy) * HW2 + px[j] + hx) * 16 + 4 * kq));
#pragma unroll
            for (int i = 0; i < TA; ++i)
#pragma unroll
                for (int j = 0; j < 2; ++j)
                    acc[i][j] = __builtin_amdgcn_mfma_f32_16x16x16bf16_1k(av[i], bv[j], acc[i][j], 0, 0, 0);
        }
    }
#pragma unroll
    for (int i = 0; i < TA; ++i) {
        const int co = co0 + cw + 16 * i + 4 * kq;
        float bias[4] = {0.f, 0.f, 0.f, 0.f};
        if (a.bias) {
            const float4 b4 = *reinterpret_cast<const float4*>(a.bias + co);
            bias[0] = b4.x, bias[1] = b4.y, bias[2] = b4.z, bias[3] = b4.w;
        }
#pragma unroll
        for (int j = 0; j < 2; ++j) {
            const int y = y0 + py[j], x = x0 + px[j];
            if (y >= a.H) continue;
            float v[4];
#pragma unroll
            for (int q = 0; q < 4; ++q) {
                v[q] = acc[i][j][q] + bias[q];
                if (a.relu) v[q] = fmaxf(v[q], 0.f);
            }
            const size_t o = ((size_t)(b * a.H + y) * a.W + x) * a.out_ld + a.out_off + co;
            if constexpr (F32OUT) {
                *reinterpret_cast<float4*>(static_cast<float*>(a.out) + o) = make_float4(v[0], v[1], v[2], v[3]);
            } else {
                const uint32_t lo = (uint32_t)f2bf(v[0]) | ((uint32_t)f2bf(v[1]) << 16);
                const uint32_t hi = (uint32_t)f2bf(v[2]) | ((uint32_t)f2bf(v[3]) << 16);
                *reinterpret_cast<uint2*>(static_cast<uint16_t*>(a.out) + o) = make_uint2(lo, hi);
            }
        }
    }
}

// The same convolution on gfx950's v_mfma_f32_16x16x32_bf16: slices of 32
// input channels (a last slice of 16 is padded with zeros), so each operand
// lane holds 8 consecutive channels (one 16-byte LDS read) and a slice costs
// half the MFMA instructions, LDS reads and staging rounds of the 16-channel
// kernel.  A lane l holds channel co + l % 16 and input channels
// 8 (l / 16) .. + 7 of the weight row; B lane l pixel l % 16 and the same 8
// input channels; the accumulator layout is that of conv3_tiled_kernel.
// RT > 1 (images 64 or more pixels wide) makes the tile RT rows of 64 pixels:
// fewer halo rows per pixel and the slice's weights staged once for RT times
// the pixels.
template <int TA, int RT, bool F32OUT>
__global__ __launch_bounds__(256) void conv3_k32_kernel(ConvArgs a)
{
    constexpr int CO_T = 32 * TA;
    constexpr int HALO = RT == 1 ? kHaloMax : (RT + 2) * 66;  // halo pixels at most
    constexpr int NHP = (HALO * 4 + 255) / 256;                // 16-byte halo parts per thread
    constexpr int NWP = (9 * CO_T * 4 + 255) / 256;           // 16-byte weight parts per thread
    __shared__ __align__(16) uint16_t s_in[HALO * 32];
    __shared__ __align__(16) uint16_t s_w[9 * CO_T * 32];
    const int tid = threadIdx.x, lane = tid & 63, wave = tid >> 6;
    const int ln = lane & 15, kq = lane >> 4;
    const int TW = min(a.W, 64), TH = 64 / TW * RT;
    const int tiles_x = a.W / TW, tiles_y = (a.H + TH - 1) / TH;
    const int t = blockIdx.x;
    const int b = t / (tiles_x * tiles_y), ty = (t / tiles_x) % tiles_y, tx = t % tiles_x;
    const int y0 = ty * TH, x0 = tx * TW;
    const int co0 = blockIdx.y * CO_T;
    const int cw = (wave & 1) * (16 * TA);
    const int pw = (wave >> 1) * 32 * RT;
    const int HW2 = TW + 2, NH = (TH + 2) * HW2;
    int py[2 * RT], px[2 * RT];
#pragma unroll
    for (int j = 0; j < 2 * RT; ++j) {
        const int p = pw + 16 * j + ln;
        py[j] = p / TW;
        px[j] = p - py[j] * TW;
    }
    // The 16-byte parts this thread stages: part q = tid + 256 k goes to LDS
    // element 8 q (halo pixel q / 4 or weight row q / 4, channels 8 (q % 4) ..);
    // its global offset without the slice's first channel, -1 for a zero
    // (outside the image), -2 for no part.
    const int part = tid & 3;
    long long hoff[NHP], woff[NWP];
#pragma unroll
    for (int k = 0; k < NHP; ++k) {
        const int q = tid + 256 * k, hp = q >> 2;
        const int hy = hp / HW2, hx = hp - hy * HW2;
        const int y = y0 - 1 + hy, x = x0 - 1 + hx;
        hoff[k] = q >= NH * 4 ? -2
                  : (y >= 0 && y < a.H && x >= 0 && x < a.W)
                      ? ((long long)(b * a.H + y) * a.W + x) * a.in_ld + a.in_off + 8 * part
                      : -1;
    }
#pragma unroll
    for (int k = 0; k < NWP; ++k) {
        const int q = tid + 256 * k, row = q >> 2;
        const int tap = row / CO_T, co = row - tap * CO_T;
        woff[k] = q >= 9 * CO_T * 4 ? -2 : ((long long)tap * a.cout + co0 + co) * a.cin + 8 * part;
    }
    uint4 hreg[NHP], wreg[NWP];
    // global loads of the slice starting at channel c0 into registers
    auto load = [&](int c0) {
        const bool in_range = c0 + 8 * part < a.cin;  // a last slice of 16 channels is padded with zeros
#pragma unroll
        for (int k = 0; k < NHP; ++k) {
            hreg[k] = make_uint4(0u, 0u, 0u, 0u);
            if (hoff[k] >= 0 && in_range) hreg[k] = *reinterpret_cast<const uint4*>(a.in + hoff[k] + c0);
        }
#pragma unroll
        for (int k = 0; k < NWP; ++k) {
            wreg[k] = make_uint4(0u, 0u, 0u, 0u);
            if (woff[k] >= 0 && in_range) wreg[k] = *reinterpret_cast<const uint4*>(a.w + woff[k] + c0);
        }
    };
    f32x4 acc[TA][2 * RT];
#pragma unroll
    for (int i = 0; i < TA; ++i)
#pragma unroll
        for (int j = 0; j < 2 * RT; ++j) acc[i][j] = f32x4{0.f, 0.f, 0.f, 0.f};

    load(0);
    for (int c0 = 0; c0 < a.cin; c0 += 32) {
        __syncthreads();  // the previous slice is consumed
#pragma unroll
        for (int k = 0; k < NHP; ++k)
            if (hoff[k] != -2) *reinterpret_cast<uint4*>(s_in + 8 * (tid + 256 * k)) = hreg[k];
#pragma unroll
        for (int k = 0; k < NWP; ++k)
            if (woff[k] != -2) *reinterpret_cast<uint4*>(s_w + 8 * (tid + 256 * k)) = wreg[k];
        __syncthreads();
        if (c0 + 32 < a.cin) load(c0 + 32);  // in flight during this slice's MFMAs
#pragma unroll
        for (int tap = 0; tap < 9; ++tap) {
            const int hy = tap / 3, hx = tap % 3;
            bf16x8 av[TA], bv[2 * RT];
#pragma unroll
            for (int i = 0; i < TA; ++i)
                av[i] = __builtin_bit_cast(
                    bf16x8, *reinterpret_cast<const uint4*>(s_w + (tap * CO_T + cw + 16 * i + ln) * 32 + 8 * kq));
#pragma unroll
            for (int j = 0; j < 2 * RT; ++j)
                bv[j] = __builtin_bit_cast(
                    bf16x8, *reinterpret_cast<const uint4*>(s_in + ((py[j] + hy) * HW2 + px[j] + hx) * 32 + 8 * kq));
#pragma unroll
            for (int i = 0; i < TA; ++i)
#pragma unroll
                for (int j = 0; j < 2 * RT; ++j)
                    acc[i][j] = __builtin_amdgcn_mfma_f32_16x16x32_bf16(av[i], bv[j], acc[i][j], 0, 0, 0);
        }
    }
#pragma unroll
    for (int i = 0; i < TA; ++i) {
        const int co = co0 + cw + 16 * i + 4 * kq;
        float bias[4] = {0.f, 0.f, 0.f, 0.f};
        if (a.bias) {
            const float4 b4 = *reinterpret_cast<const float4*>(a.bias + co);
            bias[0] = b4.x, bias[1] = b4.y, bias[2] = b4.z, bias[3] = b4.w;
        }
#pragma unroll
        for (int j = 0; j < 2 * RT; ++j) {
            const int y = y0 + py[j], x = x0 + px[j];
            if (y >= a.H) continue;
            float v[4];
#pragma unroll
            for (int q = 0; q < 4; ++q) {
                v[q] = acc[i][j][q] + bias[q];
                if (a.relu) v[q] = fmaxf(v[q], 0.f);
            }
            const size_t pix = (size_t)(b * a.H + y) * a.W + x;
            const size_t o = pix * a.out_ld + a.out_off + co;
            if constexpr (F32OUT) {
                *reinterpret_cast<float4*>(static_cast<float*>(a.out) + o) = make_float4(v[0], v[1], v[2], v[3]);
            } else {
                if (a.mask) {
                    const uint2 mv = *reinterpret_cast<const uint2*>(a.mask + pix * a.mask_ld + a.mask_off + co);
                    const uint32_t mw[2] = {mv.x, mv.y};
#pragma unroll
                    for (int q = 0; q < 4; ++q)
                        if (!(bf2f((uint16_t)(mw[q >> 1] >> (16 * (q & 1)))) > 0.f)) v[q] = 0.f;
                }
                const uint32_t lo = (uint32_t)f2bf(v[0]) | ((uint32_t)f2bf(v[1]) << 16);
                const uint32_t hi = (uint32_t)f2bf(v[2]) | ((uint32_t)f2bf(v[3]) << 16);
                *reinterpret_cast<uint2*>(static_cast<uint16_t*>(a.out) + o) = make_uint2(lo, hi);
            }
        }
    }
}

// dw[tap][co][ci] = sum_p dout[p, co] in[p + off(tap), ci]: a GEMM with
// M = output channels, N = input channels, K = pixels.  One wave owns one
// 16 x 16 tile of one tap over one chunk of pixels and writes its partial
// sum to partial[chunk][tap][co][ci].
struct WgradArgs {
    const uint16_t* in;
    const uint16_t* dout;
    float* partial;
    int in_ld, in_off, cin;
    int dout_ld, cout;
    int B, H, W, ks;
    int tiles_per_chunk;  // wgrad3_tr_kernel
};

__global__ __launch_bounds__(64 * kWaves) void wgrad_kernel(WgradArgs a)
{
    const int lane = threadIdx.x & 63, wave = threadIdx.x >> 6;
    const int ln = lane & 15, kq = lane >> 4;
    const int nco = a.cout >> 4, nci = a.cin >> 4, ntap = a.ks * a.ks;
    const int tile = blockIdx.y * kWaves + wave;
    if (tile >= ntap * nco * nci) return;
    const int tap = tile / (nco * nci), rem = tile % (nco * nci);
    const int cob = rem / nci, cib = rem % nci;
    const int dy = tap / a.ks - (a.ks >> 1), dx = tap % a.ks - (a.ks >> 1);
    const int npix = a.B * a.H * a.W;
    const int pbeg = blockIdx.x * kChunk;
    const int pend = min(npix, pbeg + kChunk);
    const int co = cob * 16 + ln, ci = cib * 16 + ln;
    // this lane's pixels are q = p + 4 kq + r, r = 0..3; their (b, y, x)
    // advance by 16 pixels per step
    int qb[4], qy[4], qx[4];
#pragma unroll
    for (int r = 0; r < 4; ++r) {
        const int q = pbeg + 4 * kq + r;
        qx[r] = q % a.W;
        const int t = q / a.W;
        qy[r] = t % a.H;
        qb[r] = t / a.H;
    }
    f32x4 acc = {0.f, 0.f, 0.f, 0.f};
    for (int p = pbeg; p < pend; p += 16) {
        bf16x4 av, bv;
#pragma unroll
        for (int r = 0; r < 4; ++r) {
            const int q = p + 4 * kq + r;
            const bool okq = q < pend;
            const int qc = okq ? q : pbeg;
            const uint16_t dv = a.dout[(size_t)qc * a.dout_ld + co];
            av[r] = (short)(okq ? dv : 0);
            const int ny = qy[r] + dy, nx = qx[r] + dx;
            const bool ok = okq && ny >= 0 && ny < a.H && nx >= 0 && nx < a.W;
            const int src = ok ? (qb[r] * a.H + ny) * a.W + nx : 0;
            const uint16_t iv = a.in[(size_t)src * a.in_ld + a.in_off + ci];
            bv[r] = (short)(ok ? iv : 0);
            // advance to pixel q + 16
            qx[r] += 16;
            while (qx[r] >= a.W) {
                qx[r] -= a.W;
                if (++qy[r] == a.H) {
                    qy[r] = 0;
                    ++qb[r];
                }
            }
        }
        acc = __builtin_amdgcn_mfma_f32_16x16x16bf16_1k(av, bv, acc, 0, 0, 0);
    }
    float* dst = a.partial + ((size_t)blockIdx.x * ntap + tap) * a.cout * a.cin;
#pragma unroll
    for (int r = 0; r < 4; ++r) dst[(size_t)(cob * 16 + 4 * kq + r) * a.cin + cib * 16 + ln] = acc[r];
}

// Weight and bias gradients of a 3x3 layer (W a power of two, 32k output and
// 16k input channels; a last block of 16 input channels is padded with
// zeros) on gfx950's transposing LDS load.  A workgroup of six
// waves owns 32 output x 32 input channels for all 9 taps and a chunk of
// a.tiles_per_chunk 64-pixel tiles (the tiles of conv3_tiled_kernel).  Per
// tile it stages the output gradient as [pixel][32 co] and the input with its
// halo as [halo pixel][32 ci], both in the NHWC order of global memory with
// 16-byte stores.  The weight gradient is a GEMM over pixels, so an MFMA
// operand lane needs one channel at 8 consecutive pixels: ds_read_b64_tr_b16
// hands lane 4a + e of each 16-lane group element e of the 8-byte reads of
// lanes a, a + 4, a + 8 and a + 12, so when lane a + 4j reads channels
// 4a .. 4a + 3 of pixel p + j, lane i receives channel i at pixels p .. p + 3;
// two such reads fill one v_mfma_f32_16x16x32_bf16 operand.  Wave w owns the
// input-channel block w % 2 and the taps w / 2, w / 2 + 3, w / 2 + 6 for both
// output-channel blocks.  Wave 0 of the first input-channel block also sums
// the output gradient over the chunk's pixels (the bias gradient).  A chunk's
// partial sums form one row of 9 cout cin + cout floats, [tap][co][ci] then
// [co], the layout of the layer's weights and bias in the parameter vector,
// so one reduction over the rows gives both gradients.
constexpr int kWgWaves = 6;
#ifndef WDG_WGRAD_TW
#define WDG_WGRAD_TW 16
#endif
#ifndef WDG_WGRAD_WG
#define WDG_WGRAD_WG 512
#endif
// Widest tile of wgrad3_tr_kernel (a power of two <= 64): 64 pixels as 4
// rows of 16 stage 6 x 18 halo pixels instead of 3 x 66 for one row of 64
// (1.2 % faster training step at 128^2 than 64-wide tiles).
constexpr int kWgTW = WDG_WGRAD_TW;
constexpr int kWgTarget = WDG_WGRAD_WG;  // workgroups a layer's chunks aim for
#ifndef WDG_WGRAD_TPC
#define WDG_WGRAD_TPC 16
#endif
constexpr int kWgTpcMax = WDG_WGRAD_TPC;  // tiles per chunk at most

__device__ __forceinline__ bf16x8 lds_tr_pair(const uint16_t* p0, const uint16_t* p1)
{
    typedef short i16x4 __attribute__((ext_vector_type(4)));
    typedef __attribute__((address_space(3))) i16x4 lds_i16x4;
    const i16x4 lo = __builtin_amdgcn_ds_read_tr16_b64_v4i16((lds_i16x4*)p0);
    const i16x4 hi = __builtin_amdgcn_ds_read_tr16_b64_v4i16((lds_i16x4*)p1);
    return __builtin_bit_cast(bf16x8, i16x8{lo[0], lo[1], lo[2], lo[3], hi[0], hi[1], hi[2], hi[3]});
}

__global__ __launch_bounds__(64 * kWgWaves) void wgrad3_tr_kernel(WgradArgs a)
{
    __shared__ __align__(16) uint16_t s_d[64 * 32];
    __shared__ __align__(16) uint16_t s_x[kHaloMax * 32];
    const int tid = threadIdx.x, lane = tid & 63, wave = tid >> 6;
    const int g = lane >> 4, ln = lane & 15, la = ln & 3, lj = ln >> 2;
    const int TW = min(a.W, kWgTW), TH = 64 / TW;
    const int sh = __builtin_ctz((unsigned)TW);
    const int tiles_x = a.W / TW, tiles_y = (a.H + TH - 1) / TH;
    const int ntiles = a.B * tiles_y * tiles_x;
    const int HW2 = TW + 2, NH = (TH + 2) * HW2;
    const int co0 = blockIdx.y * 32, ci0 = blockIdx.z * 32;
    const int t0 = blockIdx.x * a.tiles_per_chunk, t1 = min(ntiles, t0 + a.tiles_per_chunk);
    const int ib = wave & 1, tapb = wave >> 1;
    const bool bias = wave == 0 && blockIdx.z == 0;
    f32x4 acc[3][2];
#pragma unroll
    for (int u = 0; u < 3; ++u) acc[u][0] = acc[u][1] = f32x4{0.f, 0.f, 0.f, 0.f};
    float bsum[2] = {0.f, 0.f};
    // the 16-byte parts this thread stages per tile: one of the output
    // gradient (threads < 256: pixel tid / 4) and up to NXP of the halo (part
    // q = tid + 384 k: halo pixel q / 4), all channels 8 (tid % 4) ..
    constexpr int NXP = (kHaloMax * 4 + 64 * kWgWaves - 1) / (64 * kWgWaves);
    const int part = tid & 3;
    const bool ci_in = ci0 + 8 * part < a.cin;
    uint4 dreg, xreg[NXP];
    auto load = [&](int t) {
        const int b = t / (tiles_x * tiles_y), ty = (t / tiles_x) % tiles_y, tx = t % tiles_x;
        const int y0 = ty * TH, x0 = tx * TW;
        dreg = make_uint4(0u, 0u, 0u, 0u);
        if (tid < 256) {
            const int pp = tid >> 2;
            const int y = y0 + (pp >> sh), x = x0 + (pp & (TW - 1));
            if (y < a.H)
                dreg = *reinterpret_cast<const uint4*>(a.dout + ((size_t)(b * a.H + y) * a.W + x) * a.dout_ld + co0 +
                                                       8 * part);
        }
#pragma unroll
        for (int k = 0; k < NXP; ++k) {
            const int q = tid + 64 * kWgWaves * k, hp = q >> 2;
            const int hy = hp / HW2, hx = hp - hy * HW2;
            const int y = y0 - 1 + hy, x = x0 - 1 + hx;
            xreg[k] = make_uint4(0u, 0u, 0u, 0u);
            if (q < NH * 4 && y >= 0 && y < a.H && x >= 0 && x < a.W && ci_in)
                xreg[k] = *reinterpret_cast<const uint4*>(a.in + ((size_t)(b * a.H + y) * a.W + x) * a.in_ld +
                                                          a.in_off + ci0 + 8 * part);
        }
    };
    if (t0 < t1) load(t0);
    for (int t = t0; t < t1; ++t) {
        __syncthreads();  // the previous tile is consumed
        if (tid < 256) *reinterpret_cast<uint4*>(s_d + 8 * tid) = dreg;
#pragma unroll
        for (int k = 0; k < NXP; ++k) {
            const int q = tid + 64 * kWgWaves * k;
            if (q < NH * 4) *reinterpret_cast<uint4*>(s_x + 8 * q) = xreg[k];
        }
        __syncthreads();
        if (t + 1 < t1) load(t + 1);  // in flight during this tile's MFMAs
        // output-gradient operands: channel co0 + 16 cb + ln at pixels 32 kh + 8 g + 0..7
        bf16x8 A[2][2];
#pragma unroll
        for (int cb = 0; cb < 2; ++cb)
#pragma unroll
            for (int kh = 0; kh < 2; ++kh) {
                const int p = 32 * kh + 8 * g + lj;
                A[cb][kh] = lds_tr_pair(s_d + p * 32 + 16 * cb + 4 * la, s_d + (p + 4) * 32 + 16 * cb + 4 * la);
            }
        if (bias) {
#pragma unroll
            for (int cb = 0; cb < 2; ++cb)
#pragma unroll
                for (int kh = 0; kh < 2; ++kh)
#pragma unroll
                    for (int q = 0; q < 8; ++q) bsum[cb] += (float)A[cb][kh][q];
        }
#pragma unroll
        for (int u = 0; u < 3; ++u) {
            const int tap = tapb + 3 * u, hy = tap / 3, hx = tap - 3 * hy;
#pragma unroll
            for (int kh = 0; kh < 2; ++kh) {
                const int p0 = 32 * kh + 8 * g + lj, p1 = p0 + 4;
                const int h0 = ((p0 >> sh) + hy) * HW2 + (p0 & (TW - 1)) + hx;
                const int h1 = ((p1 >> sh) + hy) * HW2 + (p1 & (TW - 1)) + hx;
                const bf16x8 bv = lds_tr_pair(s_x + h0 * 32 + 16 * ib + 4 * la, s_x + h1 * 32 + 16 * ib + 4 * la);
                acc[u][0] = __builtin_amdgcn_mfma_f32_16x16x32_bf16(A[0][kh], bv, acc[u][0], 0, 0, 0);
                acc[u][1] = __builtin_amdgcn_mfma_f32_16x16x32_bf16(A[1][kh], bv, acc[u][1], 0, 0, 0);
            }
        }
    }
    // accumulator lane l: co = co0 + 16 cb + 4 (l / 16) + r, ci = ci0 + 16 ib + l % 16
    const size_t row = (size_t)9 * a.cout * a.cin + a.cout;
    float* const prow = a.partial + (size_t)blockIdx.x * row;
    if (ci0 + 16 * ib < a.cin)  // a last block of 16 input channels is padded with zeros
#pragma unroll
    for (int u = 0; u < 3; ++u) {
        float* dst = prow + (size_t)(tapb + 3 * u) * a.cout * a.cin;
#pragma unroll
        for (int cb = 0; cb < 2; ++cb)
#pragma unroll
            for (int r = 0; r < 4; ++r)
                dst[(size_t)(co0 + 16 * cb + 4 * g + r) * a.cin + ci0 + 16 * ib + ln] = acc[u][cb][r];
    }
    if (bias) {
#pragma unroll
        for (int cb = 0; cb < 2; ++cb) {
            float v = bsum[cb];
            v += __shfl_xor(v, 16);
            v += __shfl_xor(v, 32);
            if (g == 0) prow[(size_t)9 * a.cout * a.cin + co0 + 16 * cb + ln] = v;
        }
    }
}

// Weight and bias gradients of a 1x1 layer (16k output channels, at most 256;
// 16k input channels, at most 64) on the transposing LDS load: a workgroup of
// cout / 16 waves takes a chunk of a.tiles_per_chunk tiles of 64 consecutive
// pixels, staged in LDS in their NHWC order ([64][cout] and [64][cin], 16-byte
// stores); wave w owns output-channel block w for every input block and sums
// its output-gradient fragments for the bias.  A chunk's partial row is
// [co][ci] then [co], as in wgrad3_tr_kernel.
constexpr int kW1MaxIb = 4;

__global__ __launch_bounds__(1024) void wgrad1_tr_kernel(WgradArgs a)
{
    extern __shared__ __align__(16) uint16_t smem[];
    uint16_t* const s_d = smem;               // [64][cout]
    uint16_t* const s_x = smem + 64 * a.cout;  // [64][cin]
    const int tid = threadIdx.x, lane = tid & 63, wave = tid >> 6;
    const int g = lane >> 4, ln = lane & 15, la = ln & 3, lj = ln >> 2;
    const int nib = a.cin / 16, nthr = blockDim.x;
    const long long npix = (long long)a.B * a.H * a.W;
    const long long ntiles = (npix + 63) / 64;
    const long long t0 = (long long)blockIdx.x * a.tiles_per_chunk;
    const long long t1 = min(ntiles, t0 + a.tiles_per_chunk);
    const int dpp = a.cout / 8, xpp = a.cin / 8;  // 16-byte parts per pixel
    f32x4 acc[kW1MaxIb];
#pragma unroll
    for (int ib = 0; ib < kW1MaxIb; ++ib) acc[ib] = f32x4{0.f, 0.f, 0.f, 0.f};
    float bsum = 0.f;
    for (long long t = t0; t < t1; ++t) {
        const long long p0 = t * 64;
        __syncthreads();  // the previous tile is consumed
        for (int q = tid; q < 64 * dpp; q += nthr) {
            const int pp = q / dpp, part = q - pp * dpp;
            uint4 v = make_uint4(0u, 0u, 0u, 0u);
            if (p0 + pp < npix) v = *reinterpret_cast<const uint4*>(a.dout + (p0 + pp) * a.dout_ld + 8 * part);
            *reinterpret_cast<uint4*>(s_d + 8 * q) = v;
        }
        for (int q = tid; q < 64 * xpp; q += nthr) {
            const int pp = q / xpp, part = q - pp * xpp;
            uint4 v = make_uint4(0u, 0u, 0u, 0u);
            if (p0 + pp < npix)
                v = *reinterpret_cast<const uint4*>(a.in + (p0 + pp) * a.in_ld + a.in_off + 8 * part);
            *reinterpret_cast<uint4*>(s_x + 8 * q) = v;
        }
        __syncthreads();
#pragma unroll
        for (int kh = 0; kh < 2; ++kh) {
            const int p = 32 * kh + 8 * g + lj;
            const bf16x8 av = lds_tr_pair(s_d + p * a.cout + 16 * wave + 4 * la,
                                          s_d + (p + 4) * a.cout + 16 * wave + 4 * la);
#pragma unroll
            for (int q = 0; q < 8; ++q) bsum += (float)av[q];
#pragma unroll
            for (int ib = 0; ib < kW1MaxIb; ++ib) {
                if (ib >= nib) break;  // uniform
                const bf16x8 bv = lds_tr_pair(s_x + p * a.cin + 16 * ib + 4 * la, s_x + (p + 4) * a.cin + 16 * ib + 4 * la);
                acc[ib] = __builtin_amdgcn_mfma_f32_16x16x32_bf16(av, bv, acc[ib], 0, 0, 0);
            }
        }
    }
    float* const prow = a.partial + (size_t)blockIdx.x * ((size_t)a.cout * a.cin + a.cout);
#pragma unroll
    for (int ib = 0; ib < kW1MaxIb; ++ib) {
        if (ib >= nib) break;
#pragma unroll
        for (int r = 0; r < 4; ++r) prow[(size_t)(16 * wave + 4 * g + r) * a.cin + 16 * ib + ln] = acc[ib][r];
    }
    float v = bsum;
    v += __shfl_xor(v, 16);
    v += __shfl_xor(v, 32);
    if (g == 0) prow[(size_t)a.cout * a.cin + 16 * wave + ln] = v;
}

// 1x1 convolution with fp32 output (the output layer, forward with bias and
// backward-data without): out[p, c] = bias[c] + sum_k in[p, k] w[c][k] on
// 16x16x32 MFMAs with both operands loaded straight from global memory (the
// weights are a few KB and stay in cache).  A wave computes 16 pixels x all
// output channels (at most 16 kC1MaxOb); slices of 32 input channels, a last
// slice of 16 padded with zeros.
constexpr int kC1MaxOb = 8;

// Training-time epilogue of the output layer: the target t (same layout as
// the output) is read alongside, the output buffer receives e = y - t on the
// first creal channels and 0 on the padding ones, and each wave adds
// sum t^2 and sum (y - t)^2 over its 16 pixels in double (fixed xor-shuffle
// order) into part[2 * (pixel / 16)], part[2 * (pixel / 16) + 1]: a wave's
// pixels lie in one sample when H * W is a multiple of 16.
struct LossArgs {
    const float* t;
    double* part;
    int creal;
};

template <bool LOSS>
__global__ __launch_bounds__(256) void conv1_k32_kernel(ConvArgs a, LossArgs l)
{
    const int lane = threadIdx.x & 63, wave = threadIdx.x >> 6;
    const int ln = lane & 15, kq = lane >> 4;
    const long long npix = (long long)a.B * a.H * a.W;
    const long long pb = ((long long)blockIdx.x * 4 + wave) * 16;
    if (pb >= npix) return;  // no barriers below
    const long long p = pb + ln;
    const bool pok = p < npix;
    const int nob = a.cout / 16;
    f32x4 acc[kC1MaxOb];
#pragma unroll
    for (int ob = 0; ob < kC1MaxOb; ++ob) acc[ob] = f32x4{0.f, 0.f, 0.f, 0.f};
    // training: the target's values for the epilogue, loaded before the
    // operand loads so that their latency overlaps the MFMAs
    float4 tpre[LOSS ? kC1MaxOb : 1];
    if constexpr (LOSS) {
#pragma unroll
        for (int ob = 0; ob < kC1MaxOb; ++ob) {
            if (ob >= nob) break;  // uniform
            tpre[ob] = pok ? *reinterpret_cast<const float4*>(l.t + p * a.out_ld + a.out_off + 16 * ob + 4 * kq)
                           : make_float4(0.f, 0.f, 0.f, 0.f);
        }
    }
    for (int k0 = 0; k0 < a.cin; k0 += 32) {
        const int kk = k0 + 8 * kq;
        const bool kin = kk < a.cin;
        uint4 b = make_uint4(0u, 0u, 0u, 0u);
        if (pok && kin) b = *reinterpret_cast<const uint4*>(a.in + p * a.in_ld + a.in_off + kk);
        const bf16x8 bv = __builtin_bit_cast(bf16x8, b);
#pragma unroll
        for (int ob = 0; ob < kC1MaxOb; ++ob) {
            if (ob >= nob) break;  // uniform
            uint4 w = make_uint4(0u, 0u, 0u, 0u);
            if (kin) w = *reinterpret_cast<const uint4*>(a.w + (size_t)(16 * ob + ln) * a.cin + kk);
            acc[ob] = __builtin_amdgcn_mfma_f32_16x16x32_bf16(__builtin_bit_cast(bf16x8, w), bv, acc[ob], 0, 0, 0);
        }
    }
    if constexpr (!LOSS) {
        if (!pok) return;
    }
    double tt = 0.0, ee = 0.0;
#pragma unroll
    for (int ob = 0; ob < kC1MaxOb; ++ob) {
        if (ob >= nob) break;
        const int c = 16 * ob + 4 * kq;
        float v[4] = {acc[ob][0], acc[ob][1], acc[ob][2], acc[ob][3]};
        if (a.bias) {
            const float4 b4 = *reinterpret_cast<const float4*>(a.bias + c);
            v[0] += b4.x, v[1] += b4.y, v[2] += b4.z, v[3] += b4.w;
        }
        if (a.relu)
#pragma unroll
            for (int q = 0; q < 4; ++q) v[q] = fmaxf(v[q], 0.f);
        if constexpr (LOSS) {
            if (pok) {
                const float4 t4 = tpre[ob];
                const float tq[4] = {t4.x, t4.y, t4.z, t4.w};
#pragma unroll
                for (int q = 0; q < 4; ++q) {
                    if (c + q < l.creal) {
                        const double tv = tq[q], dv = (double)v[q] - tv;
                        tt += tv * tv;
                        ee += dv * dv;
                        v[q] = v[q] - tq[q];
                    } else {
                        v[q] = 0.f;
                    }
                }
            }
        }
        if (pok)
            *reinterpret_cast<float4*>(static_cast<float*>(a.out) + p * a.out_ld + a.out_off + c) =
                make_float4(v[0], v[1], v[2], v[3]);
    }
    if constexpr (LOSS) {
#pragma unroll
        for (int m = 32; m >= 1; m >>= 1) {
            tt += __shfl_xor(tt, m);
            ee += __shfl_xor(ee, m);
        }
        if (lane == 0) {
            l.part[2 * (pb / 16)] = tt;
            l.part[2 * (pb / 16) + 1] = ee;
        }
    }
}

// out[i] = sum_{c < chunks} partial[c n + i] in a fixed order: a workgroup of
// 16 waves owns 64 consecutive outputs; wave w sums the chunks c = w, w + 16,
// ... with two interleaved running sums, and the 16 wave sums are then added
// in wave order, so the result does not depend on the launch.
constexpr int kRedWaves = 16;

__global__ __launch_bounds__(64 * kRedWaves) void reduce_chunks_kernel(const float* partial, long long n, int chunks,
                                                                        float* out)
{
    __shared__ float red[kRedWaves][64];
    const int o = threadIdx.x & 63, w = threadIdx.x >> 6;
    const long long i = (long long)blockIdx.x * 64 + o;
    float s0 = 0.f, s1 = 0.f;
    if (i < n) {
        int c = w;
        for (; c + kRedWaves < chunks; c += 2 * kRedWaves) {
            s0 += partial[(size_t)c * n + i];
            s1 += partial[(size_t)(c + kRedWaves) * n + i];
        }
        if (c < chunks) s0 += partial[(size_t)c * n + i];
    }
    red[w][o] = s0 + s1;
    __syncthreads();
    if (w == 0 && i < n) {
        float tot = 0.f;
#pragma unroll
        for (int k = 0; k < kRedWaves; ++k) tot += red[k][o];
        out[i] = tot;
    }
}

// The same sums for chunks <= kRedWaves, one thread per output: wave c of
// reduce_chunks_kernel then holds chunk c alone, so both add
// 0 + c_0 + c_1 + ... in chunk order and give the same bits, but here every
// thread has work (layers with many weights and few chunks).
__global__ __launch_bounds__(256) void reduce_few_chunks_kernel(const float* partial, long long n, int chunks,
                                                                float* out)
{
    const long long i = (long long)blockIdx.x * blockDim.x + threadIdx.x;
    if (i >= n) return;
    float tot = 0.f;
    for (int c = 0; c < chunks; ++c) tot += partial[(size_t)c * n + i];
    out[i] = tot;
}

void reduce_chunks(hipStream_t s, const float* partial, long long n, int chunks, float* out)
{
    if (chunks <= kRedWaves)
        hipLaunchKernelGGL(reduce_few_chunks_kernel, dim3((unsigned)((n + 255) / 256)), dim3(256), 0, s, partial, n,
                           chunks, out);
    else
        hipLaunchKernelGGL(reduce_chunks_kernel, dim3((unsigned)((n + 63) / 64)), dim3(64 * kRedWaves), 0, s, partial,
                           n, chunks, out);
}

// partial[chunk][c] = sum over the chunk's kBiasChunk pixels of d[p, c].
// Thread (g, c) of the block (g < G = 256 / C groups) sums pixels g, g + G,
// ... for channel c (consecutive threads read consecutive channels) with two
// interleaved running sums; the G group sums are then added in group order.
constexpr int kBiasChunk = 256;

__global__ __launch_bounds__(256) void bias_partial_kernel(const uint16_t* d, int C, int npix, float* partial)
{
    __shared__ float red[256];
    const int G = max(1, 256 / C);
    const int t = threadIdx.x, g = t / C, c = t % C;
    const int pbeg = blockIdx.x * kBiasChunk, pend = min(npix, pbeg + kBiasChunk);
    float s0 = 0.f, s1 = 0.f;
    if (g < G) {
        int p = pbeg + g;
        for (; p + G < pend; p += 2 * G) {
            s0 += bf2f(d[(size_t)p * C + c]);
            s1 += bf2f(d[(size_t)(p + G) * C + c]);
        }
        if (p < pend) s0 += bf2f(d[(size_t)p * C + c]);
    }
    red[t] = s0 + s1;
    __syncthreads();
    if (t < C) {
        float tot = 0.f;
        for (int k = 0; k < G; ++k) tot += red[k * C + t];
        partial[(size_t)blockIdx.x * C + t] = tot;
    }
}

// ------------------------------------------------------------- elementwise
// 2 x 2 average pool; out has H x W pixels, in 2H x 2W.
__global__ void pool_kernel(const uint16_t* in, int in_ld, int in_off, uint16_t* out, int B, int H, int W, int C)
{
    const int i = blockIdx.x * blockDim.x + threadIdx.x;
    if (i >= B * H * W * C) return;
    const int c = i % C, p = i / C;
    const int x = p % W, t = p / W;
    const int y = t % H, b = t / H;
    const size_t q = ((size_t)b * 2 * H + 2 * y) * (2 * W) + 2 * x;
    const size_t s0 = q * in_ld + in_off + c, s1 = (q + 2 * W) * in_ld + in_off + c;
    const float v = (bf2f(in[s0]) + bf2f(in[s0 + in_ld])) + (bf2f(in[s1]) + bf2f(in[s1 + in_ld]));
    out[i] = f2bf(0.25f * v);
}

// nearest upsample; in has H/2 x W/2 pixels (contiguous, C channels)
__global__ void upsample_kernel(const uint16_t* in, uint16_t* out, int out_ld, int out_off, int B, int H, int W, int C)
{
    const int i = blockIdx.x * blockDim.x + threadIdx.x;
    if (i >= B * H * W * C) return;
    const int c = i % C, p = i / C;
    const int x = p % W, t = p / W;
    const int y = t % H, b = t / H;
    const size_t q = ((size_t)b * (H / 2) + y / 2) * (W / 2) + x / 2;
    out[(size_t)p * out_ld + out_off + c] = in[q * C + c];
}

// The same with 8 channels per thread (C, out_ld and out_off multiples of 8):
// one 16-byte load and store per thread.
__global__ void upsample8_kernel(const uint16_t* in, uint16_t* out, int out_ld, int out_off, int B, int H, int W,
                                 int C)
{
    const int C8 = C >> 3;
    const long long i = (long long)blockIdx.x * blockDim.x + threadIdx.x;
    if (i >= (long long)B * H * W * C8) return;
    const int c = (int)(i % C8) * 8;
    const long long p = i / C8;
    const int x = (int)(p % W);
    const long long t = p / W;
    const int y = (int)(t % H), b = (int)(t / H);
    const size_t q = ((size_t)b * (H / 2) + y / 2) * (W / 2) + x / 2;
    *reinterpret_cast<uint4*>(out + (size_t)p * out_ld + out_off + c) = *reinterpret_cast<const uint4*>(in + q * C + c);
}

// ReLU backward with the incoming gradient assembled from its sources:
//   g = g1[p, off1 + c]                        (up = 0), or
//   g = sum of g1 over the 2 x 2 children of p (up = 1, g1 at 2H x 2W);
//   plus 0.25 g2[parent(p), c] if g2 (average-pool backward, g2 at H/2 x W/2);
// out[p, c] = bf16(g) where act[p, c] > 0, else 0 (contiguous, C channels).
struct PrepArgs {
    const float* g1;
    const float* g2;
    const uint16_t* act;
    uint16_t* out;
    int g1_ld, g1_off, up, g2_ld, act_ld, act_off;
    int B, H, W, C;
};

__global__ void relu_grad_kernel(PrepArgs a)
{
    const int i = blockIdx.x * blockDim.x + threadIdx.x;
    if (i >= a.B * a.H * a.W * a.C) return;
    const int c = i % a.C, p = i / a.C;
    const int x = p % a.W, t = p / a.W;
    const int y = t % a.H, b = t / a.H;
    float g;
    if (a.up) {
        const size_t q = ((size_t)b * 2 * a.H + 2 * y) * (2 * a.W) + 2 * x;
        const size_t s0 = q * a.g1_ld + a.g1_off + c, s1 = (q + 2 * a.W) * a.g1_ld + a.g1_off + c;
        g = (a.g1[s0] + a.g1[s0 + a.g1_ld]) + (a.g1[s1] + a.g1[s1 + a.g1_ld]);
    } else {
        g = a.g1[(size_t)p * a.g1_ld + a.g1_off + c];
    }
    if (a.g2) {
        const size_t q = ((size_t)b * (a.H / 2) + y / 2) * (a.W / 2) + x / 2;
        g += 0.25f * a.g2[q * a.g2_ld + c];
    }
    const bool on = bf2f(a.act[(size_t)p * a.act_ld + a.act_off + c]) > 0.f;
    a.out[i] = on ? f2bf(g) : (uint16_t)0;
}

// relu_grad_kernel with 8 channels per thread (C and every leading dimension
// and offset multiples of 8): 16- and 32-byte accesses, the same arithmetic
// in the same order per element.
__global__ void relu_grad8_kernel(PrepArgs a)
{
    const int C8 = a.C >> 3;
    const long long i = (long long)blockIdx.x * blockDim.x + threadIdx.x;
    if (i >= (long long)a.B * a.H * a.W * C8) return;
    const int c = (int)(i % C8) * 8;
    const long long p = i / C8;
    const int x = (int)(p % a.W);
    const long long t = p / a.W;
    const int y = (int)(t % a.H), b = (int)(t / a.H);
    auto load8 = [](const float* src, float* d) {
        const float4 u = *reinterpret_cast<const float4*>(src), v = *reinterpret_cast<const float4*>(src + 4);
        d[0] = u.x, d[1] = u.y, d[2] = u.z, d[3] = u.w, d[4] = v.x, d[5] = v.y, d[6] = v.z, d[7] = v.w;
    };
    float g[8];
    if (a.up) {
        const size_t q = ((size_t)b * 2 * a.H + 2 * y) * (2 * a.W) + 2 * x;
        const size_t s0 = q * a.g1_ld + a.g1_off + c, s1 = (q + 2 * a.W) * a.g1_ld + a.g1_off + c;
        float g00[8], g01[8], g10[8], g11[8];
        load8(a.g1 + s0, g00);
        load8(a.g1 + s0 + a.g1_ld, g01);
        load8(a.g1 + s1, g10);
        load8(a.g1 + s1 + a.g1_ld, g11);
#pragma unroll
        for (int k = 0; k < 8; ++k) g[k] = (g00[k] + g01[k]) + (g10[k] + g11[k]);
    } else {
        load8(a.g1 + (size_t)p * a.g1_ld + a.g1_off + c, g);
    }
    if (a.g2) {
        const size_t q = ((size_t)b * (a.H / 2) + y / 2) * (a.W / 2) + x / 2;
        float h[8];
        load8(a.g2 + q * a.g2_ld + c, h);
#pragma unroll
        for (int k = 0; k < 8; ++k) g[k] += 0.25f * h[k];
    }
    const uint4 av = *reinterpret_cast<const uint4*>(a.act + (size_t)p * a.act_ld + a.act_off + c);
    const uint32_t aw[4] = {av.x, av.y, av.z, av.w};
    uint32_t ow[4];
#pragma unroll
    for (int k = 0; k < 4; ++k) {
        const uint16_t lo = bf2f((uint16_t)(aw[k] & 0xffffu)) > 0.f ? f2bf(g[2 * k]) : (uint16_t)0;
        const uint16_t hi = bf2f((uint16_t)(aw[k] >> 16)) > 0.f ? f2bf(g[2 * k + 1]) : (uint16_t)0;
        ow[k] = (uint32_t)lo | ((uint32_t)hi << 16);
    }
    *reinterpret_cast<uint4*>(a.out + (size_t)p * a.C + c) = make_uint4(ow[0], ow[1], ow[2], ow[3]);
}

// Per-sample relative squared error ||y - t||^2 / ||t||^2 over the first
// creal channels and dy = 2 (y - t) / (B ||t||^2) as bf16 (0 on padding), in
// three deterministic phases: per-(sample, chunk) partial sums, per-sample
// totals in chunk order, then dy.  Training normally takes the fused route
// (conv1_k32_kernel<true>, loss_final_kernel, loss_grad_e_kernel); these
// kernels serve output layers that the 1x1 MFMA kernel does not take.
constexpr int kLossChunk = 4096;  // elements per partial

__global__ __launch_bounds__(256) void loss_partial_kernel(const float* y, const float* t, int ld, int creal, int n,
                                                           double* partial)
{
    __shared__ double red[2][256];
    const int b = blockIdx.y, chunk = blockIdx.x;
    const size_t base = (size_t)b * n;
    const int i0 = chunk * kLossChunk, i1 = min(n, i0 + kLossChunk);
    double tt = 0.0, ee = 0.0;
    for (int i = i0 + threadIdx.x; i < i1; i += 256) {
        if (i % ld >= creal) continue;
        const double tv = t[base + i], dv = (double)y[base + i] - tv;
        tt += tv * tv;
        ee += dv * dv;
    }
    red[0][threadIdx.x] = tt;
    red[1][threadIdx.x] = ee;
    __syncthreads();
    for (int w = 128; w > 0; w >>= 1) {
        if ((int)threadIdx.x < w) {
            red[0][threadIdx.x] += red[0][threadIdx.x + w];
            red[1][threadIdx.x] += red[1][threadIdx.x + w];
        }
        __syncthreads();
    }
    if (threadIdx.x == 0) {
        partial[2 * ((size_t)b * gridDim.x + chunk)] = red[0][0];
        partial[2 * ((size_t)b * gridDim.x + chunk) + 1] = red[1][0];
    }
}

// One workgroup per sample: thread k sums the chunks k, k + 256, ... in order,
// then a fixed-shape tree adds the 256 thread sums.
__global__ __launch_bounds__(256) void loss_final_kernel(const double* partial, int chunks, int B, float* loss,
                                                         float* scale)
{
    __shared__ double red[2][256];
    const int b = blockIdx.x, k = threadIdx.x;
    double tt = 0.0, ee = 0.0;
    for (int c = k; c < chunks; c += 256) {
        tt += partial[2 * ((size_t)b * chunks + c)];
        ee += partial[2 * ((size_t)b * chunks + c) + 1];
    }
    red[0][k] = tt;
    red[1][k] = ee;
    __syncthreads();
    for (int w = 128; w > 0; w >>= 1) {
        if (k < w) {
            red[0][k] += red[0][k + w];
            red[1][k] += red[1][k + w];
        }
        __syncthreads();
    }
    if (k == 0) {
        const double tn = red[0][0] > 1e-300 ? red[0][0] : 1e-300;
        loss[b] = (float)(red[1][0] / tn);
        scale[b] = (float)(2.0 / (B * tn));
    }
}

__global__ void loss_grad_kernel(const float* y, const float* t, int ld, int creal, int n, int B, const float* scale,
                                 uint16_t* dy)
{
    const size_t i = (size_t)blockIdx.x * blockDim.x + threadIdx.x;
    if (i >= (size_t)B * n) return;
    const int b = (int)(i / n), c = (int)(i % n) % ld;
    dy[i] = c < creal ? f2bf(scale[b] * (y[i] - t[i])) : (uint16_t)0;
}

// dy = scale[b] e as bf16, four per thread, from e = y - t (0 on padding
// channels) written by the output layer's training epilogue: the same bits
// as loss_grad_kernel gives from y and t.
__global__ void loss_grad_e_kernel(const float* e, int n, long long total, const float* scale, uint16_t* dy)
{
    const long long i = ((long long)blockIdx.x * blockDim.x + threadIdx.x) * 4;
    if (i >= total) return;
    const float sc = scale[i / n];
    const float4 v = *reinterpret_cast<const float4*>(e + i);
    const uint32_t lo = (uint32_t)f2bf(sc * v.x) | ((uint32_t)f2bf(sc * v.y) << 16);
    const uint32_t hi = (uint32_t)f2bf(sc * v.z) | ((uint32_t)f2bf(sc * v.w) << 16);
    *reinterpret_cast<uint2*>(dy + i) = make_uint2(lo, hi);
}

// Adam (Kingma & Ba 2015) on the fp32 master copy; c1 = 1/(1-b1^t),
// c2 = 1/(1-b2^t).  Writes the bf16 copy the forward kernels read.
__global__ void adam_kernel(float* w, float* m, float* v, const float* g, uint16_t* wbf, long long n, float lr, float b1,
                            float b2, float eps, float c1, float c2)
{
    const long long i = (long long)blockIdx.x * blockDim.x + threadIdx.x;
    if (i >= n) return;
    const float gi = g[i];
    const float mi = b1 * m[i] + (1.f - b1) * gi;
    const float vi = b2 * v[i] + (1.f - b2) * gi * gi;
    m[i] = mi;
    v[i] = vi;
    const float wi = w[i] - lr * (mi * c1) / (sqrtf(vi * c2) + eps);
    w[i] = wi;
    wbf[i] = f2bf(wi);
}

__global__ void to_bf16_kernel(const float* src, uint16_t* dst, long long n)
{
    const long long i = (long long)blockIdx.x * blockDim.x + threadIdx.x;
    if (i < n) dst[i] = f2bf(src[i]);
}

// Backward-data weights: wb[t][ci][co] = bf16(w[ntap-1-t][co][ci]).
// Every layer's flipped weights in one launch: element i of the flipped
// array belongs to the first layer l with i < end[l].
struct FlipTable {
    long long w[NLAYER_MAX], wb[NLAYER_MAX], end[NLAYER_MAX];
    long long tend[NLAYER_MAX];  // wflip_tiled_kernel: cumulative 16 x 16 tile counts
    int cout[NLAYER_MAX], cin[NLAYER_MAX], ntap[NLAYER_MAX];
    int nl;
};

// The same copies through 16 x 16 LDS tiles (every layer's channel counts
// multiples of 16): a workgroup reads 16 output x 16 input channels of one
// tap along the input channels and writes them transposed along the output
// channels, so both sides are contiguous.  Each copy is f2bf of the same
// weight, so the bits equal wflip_all_kernel's.
__global__ __launch_bounds__(256) void wflip_tiled_kernel(const float* w, uint16_t* wb, FlipTable t)
{
    __shared__ float tile[16][17];
    const long long b = blockIdx.x;
    int l = 0;
    while (b >= t.tend[l]) ++l;  // uniform across the workgroup
    const long long tb = b - (l ? t.tend[l - 1] : 0);
    const int cout = t.cout[l], cin = t.cin[l], ntap = t.ntap[l];
    const int nci = cin >> 4, nco = cout >> 4;
    const int cit = (int)(tb % nci);
    const long long r = tb / nci;
    const int cot = (int)(r % nco), tp = (int)(r / nco);
    const int ty = threadIdx.x >> 4, tx = threadIdx.x & 15;
    tile[ty][tx] = w[t.w[l] + ((long long)tp * cout + 16 * cot + ty) * cin + 16 * cit + tx];
    __syncthreads();
    wb[t.wb[l] + ((long long)(ntap - 1 - tp) * cin + 16 * cit + ty) * cout + 16 * cot + tx] = f2bf(tile[tx][ty]);
}

__global__ void wflip_all_kernel(const float* w, uint16_t* wb, FlipTable t)
{
    const long long i = (long long)blockIdx.x * blockDim.x + threadIdx.x;
    if (i >= t.end[t.nl - 1]) return;
    int l = 0;
    while (i >= t.end[l]) ++l;
    const long long j = i - t.wb[l];
    const int cout = t.cout[l], cin = t.cin[l], ntap = t.ntap[l];
    const int co = (int)(j % cout);
    const long long r = j / cout;
    const int ci = (int)(r % cin), tp = (int)(r / cin);
    wb[i] = f2bf(w[t.w[l] + ((long long)(ntap - 1 - tp) * cout + co) * cin + ci]);
}


// ------------------------------------------------------- DG <-> image maps
// Bessel functions J0 and Y0 of the first and second kind, order 0, from the
// polynomial approximations of Abramowitz & Stegun 9.4.1-9.4.3 (absolute
// error below 1e-7; ample for input features).
__device__ __forceinline__ void bessel_j0_y0(double x, double& j0, double& y0)
{
    if (x <= 3.0) {
        const double y = (x / 3.0) * (x / 3.0);
        j0 = 1.0 + y * (-2.2499997 + y * (1.2656208 + y * (-0.3163866 + y * (0.0444479 + y * (-0.0039444 + y * 0.0002100)))));
        y0 = 0.63661977236758134 * log(0.5 * x) * j0 + 0.36746691 +
             y * (0.60559366 + y * (-0.74350384 + y * (0.25300117 + y * (-0.04261214 + y * (0.00427916 - y * 0.00024846)))));
    } else {
        const double z = 3.0 / x;
        const double f0 = 0.79788456 + z * (-0.00000077 + z * (-0.00552740 + z * (-0.00009512 + z * (0.00137237 +
                          z * (-0.00072805 + z * 0.00014476)))));
        const double t0 = x - 0.78539816 + z * (-0.04166397 + z * (-0.00003954 + z * (0.00262573 + z * (-0.00054125 +
                          z * (-0.00029333 + z * 0.00013558)))));
        const double a = f0 / sqrt(x);
        j0 = a * cos(t0);
        y0 = a * sin(t0);
    }
}

// Input channels of element e (nn = n^2, G = green_gain != 0):
//   [0, nn)          gain * f(node i + n j), f = sum of the Gaussian sources
//   [nn, 2nn)   (G)  green_gain * sum_k m_k exp(-omega^2 sigma_k^2 / 2) J0(omega r_k)
//   [2nn, 3nn)  (G)  the same with Y0: the free-space Helmholtz response of
//                    point sources of mass m_k = amp_k 2 pi sigma_k^2, damped as
//                    the Gaussian's Fourier transform at wavenumber omega, with
//                    r_k = sqrt(|x - c_k|^2 + sigma_k^2) to stay finite
//   base = nn (or 3nn with G): omega_feature; base+1, base+2: element centre
//   mapped from the box to [-1, 1]; the rest 0.
__global__ void features_kernel(const double* nodes, int n, int nel, const double* src, int ns, double gain,
                                double omega, float omega_feat, double green_gain, double x0, double y0, double x1,
                                double y1, uint16_t* x, int ld)
{
    const int i = blockIdx.x * blockDim.x + threadIdx.x;
    if (i >= nel * ld) return;
    const int c = i % ld, e = i / ld;
    const int nn = n * n;
    const bool green = green_gain != 0.0;
    const int base = green ? 3 * nn : nn;
    float v = 0.f;
    if (c < base) {
        const int k0 = c % nn, part = c / nn;
        const double px = nodes[2 * ((size_t)k0 + (size_t)nn * e)], py = nodes[2 * ((size_t)k0 + (size_t)nn * e) + 1];
        double f = 0.0;
        for (int k = 0; k < ns; ++k) {
            const double dx = px - src[4 * k], dy = py - src[4 * k + 1], sg = src[4 * k + 2], amp = src[4 * k + 3];
            const double r2 = dx * dx + dy * dy;
            if (part == 0) {
                f += amp * exp(-r2 / (2.0 * sg * sg));
            } else {
                double jv, yv;
                bessel_j0_y0(omega * sqrt(r2 + sg * sg), jv, yv);
                const double m = amp * 6.283185307179586 * sg * sg * exp(-0.5 * omega * omega * sg * sg);
                f += m * (part == 1 ? jv : yv);
            }
        }
        v = (float)((part == 0 ? gain : green_gain) * f);
    } else if (c == base) {
        v = omega_feat;
    } else if (c == base + 1 || c == base + 2) {
        const int d = c - base - 1;
        double m = 0.0;
        for (int k = 0; k < nn; ++k) m += nodes[2 * ((size_t)k + (size_t)nn * e) + d];
        m /= nn;
        const double lo = d ? y0 : x0, hi = d ? y1 : x1;
        v = (float)(2.0 * (m - lo) / (hi - lo) - 1.0);
    }
    x[i] = f2bf(v);
}

__global__ void dg_to_image_kernel(const double* W, int nd, long long nel, double inv_scale, float* t, int ld)
{
    const long long i = (long long)blockIdx.x * blockDim.x + threadIdx.x;
    if (i >= nel * ld) return;
    const int c = (int)(i % ld);
    const long long e = i / ld;
    t[i] = c < nd ? (float)(W[e * nd + c] * inv_scale) : 0.f;
}

__global__ void image_to_dg_kernel(const float* y, int ld, int nd, long long nel, double scale, double* W)
{
    const long long i = (long long)blockIdx.x * blockDim.x + threadIdx.x;
    if (i >= nel * nd) return;
    const int c = (int)(i % nd);
    const long long e = i / nd;
    W[i] = (double)y[e * ld + c] * scale;
}

// ------------------------------------------------------------------- host
struct Fail {
    int code;
    std::string msg;
};

void ck(hipError_t e, const char* what)
{
    if (e != hipSuccess) throw Fail{WDG_ERR_HIP, std::string(what) + ": " + hipGetErrorString(e)};
}

template <typename F>
int guard(F&& f)
{
    try {
        f();
        return WDG_OK;
    } catch (const Fail& x) {
        wdg::g_last_error = x.msg;
        return x.code;
    } catch (const std::exception& x) {
        wdg::g_last_error = x.what();
        return WDG_ERR_INVALID;
    }
}

template <typename T>
T* dalloc(size_t n)
{
    void* p = nullptr;
    if (hipMalloc(&p, sizeof(T) * std::max<size_t>(n, 1)) != hipSuccess) {
        (void)hipGetLastError();
        throw Fail{WDG_ERR_ALLOC, "hipMalloc failed (" + std::to_string(sizeof(T) * n) + " bytes)"};
    }
    return static_cast<T*>(p);
}

inline unsigned blocks_for(long long n, int bs = 256) { return (unsigned)((n + bs - 1) / bs); }

struct Layer {
    int cin, cout, ks, level;
    long long w, b, wb;  // offsets: weights and bias in the parameter vector, flipped bf16 weights
};

// Layer indices in the parameter vector.
enum { E1A, E1B, E2A, E2B, E3A, E3B, BA, BB, D3A, D3B, D2A, D2B, D1A, D1B, OUT, NLAYER };
static_assert(NLAYER == NLAYER_MAX, "FlipTable holds every layer");

}  // namespace

struct wdg_cnn_s {
    int dev = 0, M = 0, in_ld = 0, out_ld = 0, out_real = 0, maxB = 0;
    int C[4] = {0, 0, 0, 0};
    long long nparam = 0, nflip = 0, step = 0;
    Layer L[NLAYER];
    float *w = nullptr, *m = nullptr, *v = nullptr, *g = nullptr;
    uint16_t *wbf = nullptr, *wflip = nullptr;
    // activations (bf16), per level l = 0..3 at resolution M >> l
    uint16_t *e1a = nullptr, *cat1 = nullptr, *d1a = nullptr, *d1 = nullptr;
    uint16_t *p1 = nullptr, *e2a = nullptr, *cat2 = nullptr, *d2a = nullptr, *d2 = nullptr;
    uint16_t *p2 = nullptr, *e3a = nullptr, *cat3 = nullptr, *d3a = nullptr, *d3 = nullptr;
    uint16_t *p3 = nullptr, *ba = nullptr, *bb = nullptr;
    float* y = nullptr;
    // gradients
    uint16_t *dy = nullptr, *dpre = nullptr, *dpre2 = nullptr;  // dpre2: the other pre-activation gradient
    float *gA = nullptr, *gcat1 = nullptr, *gcat2 = nullptr, *gcat3 = nullptr, *gp1 = nullptr, *gp2 = nullptr,
          *gp3 = nullptr;
    float *partial = nullptr, *loss = nullptr;
    long long partial_len = 0;
    std::vector<void*> owned;

    template <typename T>
    T* alloc(size_t n)
    {
        T* p = dalloc<T>(n);
        owned.push_back(p);
        return p;
    }
    ~wdg_cnn_s()
    {
        for (void* p : owned) (void)hipFree(p);
    }
    int res(int level) const { return M >> level; }
};

namespace {

// The transposing-load weight gradient covers 3x3 layers with 32k output and
// 16k input channels on power-of-two images.
bool wgrad_tr_ok(const Layer& L, int H, int in_ld, int in_off)
{
    return L.ks == 3 && H > 0 && (H & (H - 1)) == 0 && L.cout % 32 == 0 && L.cin % 16 == 0 && in_ld % 8 == 0 &&
           in_off % 8 == 0;
}

// The 1x1 transposing-load weight gradient: 16k output channels up to 256, 16k
// input channels up to 64.
bool wgrad1_tr_ok(const Layer& L, int in_ld, int in_off)
{
    return L.ks == 1 && L.cout % 16 == 0 && L.cout <= 256 && L.cin % 16 == 0 && L.cin <= 16 * kW1MaxIb &&
           in_ld % 8 == 0 && in_off % 8 == 0;
}

// Chunks of 64-pixel tiles for wgrad1_tr_kernel: about 512 workgroups, at
// most 16 tiles each.  Returns the chunk count.
int wgrad1_chunks(int B, int H, int* tiles_per_chunk)
{
    const long long tiles = ((long long)B * H * H + 63) / 64;
    const int tpc = (int)std::max<long long>(1, std::min<long long>(16, tiles / 512));
    if (tiles_per_chunk) *tiles_per_chunk = tpc;
    return (int)((tiles + tpc - 1) / tpc);
}

// Tiles per chunk of the transposing-load weight gradient: enough chunks for about kWgTarget
// workgroups, at most 16 tiles each.  Returns the chunk count.
int wgrad_tiled_chunks(const Layer& L, int B, int H, int* tiles_per_chunk)
{
    const int TW = std::min(H, kWgTW), TH = 64 / TW;
    const long long tiles = (long long)B * ((H + TH - 1) / TH) * (H / TW);
    const long long per = (long long)(L.cout / 32) * ((L.cin + 31) / 32);
    const int tpc = (int)std::max<long long>(1, std::min<long long>(kWgTpcMax, tiles * per / kWgTarget));
    if (tiles_per_chunk) *tiles_per_chunk = tpc;
    return (int)((tiles + tpc - 1) / tpc);
}

// The 3x3 convolutions that run on conv3_k32_kernel (the only one that
// applies a ReLU-backward mask).
bool conv_k32_ok(int cin, int cout, int ks, int W, int in_ld, int in_off)
{
    return ks == 3 && W > 0 && (W & (W - 1)) == 0 && cout % 32 == 0 && cin % 16 == 0 && in_ld % 8 == 0 &&
           in_off % 8 == 0 && (cin % 32 == 0 || cin > 32);
}

void launch_conv(const wdg_cnn_s& n, hipStream_t s, const uint16_t* in, int in_ld, int in_off, int cin,
                 const uint16_t* w, const float* bias, void* out, int out_ld, int out_off, int cout, bool f32out,
                 int B, int H, int W, int ks, bool relu, const uint16_t* mask = nullptr, int mask_ld = 0,
                 int mask_off = 0)
{
    (void)n;
    ConvArgs a{in,  w, bias, out, in_ld, in_off, cin, out_ld, out_off, cout, B, H, W, ks, relu ? 1 : 0,
               mask, mask_ld, mask_off};
    if (mask && (f32out || !conv_k32_ok(cin, cout, ks, W, in_ld, in_off)))
        throw Fail{WDG_ERR_INVALID, "a ReLU-backward mask needs the 32-channel 3x3 convolution with bf16 output"};
    const long long npix = (long long)B * H * W;
    const bool pow2w = W > 0 && (W & (W - 1)) == 0;
    if (ks == 3 && pow2w && cout % 32 == 0 && cin % 16 == 0 && in_ld % 8 == 0 && in_off % 8 == 0) {
        const int TW = std::min(W, 64), TH = 64 / TW;
        const long long tiles = (long long)B * ((H + TH - 1) / TH) * (W / TW);
        const int TA = cout % 64 == 0 ? 2 : 1;
        const dim3 grid((unsigned)tiles, (unsigned)(cout / (32 * TA)));
        if (cin % 32 == 0 || cin > 32) {  // 32-channel slices; a last slice of 16 is zero-padded
            // taller tiles on images 64 or more pixels wide while about 512 workgroups remain
            int RT = 1;
            while (RT < 4 && W >= 64 && H % (2 * RT) == 0 && tiles / (2 * RT) * (cout / (32 * TA)) >= 512) RT *= 2;
            const dim3 g2((unsigned)(tiles / RT), (unsigned)(cout / (32 * TA)));
#define WDG_K32_CASE(ta, rt)                                                                       \
    if (TA == ta && RT == rt) {                                                                    \
        if (f32out)                                                                                \
            hipLaunchKernelGGL((conv3_k32_kernel<ta, rt, true>), g2, dim3(256), 0, s, a);          \
        else                                                                                       \
            hipLaunchKernelGGL((conv3_k32_kernel<ta, rt, false>), g2, dim3(256), 0, s, a);         \
        return;                                                                                    \
    }
            WDG_K32_CASE(1, 1) WDG_K32_CASE(1, 2) WDG_K32_CASE(1, 4)
            WDG_K32_CASE(2, 1) WDG_K32_CASE(2, 2) WDG_K32_CASE(2, 4)
#undef WDG_K32_CASE
        }
        if (TA == 2) {
            if (f32out)
                hipLaunchKernelGGL((conv3_tiled_kernel<2, true>), grid, dim3(256), 0, s, a);
            else
                hipLaunchKernelGGL((conv3_tiled_kernel<2, false>), grid, dim3(256), 0, s, a);
        } else {
            if (f32out)
                hipLaunchKernelGGL((conv3_tiled_kernel<1, true>), grid, dim3(256), 0, s, a);
            else
                hipLaunchKernelGGL((conv3_tiled_kernel<1, false>), grid, dim3(256), 0, s, a);
        }
        return;
    }
    if (ks == 1 && f32out && cout % 16 == 0 && cout <= 16 * kC1MaxOb && cin % 8 == 0 && in_ld % 8 == 0 &&
        in_off % 8 == 0 && out_ld % 4 == 0 && out_off % 4 == 0) {
        hipLaunchKernelGGL(conv1_k32_kernel<false>, dim3((unsigned)((npix + 63) / 64)), dim3(256), 0, s, a,
                           LossArgs{});
        return;
    }
    const int nb = cout / 16;
    const int TA = nb % 4 == 0 ? 4 : (nb % 2 == 0 ? 2 : 1);
    int TB = 4;
    while (TB > 1 && (long long)((npix + 64 * TB - 1) / (64 * TB)) * (nb / TA) < 512) TB >>= 1;
    const dim3 grid((unsigned)((npix + 16LL * kWaves * TB - 1) / (16LL * kWaves * TB)), (unsigned)(nb / TA));
    const dim3 blk(64 * kWaves);
#define WDG_CONV_CASE(ta, tb)                                                                      \
    if (TA == ta && TB == tb) {                                                                    \
        if (f32out)                                                                                \
            hipLaunchKernelGGL((conv_kernel<ta, tb, true>), grid, blk, 0, s, a);                   \
        else                                                                                       \
            hipLaunchKernelGGL((conv_kernel<ta, tb, false>), grid, blk, 0, s, a);                  \
        return;                                                                                    \
    }
    WDG_CONV_CASE(1, 1) WDG_CONV_CASE(1, 2) WDG_CONV_CASE(1, 4)
    WDG_CONV_CASE(2, 1) WDG_CONV_CASE(2, 2) WDG_CONV_CASE(2, 4)
    WDG_CONV_CASE(4, 1) WDG_CONV_CASE(4, 2) WDG_CONV_CASE(4, 4)
#undef WDG_CONV_CASE
}

// Forward of one conv layer (bias + ReLU, bf16 out).
void fwd(const wdg_cnn_s& n, hipStream_t s, int li, const uint16_t* in, int in_ld, int in_off, uint16_t* out,
         int out_ld, int out_off, int B)
{
    const Layer& L = n.L[li];
    const int H = n.res(L.level);
    launch_conv(n, s, in, in_ld, in_off, L.cin, n.wbf + L.w, n.w + L.b, out, out_ld, out_off, L.cout, false, B, H, H,
                L.ks, true);
}

// Parameter gradients of layer li from its input and dpre (bf16, contiguous
// cout channels), then (if gin) the input gradient (fp32, cin channels).
void bwd(wdg_cnn_s& n, hipStream_t s, int li, const uint16_t* in, int in_ld, int in_off, const uint16_t* dpre,
         float* gin, int B)
{
    const Layer& L = n.L[li];
    const int H = n.res(L.level);
    const long long npix = (long long)B * H * H;
    const int chunks = (int)((npix + kChunk - 1) / kChunk);
    const int ntap = L.ks * L.ks;
    const long long nw = (long long)ntap * L.cout * L.cin;
    WgradArgs wa{in, dpre, n.partial, in_ld, in_off, L.cin, L.cout, L.cout, B, H, H, L.ks, 1};
    if (wgrad_tr_ok(L, H, in_ld, in_off)) {
        const int wchunks = wgrad_tiled_chunks(L, B, H, &wa.tiles_per_chunk);
        hipLaunchKernelGGL(wgrad3_tr_kernel, dim3(wchunks, L.cout / 32, (L.cin + 31) / 32), dim3(64 * kWgWaves), 0, s,
                           wa);
        // rows of [weights][bias]: the bias follows the weights in the parameter vector
        reduce_chunks(s, n.partial, nw + L.cout, wchunks, n.g + L.w);
        if (gin)
            launch_conv(n, s, dpre, L.cout, 0, L.cout, n.wflip + L.wb, nullptr, gin, L.cin, 0, L.cin, true, B, H, H,
                        L.ks, false);
        return;
    }
    if (wgrad1_tr_ok(L, in_ld, in_off)) {
        const int wchunks = wgrad1_chunks(B, H, &wa.tiles_per_chunk);
        const size_t lds = (size_t)64 * (L.cout + L.cin) * sizeof(uint16_t);
        hipLaunchKernelGGL(wgrad1_tr_kernel, dim3(wchunks), dim3(64 * (L.cout / 16)), lds, s, wa);
        reduce_chunks(s, n.partial, nw + L.cout, wchunks, n.g + L.w);
        if (gin)
            launch_conv(n, s, dpre, L.cout, 0, L.cout, n.wflip + L.wb, nullptr, gin, L.cin, 0, L.cin, true, B, H, H,
                        L.ks, false);
        return;
    }
    const int wchunks = chunks;
    const int tiles = ntap * (L.cout / 16) * (L.cin / 16);
    hipLaunchKernelGGL(wgrad_kernel, dim3(chunks, (tiles + kWaves - 1) / kWaves), dim3(64 * kWaves), 0, s, wa);
    reduce_chunks(s, n.partial, nw, wchunks, n.g + L.w);
    const int bchunks = (int)((npix + kBiasChunk - 1) / kBiasChunk);
    hipLaunchKernelGGL(bias_partial_kernel, dim3(bchunks), dim3(256), 0, s, dpre, L.cout, (int)npix, n.partial);
    reduce_chunks(s, n.partial, (long long)L.cout, bchunks, n.g + L.b);
    if (gin)
        launch_conv(n, s, dpre, L.cout, 0, L.cout, n.wflip + L.wb, nullptr, gin, L.cin, 0, L.cin, true, B, H, H, L.ks,
                    false);
}

void relu_grad(wdg_cnn_s& n, hipStream_t s, int level, int C, const float* g1, int g1_ld, int g1_off, bool up,
               const float* g2, int g2_ld, const uint16_t* act, int act_ld, int act_off, int B)
{
    const int H = n.res(level);
    PrepArgs a{g1, g2, act, n.dpre, g1_ld, g1_off, up ? 1 : 0, g2_ld, act_ld, act_off, B, H, H, C};
    if ((C | g1_ld | g1_off | g2_ld | act_ld | act_off) % 8 == 0)
        hipLaunchKernelGGL(relu_grad8_kernel, dim3(blocks_for((long long)B * H * H * C / 8)), dim3(256), 0, s, a);
    else
        hipLaunchKernelGGL(relu_grad_kernel, dim3(blocks_for((long long)B * H * H * C)), dim3(256), 0, s, a);
}

// Parameter gradients of layer li, then the gradient w.r.t. its pre-ReLU
// input where that input is act (leading dimension act_ld, offset act_off):
// the backward-data convolution and the ReLU backward in one pass when
// conv3_k32_kernel applies (bf16 output masked by act > 0), else the fp32
// input gradient and relu_grad_kernel.  Returns the buffer that holds the
// result (the other of n.dpre and n.dpre2 than dpre when fused, n.dpre else).
const uint16_t* bwd_relu(wdg_cnn_s& n, hipStream_t s, int li, const uint16_t* in, int in_ld, int in_off,
                         const uint16_t* dpre, const uint16_t* act, int act_ld, int act_off, int B);

void pool(wdg_cnn_s& n, hipStream_t s, int level_out, const uint16_t* in, int in_ld, int in_off, uint16_t* out, int C,
          int B)
{
    const int H = n.res(level_out);
    hipLaunchKernelGGL(pool_kernel, dim3(blocks_for((long long)B * H * H * C)), dim3(256), 0, s, in, in_ld, in_off, out,
                       B, H, H, C);
}

void upsample(wdg_cnn_s& n, hipStream_t s, int level_out, const uint16_t* in, uint16_t* out, int out_ld, int C, int B)
{
    const int H = n.res(level_out);
    if ((C | out_ld) % 8 == 0)
        hipLaunchKernelGGL(upsample8_kernel, dim3(blocks_for((long long)B * H * H * C / 8)), dim3(256), 0, s, in, out,
                           out_ld, 0, B, H, H, C);
    else
        hipLaunchKernelGGL(upsample_kernel, dim3(blocks_for((long long)B * H * H * C)), dim3(256), 0, s, in, out,
                           out_ld, 0, B, H, H, C);
}

// y_out: where the output layer writes ([B][M][M][out_ld] fp32); the
// handle's own buffer when null.
const uint16_t* bwd_relu(wdg_cnn_s& n, hipStream_t s, int li, const uint16_t* in, int in_ld, int in_off,
                         const uint16_t* dpre, const uint16_t* act, int act_ld, int act_off, int B)
{
    const Layer& L = n.L[li];
    const int H = n.res(L.level);
    if (conv_k32_ok(L.cout, L.cin, L.ks, H, L.cout, 0) && act_ld % 4 == 0 && act_off % 4 == 0) {
        uint16_t* out = dpre == n.dpre ? n.dpre2 : n.dpre;
        bwd(n, s, li, in, in_ld, in_off, dpre, nullptr, B);
        launch_conv(n, s, dpre, L.cout, 0, L.cout, n.wflip + L.wb, nullptr, out, L.cin, 0, L.cin, false, B, H, H,
                    L.ks, false, act, act_ld, act_off);
        return out;
    }
    bwd(n, s, li, in, in_ld, in_off, dpre, n.gA, B);
    relu_grad(n, s, L.level, L.cin, n.gA, L.cin, 0, false, nullptr, 0, act, act_ld, act_off, B);
    return n.dpre;
}

// The output layer's fused training epilogue applies (1x1 MFMA path, wave
// partials of 16 pixels inside one sample).
bool loss_fused_ok(const wdg_cnn_s& n)
{
    return n.out_ld % 16 == 0 && n.out_ld <= 16 * kC1MaxOb && n.L[OUT].cin % 8 == 0 && (n.M * n.M) % 16 == 0;
}

// With loss (training), the output layer writes e = y - t into n.y and the
// loss partials instead of y (see conv1_k32_kernel<true>).
void forward(wdg_cnn_s& n, hipStream_t s, const uint16_t* x, int B, float* y_out = nullptr,
             const LossArgs* loss = nullptr)
{
    const int C1 = n.C[0], C2 = n.C[1], C3 = n.C[2], C4 = n.C[3];
    fwd(n, s, E1A, x, n.in_ld, 0, n.e1a, C1, 0, B);
    fwd(n, s, E1B, n.e1a, C1, 0, n.cat1, C2 + C1, C2, B);
    pool(n, s, 1, n.cat1, C2 + C1, C2, n.p1, C1, B);
    fwd(n, s, E2A, n.p1, C1, 0, n.e2a, C2, 0, B);
    fwd(n, s, E2B, n.e2a, C2, 0, n.cat2, C3 + C2, C3, B);
    pool(n, s, 2, n.cat2, C3 + C2, C3, n.p2, C2, B);
    fwd(n, s, E3A, n.p2, C2, 0, n.e3a, C3, 0, B);
    fwd(n, s, E3B, n.e3a, C3, 0, n.cat3, C4 + C3, C4, B);
    pool(n, s, 3, n.cat3, C4 + C3, C4, n.p3, C3, B);
    fwd(n, s, BA, n.p3, C3, 0, n.ba, C4, 0, B);
    fwd(n, s, BB, n.ba, C4, 0, n.bb, C4, 0, B);
    upsample(n, s, 2, n.bb, n.cat3, C4 + C3, C4, B);
    fwd(n, s, D3A, n.cat3, C4 + C3, 0, n.d3a, C3, 0, B);
    fwd(n, s, D3B, n.d3a, C3, 0, n.d3, C3, 0, B);
    upsample(n, s, 1, n.d3, n.cat2, C3 + C2, C3, B);
    fwd(n, s, D2A, n.cat2, C3 + C2, 0, n.d2a, C2, 0, B);
    fwd(n, s, D2B, n.d2a, C2, 0, n.d2, C2, 0, B);
    upsample(n, s, 0, n.d2, n.cat1, C2 + C1, C2, B);
    fwd(n, s, D1A, n.cat1, C2 + C1, 0, n.d1a, C1, 0, B);
    fwd(n, s, D1B, n.d1a, C1, 0, n.d1, C1, 0, B);
    const Layer& L = n.L[OUT];
    if (loss) {
        const ConvArgs a{n.d1, n.wbf + L.w, n.w + L.b, n.y, C1, 0, C1, n.out_ld, 0, n.out_ld, B, n.M, n.M, 1, 0,
                         nullptr, 0, 0};
        const long long npix = (long long)B * n.M * n.M;
        hipLaunchKernelGGL(conv1_k32_kernel<true>, dim3((unsigned)((npix + 63) / 64)), dim3(256), 0, s, a, *loss);
        return;
    }
    launch_conv(n, s, n.d1, C1, 0, C1, n.wbf + L.w, n.w + L.b, y_out ? y_out : n.y, n.out_ld, 0, n.out_ld, true, B, n.M,
                n.M, 1, false);
}

// Gradients of the mean per-sample loss w.r.t. every parameter (into n.g),
// given n.dy from loss_kernel and the activations of forward().
void backward(wdg_cnn_s& n, hipStream_t s, const uint16_t* x, int B)
{
    const int C1 = n.C[0], C2 = n.C[1], C3 = n.C[2], C4 = n.C[3];
    // output 1x1 layer: dpre is dy itself (no activation)
    bwd(n, s, OUT, n.d1, C1, 0, n.dy, n.gA, B);
    relu_grad(n, s, 0, C1, n.gA, C1, 0, false, nullptr, 0, n.d1, C1, 0, B);
    const uint16_t* dp = bwd_relu(n, s, D1B, n.d1a, C1, 0, n.dpre, n.d1a, C1, 0, B);
    bwd(n, s, D1A, n.cat1, C2 + C1, 0, dp, n.gcat1, B);
    // level 1 decoder: d2 feeds up() into cat1[:, 0:C2]
    relu_grad(n, s, 1, C2, n.gcat1, C2 + C1, 0, true, nullptr, 0, n.d2, C2, 0, B);
    dp = bwd_relu(n, s, D2B, n.d2a, C2, 0, n.dpre, n.d2a, C2, 0, B);
    bwd(n, s, D2A, n.cat2, C3 + C2, 0, dp, n.gcat2, B);
    relu_grad(n, s, 2, C3, n.gcat2, C3 + C2, 0, true, nullptr, 0, n.d3, C3, 0, B);
    dp = bwd_relu(n, s, D3B, n.d3a, C3, 0, n.dpre, n.d3a, C3, 0, B);
    bwd(n, s, D3A, n.cat3, C4 + C3, 0, dp, n.gcat3, B);
    // bottleneck
    relu_grad(n, s, 3, C4, n.gcat3, C4 + C3, 0, true, nullptr, 0, n.bb, C4, 0, B);
    dp = bwd_relu(n, s, BB, n.ba, C4, 0, n.dpre, n.ba, C4, 0, B);
    bwd(n, s, BA, n.p3, C3, 0, dp, n.gp3, B);
    // encoder, deepest first: e_l receives the skip gradient and the pool gradient
    relu_grad(n, s, 2, C3, n.gcat3, C4 + C3, C4, false, n.gp3, C3, n.cat3, C4 + C3, C4, B);
    dp = bwd_relu(n, s, E3B, n.e3a, C3, 0, n.dpre, n.e3a, C3, 0, B);
    bwd(n, s, E3A, n.p2, C2, 0, dp, n.gp2, B);
    relu_grad(n, s, 1, C2, n.gcat2, C3 + C2, C3, false, n.gp2, C2, n.cat2, C3 + C2, C3, B);
    dp = bwd_relu(n, s, E2B, n.e2a, C2, 0, n.dpre, n.e2a, C2, 0, B);
    bwd(n, s, E2A, n.p1, C1, 0, dp, n.gp1, B);
    relu_grad(n, s, 0, C1, n.gcat1, C2 + C1, C2, false, n.gp1, C1, n.cat1, C2 + C1, C2, B);
    dp = bwd_relu(n, s, E1B, n.e1a, C1, 0, n.dpre, n.e1a, C1, 0, B);
    bwd(n, s, E1A, x, n.in_ld, 0, dp, nullptr, B);
}

// The flipped bf16 weights of every layer from the fp32 master weights.
void flip_weights(wdg_cnn_s& n, hipStream_t s)
{
    FlipTable t{};
    t.nl = NLAYER;
    bool tiled = true;
    long long tiles = 0;
    for (int l = 0; l < NLAYER; ++l) {
        const Layer& L = n.L[l];
        t.w[l] = L.w;
        t.wb[l] = L.wb;
        t.end[l] = L.wb + (long long)L.ks * L.ks * L.cout * L.cin;
        t.cout[l] = L.cout;
        t.cin[l] = L.cin;
        t.ntap[l] = L.ks * L.ks;
        tiled = tiled && L.cout % 16 == 0 && L.cin % 16 == 0;
        tiles += (long long)L.ks * L.ks * (L.cout / 16) * (L.cin / 16);
        t.tend[l] = tiles;
    }
    if (tiled)
        hipLaunchKernelGGL(wflip_tiled_kernel, dim3((unsigned)tiles), dim3(256), 0, s, n.w, n.wflip, t);
    else
        hipLaunchKernelGGL(wflip_all_kernel, dim3(blocks_for(n.nflip)), dim3(256), 0, s, n.w, n.wflip, t);
}

void refresh_copies(wdg_cnn_s& n, hipStream_t s)
{
    hipLaunchKernelGGL(to_bf16_kernel, dim3(blocks_for(n.nparam)), dim3(256), 0, s, n.w, n.wbf, n.nparam);
    flip_weights(n, s);
}

void loss_grad(wdg_cnn_s& n, hipStream_t s, int B, const uint16_t* x, const float* t)
{
    const int nper = n.M * n.M * n.out_ld;
    double* part = reinterpret_cast<double*>(n.partial);  // free between the weight-gradient passes
    if (loss_fused_ok(n)) {
        // the output layer writes e = y - t and per-wave partials; dy from e alone
        const LossArgs la{t, part, n.out_real};
        forward(n, s, x, B, nullptr, &la);
        const int chunks = n.M * n.M / 16;
        hipLaunchKernelGGL(loss_final_kernel, dim3(B), dim3(256), 0, s, part, chunks, B, n.loss, n.loss + n.maxB);
        const long long total = (long long)B * nper;
        hipLaunchKernelGGL(loss_grad_e_kernel, dim3((unsigned)((total / 4 + 255) / 256)), dim3(256), 0, s, n.y, nper,
                           total, n.loss + n.maxB, n.dy);
    } else {
        forward(n, s, x, B);
        const int chunks = (nper + kLossChunk - 1) / kLossChunk;
        hipLaunchKernelGGL(loss_partial_kernel, dim3(chunks, B), dim3(256), 0, s, n.y, t, n.out_ld, n.out_real, nper,
                           part);
        hipLaunchKernelGGL(loss_final_kernel, dim3(B), dim3(256), 0, s, part, chunks, B, n.loss, n.loss + n.maxB);
        hipLaunchKernelGGL(loss_grad_kernel, dim3(blocks_for((long long)B * nper)), dim3(256), 0, s, n.y, t,
                           n.out_ld, n.out_real, nper, B, n.loss + n.maxB, n.dy);
    }
    backward(n, s, x, B);
}

float read_loss(wdg_cnn_s& n, hipStream_t s, int B)
{
    std::vector<float> l(B);
    ck(hipMemcpyAsync(l.data(), n.loss, sizeof(float) * B, hipMemcpyDeviceToHost, s), "copy loss");
    ck(hipStreamSynchronize(s), "hipStreamSynchronize");
    double sum = 0.0;
    for (float v : l) sum += v;
    return (float)(sum / B);
}

void check_batch(const wdg_cnn_s* n, int B)
{
    if (!n) throw Fail{WDG_ERR_INVALID, "null network handle"};
    if (B < 1 || B > n->maxB) throw Fail{WDG_ERR_INVALID, "batch size outside 1..max_batch"};
}

void set_device(const wdg_cnn_s& n)
{
    int cur = -1;
    ck(hipGetDevice(&cur), "hipGetDevice");
    if (cur != n.dev) ck(hipSetDevice(n.dev), "hipSetDevice");
}

}  // namespace

extern "C" {

int wdg_hip_cnn_create(int M, int in_ld, int out_channels, int base_width, int max_batch, wdg_cnn* out)
{
    return guard([&] {
        if (!out) throw Fail{WDG_ERR_INVALID, "null output handle"};
        *out = nullptr;
        if (M < 8 || M % 8 != 0) throw Fail{WDG_ERR_INVALID, "image size must be a positive multiple of 8"};
        if (in_ld < 16 || in_ld % 16 != 0) throw Fail{WDG_ERR_INVALID, "input channels must be a multiple of 16"};
        if (out_channels < 1) throw Fail{WDG_ERR_INVALID, "output channels must be positive"};
        if (base_width < 16 || base_width % 16 != 0) throw Fail{WDG_ERR_INVALID, "base width must be a multiple of 16"};
        if (max_batch < 1) throw Fail{WDG_ERR_INVALID, "max_batch must be positive"};
        {
            // 32-bit element indices inside the kernels
            const long long widest = std::max<long long>({3LL * base_width, (long long)in_ld, out_channels + 15LL});
            if ((long long)max_batch * M * M * widest >= (1LL << 31))
                throw Fail{WDG_ERR_INVALID, "max_batch * M^2 * channels must stay below 2^31"};
        }
        auto n = std::make_unique<wdg_cnn_s>();
        ck(hipGetDevice(&n->dev), "hipGetDevice");
        n->M = M;
        n->in_ld = in_ld;
        n->out_real = out_channels;
        n->out_ld = (out_channels + 15) / 16 * 16;
        n->maxB = max_batch;
        for (int l = 0; l < 4; ++l) n->C[l] = base_width << l;
        const int C1 = n->C[0], C2 = n->C[1], C3 = n->C[2], C4 = n->C[3];
        const int spec[NLAYER][4] = {
            {in_ld, C1, 3, 0}, {C1, C1, 3, 0}, {C1, C2, 3, 1},      {C2, C2, 3, 1}, {C2, C3, 3, 2},
            {C3, C3, 3, 2},    {C3, C4, 3, 3}, {C4, C4, 3, 3},      {C4 + C3, C3, 3, 2}, {C3, C3, 3, 2},
            {C3 + C2, C2, 3, 1}, {C2, C2, 3, 1}, {C2 + C1, C1, 3, 0}, {C1, C1, 3, 0}, {C1, n->out_ld, 1, 0}};
        long long off = 0, foff = 0, maxpart = 0;
        for (int i = 0; i < NLAYER; ++i) {
            Layer& L = n->L[i];
            L.cin = spec[i][0];
            L.cout = spec[i][1];
            L.ks = spec[i][2];
            L.level = spec[i][3];
            const long long nw = (long long)L.ks * L.ks * L.cout * L.cin;
            L.w = off;
            L.b = off + nw;
            off += nw + L.cout;
            L.wb = foff;
            foff += nw;
            const int H = M >> L.level;
            // the largest chunk count any batch size up to max_batch uses
            long long tchunks = 0;
            for (int bb = 1; bb <= max_batch; ++bb) tchunks = std::max<long long>(tchunks, wgrad_tiled_chunks(L, bb, H, nullptr));
            const long long chunks =
                std::max<long long>(((long long)max_batch * H * H + kChunk - 1) / kChunk, tchunks);
            const long long bchunks = ((long long)max_batch * H * H + kBiasChunk - 1) / kBiasChunk;
            long long c1 = 0;  // 1x1 transposing-load chunks
            for (int bb = 1; bb <= max_batch; ++bb) c1 = std::max<long long>(c1, wgrad1_chunks(bb, H, nullptr));
            maxpart = std::max({maxpart, chunks * nw, tchunks * (nw + L.cout), c1 * (nw + L.cout), bchunks * L.cout});
        }
        // the loss partials (two doubles per sample and chunk) reuse the buffer
        const long long loss_chunks =
            std::max(((long long)M * M * n->out_ld + kLossChunk - 1) / kLossChunk, (long long)M * M / 16);
        maxpart = std::max(maxpart, 4LL * max_batch * loss_chunks);
        n->nparam = off;
        n->nflip = foff;
        n->partial_len = maxpart;
        n->w = n->alloc<float>(off);
        n->m = n->alloc<float>(off);
        n->v = n->alloc<float>(off);
        n->g = n->alloc<float>(off);
        n->wbf = n->alloc<uint16_t>(off);
        n->wflip = n->alloc<uint16_t>(foff);
        n->partial = n->alloc<float>(maxpart);
        n->loss = n->alloc<float>(2 * (size_t)max_batch);  // per-sample loss, then dy scale
        const long long B = max_batch;
        auto px = [&](int level) { return B * (long long)(M >> level) * (M >> level); };
        n->e1a = n->alloc<uint16_t>(px(0) * C1);
        n->cat1 = n->alloc<uint16_t>(px(0) * (C2 + C1));
        n->d1a = n->alloc<uint16_t>(px(0) * C1);
        n->d1 = n->alloc<uint16_t>(px(0) * C1);
        n->p1 = n->alloc<uint16_t>(px(1) * C1);
        n->e2a = n->alloc<uint16_t>(px(1) * C2);
        n->cat2 = n->alloc<uint16_t>(px(1) * (C3 + C2));
        n->d2a = n->alloc<uint16_t>(px(1) * C2);
        n->d2 = n->alloc<uint16_t>(px(1) * C2);
        n->p2 = n->alloc<uint16_t>(px(2) * C2);
        n->e3a = n->alloc<uint16_t>(px(2) * C3);
        n->cat3 = n->alloc<uint16_t>(px(2) * (C4 + C3));
        n->d3a = n->alloc<uint16_t>(px(2) * C3);
        n->d3 = n->alloc<uint16_t>(px(2) * C3);
        n->p3 = n->alloc<uint16_t>(px(3) * C3);
        n->ba = n->alloc<uint16_t>(px(3) * C4);
        n->bb = n->alloc<uint16_t>(px(3) * C4);
        n->y = n->alloc<float>(px(0) * n->out_ld);
        n->dy = n->alloc<uint16_t>(px(0) * n->out_ld);
        long long maxg = 0;
        for (int l = 0; l < 4; ++l) maxg = std::max(maxg, px(l) * (long long)n->C[l]);
        n->dpre = n->alloc<uint16_t>(std::max(maxg, px(0) * n->out_ld));
        n->dpre2 = n->alloc<uint16_t>(std::max(maxg, px(0) * n->out_ld));
        n->gA = n->alloc<float>(maxg);
        n->gcat1 = n->alloc<float>(px(0) * (C2 + C1));
        n->gcat2 = n->alloc<float>(px(1) * (C3 + C2));
        n->gcat3 = n->alloc<float>(px(2) * (C4 + C3));
        n->gp1 = n->alloc<float>(px(1) * C1);
        n->gp2 = n->alloc<float>(px(2) * C2);
        n->gp3 = n->alloc<float>(px(3) * C3);
        ck(hipMemset(n->w, 0, sizeof(float) * off), "hipMemset");
        ck(hipMemset(n->m, 0, sizeof(float) * off), "hipMemset");
        ck(hipMemset(n->v, 0, sizeof(float) * off), "hipMemset");
        refresh_copies(*n, nullptr);
        ck(hipDeviceSynchronize(), "hipDeviceSynchronize");
        *out = n.release();
    });
}

int wdg_hip_cnn_destroy(wdg_cnn n)
{
    delete n;
    return WDG_OK;
}

int wdg_hip_cnn_info(wdg_cnn n, long long* n_params, int* in_ld, int* out_ld)
{
    return guard([&] {
        if (!n) throw Fail{WDG_ERR_INVALID, "null network handle"};
        if (n_params) *n_params = n->nparam;
        if (in_ld) *in_ld = n->in_ld;
        if (out_ld) *out_ld = n->out_ld;
    });
}

int wdg_hip_cnn_set_params(wdg_cnn n, const float* params)
{
    return guard([&] {
        if (!n || !params) throw Fail{WDG_ERR_INVALID, "null argument"};
        set_device(*n);
        ck(hipMemcpy(n->w, params, sizeof(float) * n->nparam, hipMemcpyHostToDevice), "upload params");
        ck(hipMemset(n->m, 0, sizeof(float) * n->nparam), "hipMemset");
        ck(hipMemset(n->v, 0, sizeof(float) * n->nparam), "hipMemset");
        n->step = 0;
        refresh_copies(*n, nullptr);
        ck(hipDeviceSynchronize(), "hipDeviceSynchronize");
    });
}

int wdg_hip_cnn_get_params(wdg_cnn n, float* params)
{
    return guard([&] {
        if (!n || !params) throw Fail{WDG_ERR_INVALID, "null argument"};
        set_device(*n);
        ck(hipMemcpy(params, n->w, sizeof(float) * n->nparam, hipMemcpyDeviceToHost), "download params");
    });
}

int wdg_hip_cnn_forward(wdg_cnn n, int batch, const uint16_t* d_x, float* d_y, void* stream)
{
    return guard([&] {
        check_batch(n, batch);
        if (!d_x || !d_y) throw Fail{WDG_ERR_INVALID, "null device pointer"};
        set_device(*n);
        const hipStream_t s = static_cast<hipStream_t>(stream);
        forward(*n, s, d_x, batch, d_y);
        ck(hipGetLastError(), "forward");
    });
}

int wdg_hip_cnn_loss_grad(wdg_cnn n, int batch, const uint16_t* d_x, const float* d_t, float* loss, float* grad,
                          void* stream)
{
    return guard([&] {
        check_batch(n, batch);
        if (!d_x || !d_t) throw Fail{WDG_ERR_INVALID, "null device pointer"};
        set_device(*n);
        const hipStream_t s = static_cast<hipStream_t>(stream);
        loss_grad(*n, s, batch, d_x, d_t);
        ck(hipGetLastError(), "loss_grad");
        if (grad)
            ck(hipMemcpyAsync(grad, n->g, sizeof(float) * n->nparam, hipMemcpyDeviceToHost, s), "download gradient");
        const float l = read_loss(*n, s, batch);
        if (loss) *loss = l;
    });
}

int wdg_hip_cnn_train_step(wdg_cnn n, int batch, const uint16_t* d_x, const float* d_t, float lr, float* loss,
                           void* stream)
{
    return guard([&] {
        check_batch(n, batch);
        if (!d_x || !d_t) throw Fail{WDG_ERR_INVALID, "null device pointer"};
        if (!(lr > 0.f)) throw Fail{WDG_ERR_INVALID, "learning rate must be positive"};
        set_device(*n);
        const hipStream_t s = static_cast<hipStream_t>(stream);
        loss_grad(*n, s, batch, d_x, d_t);
        n->step += 1;
        const float b1 = 0.9f, b2 = 0.999f, eps = 1e-8f;
        const float c1 = (float)(1.0 / (1.0 - std::pow((double)b1, (double)n->step)));
        const float c2 = (float)(1.0 / (1.0 - std::pow((double)b2, (double)n->step)));
        hipLaunchKernelGGL(adam_kernel, dim3(blocks_for(n->nparam)), dim3(256), 0, s, n->w, n->m, n->v, n->g, n->wbf,
                           n->nparam, lr, b1, b2, eps, c1, c2);
        flip_weights(*n, s);
        ck(hipGetLastError(), "train_step");
        if (loss) *loss = read_loss(*n, s, batch);
    });
}

int wdg_hip_f32_to_bf16(const float* d_src, uint16_t* d_dst, long long n, void* stream)
{
    return guard([&] {
        if (n < 0 || (n && (!d_src || !d_dst))) throw Fail{WDG_ERR_INVALID, "bad arguments"};
        if (n == 0) return;
        hipLaunchKernelGGL(to_bf16_kernel, dim3(blocks_for(n)), dim3(256), 0, static_cast<hipStream_t>(stream), d_src,
                           d_dst, n);
        ck(hipGetLastError(), "to_bf16");
    });
}

int wdg_hip_ws_features(const double* d_nodes, int n_colloc, int nel, int n_sources, const double* d_sources,
                        double gain, double omega, double omega_feature, double green_gain, const double* box, int ld,
                        uint16_t* d_x, void* stream)
{
    return guard([&] {
        if (!d_nodes || !d_x || !box || (n_sources > 0 && !d_sources)) throw Fail{WDG_ERR_INVALID, "null argument"};
        const int nn = n_colloc * n_colloc, need = (green_gain != 0.0 ? 3 * nn : nn) + 3;
        if (n_colloc < 2 || nel < 1 || n_sources < 0 || ld < need)
            throw Fail{WDG_ERR_INVALID, "bad sizes (ld must hold every feature channel)"};
        if ((long long)nel * ld >= (1LL << 31)) throw Fail{WDG_ERR_INVALID, "image too large"};
        if (!(box[2] > box[0]) || !(box[3] > box[1])) throw Fail{WDG_ERR_INVALID, "empty bounding box"};
        hipLaunchKernelGGL(features_kernel, dim3(blocks_for((long long)nel * ld)), dim3(256), 0,
                           static_cast<hipStream_t>(stream), d_nodes, n_colloc, nel, d_sources, n_sources, gain, omega,
                           (float)omega_feature, green_gain, box[0], box[1], box[2], box[3], d_x, ld);
        ck(hipGetLastError(), "features");
    });
}

int wdg_hip_ws_dg_to_image(const double* d_W, int n_colloc, int nel, double scale, int ld, float* d_t, void* stream)
{
    return guard([&] {
        const int nd = 3 * n_colloc * n_colloc;
        if (!d_W || !d_t || nel < 1 || ld < nd || !(scale > 0.0)) throw Fail{WDG_ERR_INVALID, "bad arguments"};
        hipLaunchKernelGGL(dg_to_image_kernel, dim3(blocks_for((long long)nel * ld)), dim3(256), 0,
                           static_cast<hipStream_t>(stream), d_W, nd, (long long)nel, 1.0 / scale, d_t, ld);
        ck(hipGetLastError(), "dg_to_image");
    });
}

int wdg_hip_ws_image_to_dg(const float* d_y, int ld, int n_colloc, int nel, double scale, double* d_W, void* stream)
{
    return guard([&] {
        const int nd = 3 * n_colloc * n_colloc;
        if (!d_W || !d_y || nel < 1 || ld < nd) throw Fail{WDG_ERR_INVALID, "bad arguments"};
        hipLaunchKernelGGL(image_to_dg_kernel, dim3(blocks_for((long long)nel * nd)), dim3(256), 0,
                           static_cast<hipStream_t>(stream), d_y, ld, nd, (long long)nel, scale, d_W);
        ck(hipGetLastError(), "image_to_dg");
    });
}

}  // extern "C"
