// cnn_ops.hip -- the learned orbital preconditioner: a 4-layer 3D CNN
// (1 -> 16 -> 16 -> 16 -> 1 channels, 3^3 taps, zero padding on each box,
// bias, ReLU after layers 0-2, scalar skip y = s x + conv3(h3)) with BF16
// weights and activations and FP32 accumulation on MFMA, its backward pass,
// Adam training, and inference on a box set.  It has no counterpart in the
// reference; it maps an orbital residual to a preconditioned correction in
// place of the V-cycle (Preconditioning::mg, src/Preconditioning.cc:155-216),
// which stays the fallback.
//
// Layouts.  Samples are cubes of edge e, sample s at s*e^3 (the packed box
// layout of a box set whose boxes all have edge e).  Activations are BF16
// [voxel][16] (channel fastest).  Master parameters are FP32, packed
//   w0[27][16][1] b0[16] w1[27][16][16] b1[16] w2[27][16][16] b2[16]
//   w3[27][1][16] b3[1] s
// with tap = (dx*3 + dy)*3 + dz for the offset (dx-1, dy-1, dz-1).
//
// Convolution as MFMA (v_mfma_f32_16x16x16_bf16): one wave computes a tile
// of 16 z-consecutive voxels x 16 output channels, D[cout][voxel] =
// sum_tap W_tap[cout][cin] X[cin][voxel + off(tap)]: A = weights (held in
// registers for all taps), B = input rows (8-byte loads, 16 voxels x 32 B
// contiguous per tap).  Lane l holds A[l%16][4(l/16)+i], B[4(l/16)+i][l%16]
// and D[4(l/16)+i][l%16].
#include "common.hpp"

#include <cmath>
#include <cstring>
#include <vector>

namespace mgx
{
namespace
{

constexpr int kCh      = 16;
constexpr int kTaps    = 27;
constexpr int kWaves   = 4;  // per block
constexpr int kDwWaves = 1024;

// parameter offsets (floats)
constexpr int kW0 = 0, kB0 = kW0 + 27 * 16, kW1 = kB0 + 16, kB1 = kW1 + 27 * 256, kW2 = kB1 + 16,
              kB2 = kW2 + 27 * 256, kW3 = kB2 + 16, kB3 = kW3 + 27 * 16, kS = kB3 + 1, kNP = kS + 1;

typedef short s4 __attribute__((ext_vector_type(4)));
typedef short s8 __attribute__((ext_vector_type(8)));
typedef __bf16 bf8 __attribute__((ext_vector_type(8)));
typedef __bf16 bf2 __attribute__((ext_vector_type(2)));
typedef float f4 __attribute__((ext_vector_type(4)));

// v_mfma_f32_16x16x32_bf16: lane l holds A[l%16][8(l/16)..8(l/16)+7],
// B[8(l/16)..8(l/16)+7][l%16], D[4(l/16)+i][l%16]
__device__ __forceinline__ f4 mfma32(s8 a, s8 b, f4 c)
{
    return __builtin_amdgcn_mfma_f32_16x16x32_bf16(__builtin_bit_cast(bf8, a), __builtin_bit_cast(bf8, b), c, 0, 0,
        0);
}

__host__ __device__ inline uint16_t f2bf(float f)
{
    uint32_t u;
    memcpy(&u, &f, 4);
    u += 0x7FFFu + ((u >> 16) & 1u);
    return (uint16_t)(u >> 16);
}

__host__ __device__ inline float bf2f(uint16_t h)
{
    const uint32_t u = (uint32_t)h << 16;
    float f;
    memcpy(&f, &u, 4);
    return f;
}

__device__ inline int tap_dx(int t) { return t / 9 - 1; }
__device__ inline int tap_dy(int t) { return (t / 3) % 3 - 1; }
__device__ inline int tap_dz(int t) { return t % 3 - 1; }

struct TileIx
{
    int s, x, y, z0;
};

__device__ inline TileIx decode(int64_t t, int e, int nzt)
{
    TileIx r;
    const int64_t per = (int64_t)e * e * nzt;
    r.s               = (int)(t / per);
    int64_t q         = t - (int64_t)r.s * per;
    r.x               = (int)(q / ((int64_t)e * nzt));
    q -= (int64_t)r.x * e * nzt;
    r.y  = (int)(q / nzt);
    r.z0 = (int)(q - (int64_t)r.y * nzt) * 16;
    return r;
}

enum
{
    CONV_FWD_RELU = 0,  // + bias, ReLU, BF16 out
    CONV_FWD_LIN  = 1,  // + bias, FP32 out (last layer)
    CONV_BWD_MASK = 2   // no bias, times (hmask > 0), BF16 out
};

// One conv layer over nt tiles.  CIN = 1: input FP32 scalar per voxel,
// weights [16 cout][32 taps] BF16 (taps >= 27 zero).  CIN = 16: input BF16
// [voxel][16], weights [27][16 cout][16 cin] BF16.  COUT = 16: output
// [voxel][16]; COUT = 1: FP32 scalar (row 0 of D; weights rows 1..15 zero).
template <int CIN, int COUT, int MODE>
__global__ void __launch_bounds__(64 * kWaves) k_conv(const void* __restrict__ in, const uint16_t* __restrict__ wt,
    const float* __restrict__ bias, void* __restrict__ out, const uint16_t* __restrict__ hmask, int e, int nzt,
    int64_t nt)
{
    const int64_t t = (int64_t)blockIdx.x * kWaves + (threadIdx.x >> 6);
    if (t >= nt) return;  // whole wave
    const int lane = threadIdx.x & 63, n = lane & 15, g = lane >> 4;
    const TileIx T = decode(t, e, nzt);
    const int64_t sbase = (int64_t)T.s * e * e * e;
    f4 acc = { 0.f, 0.f, 0.f, 0.f };
    if (CIN == 16)
    {
        // all 27 input loads in flight before the MFMA chain
        const uint16_t* X = (const uint16_t*)in;
        s4 b[kTaps];
#pragma unroll
        for (int tap = 0; tap < kTaps; tap++)
        {
            const int xi = T.x + tap_dx(tap), yi = T.y + tap_dy(tap), zi = T.z0 + n + tap_dz(tap);
            b[tap] = (s4){ 0, 0, 0, 0 };
            if (xi >= 0 && xi < e && yi >= 0 && yi < e && zi >= 0 && zi < e)
                b[tap] = *(const s4*)(X + (sbase + ((int64_t)xi * e + yi) * e + zi) * 16 + 4 * g);
        }
#pragma unroll
        for (int tap = 0; tap < kTaps; tap++)
        {
            const s4 a = *(const s4*)(wt + ((size_t)tap * 16 + n) * 16 + 4 * g);
            acc        = __builtin_amdgcn_mfma_f32_16x16x16bf16_1k(a, b[tap], acc, 0, 0, 0);
        }
    }
    else
    {
        const float* X = (const float*)in;
#pragma unroll
        for (int h = 0; h < 2; h++)
        {
            const s4 a = *(const s4*)(wt + (size_t)n * 32 + 16 * h + 4 * g);
            s4 b;
#pragma unroll
            for (int i = 0; i < 4; i++)
            {
                const int tap = 16 * h + 4 * g + i;
                float v       = 0.f;
                if (tap < kTaps)
                {
                    const int xi = T.x + tap_dx(tap), yi = T.y + tap_dy(tap), zi = T.z0 + n + tap_dz(tap);
                    if (xi >= 0 && xi < e && yi >= 0 && yi < e && zi >= 0 && zi < e)
                        v = X[sbase + ((int64_t)xi * e + yi) * e + zi];
                }
                b[i] = (short)f2bf(v);
            }
            acc = __builtin_amdgcn_mfma_f32_16x16x16bf16_1k(a, b, acc, 0, 0, 0);
        }
    }
    const int z = T.z0 + n;
    if (z >= e) return;
    const int64_t vox = sbase + ((int64_t)T.x * e + T.y) * e + z;
    if (COUT == 1)
    {
        if (g == 0) ((float*)out)[vox] = acc[0] + (MODE == CONV_BWD_MASK ? 0.f : bias[0]);
        return;
    }
    s4 o;
#pragma unroll
    for (int i = 0; i < 4; i++)
    {
        float v = acc[i];
        if (MODE == CONV_FWD_RELU) v = fmaxf(v + bias[4 * g + i], 0.f);
        if (MODE == CONV_FWD_LIN) v += bias[4 * g + i];
        if (MODE == CONV_BWD_MASK && hmask[vox * 16 + 4 * g + i] == 0) v = 0.f;
        o[i] = (short)f2bf(v);
    }
    *(s4*)((uint16_t*)out + vox * 16 + 4 * g) = o;
}

// A 16 -> 16 layer with the input staged in LDS: a block (4 waves) owns
// rows y0..y0+3 and z0..z0+31 of one sample and marches along x, keeping the
// planes x-1, x, x+1 of its (6 x 34)-voxel window in a ring of 3 LDS slots
// (one new plane loaded per step, 16-byte coalesced loads, zeros outside the
// box).  Wave w computes row y0+w, two 16-voxel z-tiles, 27 MFMAs each, with
// all 27 taps' weights in registers and the B operands read from LDS.
constexpr int kLY = 4, kLZ = 32, kLW = kLZ + 2, kLH = kLY + 2;
constexpr int kPlane = kLH * kLW * 16;  // BF16 values per staged plane

template <int MODE>
__global__ void __launch_bounds__(256) k_conv16_lds(const uint16_t* __restrict__ in, const uint16_t* __restrict__ wt,
    const float* __restrict__ bias, uint16_t* __restrict__ out, const uint16_t* __restrict__ hmask, int e, int nyb,
    int nzb)
{
    __shared__ __align__(16) uint16_t lds[3 * kPlane];
    const int b = blockIdx.x;
    const int s = b / (nyb * nzb), yb = (b / nzb) % nyb, zb = b % nzb;
    const int y0 = yb * kLY, z0 = zb * kLZ;
    const int lane = threadIdx.x & 63, wid = threadIdx.x >> 6, n = lane & 15, g = lane >> 4;
    const int64_t e3 = (int64_t)e * e * e;
    const uint16_t* X = in + (int64_t)s * e3 * 16;
    // K = 32 per MFMA: tap pair tp, lane group g: tap 2tp + (g >> 1), channels
    // 8(g & 1)..8(g & 1)+7 (the 28th tap is zero)
    constexpr int kPairs = (kTaps + 1) / 2;
    const int tsel = g >> 1, cs = 8 * (g & 1);
    s8 wa[kPairs];
#pragma unroll
    for (int tp = 0; tp < kPairs; tp++)
    {
        const int tap = 2 * tp + tsel;
        wa[tp]        = tap < kTaps ? *(const s8*)(wt + ((size_t)tap * 16 + n) * 16 + cs) : (s8){ 0, 0, 0, 0, 0, 0, 0, 0 };
    }
    // stage plane xp into slot, as 16-byte chunks (2 per voxel)
    auto stage = [&](int xp, int slot) {
        uint4* dst = (uint4*)(lds + slot * kPlane);
        for (int q = threadIdx.x; q < kLH * kLW * 2; q += blockDim.x)
        {
            const int vox = q >> 1, half = q & 1;
            const int yy = vox / kLW, zz = vox - (vox / kLW) * kLW;
            const int y = y0 - 1 + yy, z = z0 - 1 + zz;
            uint4 v = { 0u, 0u, 0u, 0u };
            if (xp >= 0 && xp < e && y >= 0 && y < e && z >= 0 && z < e)
                v = *(const uint4*)(X + (((int64_t)xp * e + y) * e + z) * 16 + 8 * half);
            dst[q] = v;
        }
    };
    stage(-1, 2);  // zeros
    stage(0, 0);
    const int y = y0 + wid;
    for (int x = 0; x < e; x++)
    {
        stage(x + 1, (x + 1) % 3);
        __syncthreads();
        f4 acc0 = { 0.f, 0.f, 0.f, 0.f }, acc1 = acc0;
#pragma unroll
        for (int tp = 0; tp < kPairs; tp++)
        {
            // the padding tap (27) reads tap 26's operands against zero weights
            const int tap  = min(2 * tp + tsel, kTaps - 1);
            const int slot = (x + 3 + tap_dx(tap)) % 3;
            const int yy   = wid + 1 + tap_dy(tap);
            const uint16_t* row = lds + slot * kPlane + (yy * kLW) * 16;
            const int zz = n + 1 + tap_dz(tap);
            const s8 b0 = *(const s8*)(row + zz * 16 + cs);
            const s8 b1 = *(const s8*)(row + (zz + 16) * 16 + cs);
            acc0 = mfma32(wa[tp], b0, acc0);
            acc1 = mfma32(wa[tp], b1, acc1);
        }
        if (y < e)
        {
#pragma unroll
            for (int t = 0; t < 2; t++)
            {
                const int z = z0 + 16 * t + n;
                if (z >= e) continue;
                const int64_t vox = (int64_t)s * e3 + ((int64_t)x * e + y) * e + z;
                const f4 a = t == 0 ? acc0 : acc1;
                s4 o;
#pragma unroll
                for (int i = 0; i < 4; i++)
                {
                    float v = a[i];
                    if (MODE == CONV_FWD_RELU) v = fmaxf(v + bias[4 * g + i], 0.f);
                    if (MODE == CONV_BWD_MASK && hmask[vox * 16 + 4 * g + i] == 0) v = 0.f;
                    o[i] = (short)f2bf(v);
                }
                *(s4*)(out + vox * 16 + 4 * g) = o;
            }
        }
        __syncthreads();
    }
}

// The first layer (1 -> 16 channels) with the scalar input staged in LDS
// as for k_conv16_lds: a ring of 3 planes of (6 x 34) values; one 16x16x32
// MFMA per 16-voxel tile covers all 27 taps (K = tap, lane group g holds taps
// 8g..8g+7, the last 5 zero).  Bias and ReLU; BF16 out.
template <typename TIN>
__global__ void __launch_bounds__(256) k_conv1_lds(const TIN* __restrict__ in, const double* __restrict__ rms,
    const uint16_t* __restrict__ wt, const float* __restrict__ bias, uint16_t* __restrict__ out, int e, int nyb,
    int nzb)
{
    constexpr int kP1 = kLH * kLW;
    __shared__ float lds[3 * kP1];
    const int b = blockIdx.x;
    const int s = b / (nyb * nzb), yb = (b / nzb) % nyb, zb = b % nzb;
    const int y0 = yb * kLY, z0 = zb * kLZ;
    const int lane = threadIdx.x & 63, wid = threadIdx.x >> 6, n = lane & 15, g = lane >> 4;
    const int64_t e3 = (int64_t)e * e * e;
    const TIN* X = in + (int64_t)s * e3;
    // inputs R / rms (the box's normalised residual) when rms is given
    const double rs = rms ? rms[s] : 1.;
    // A: W0[cout = n][taps 8g..8g+7] from the [16][32] layout
    const s8 wa = *(const s8*)(wt + (size_t)n * 32 + 8 * g);
    float bv4[4];
#pragma unroll
    for (int i = 0; i < 4; i++)
        bv4[i] = bias[4 * g + i];
    auto stage = [&](int xp, int slot) {
        float* dst = lds + slot * kP1;
        for (int q = threadIdx.x; q < kP1; q += blockDim.x)
        {
            const int yy = q / kLW, zz = q - (q / kLW) * kLW;
            const int y = y0 - 1 + yy, z = z0 - 1 + zz;
            float v = 0.f;
            if (xp >= 0 && xp < e && y >= 0 && y < e && z >= 0 && z < e)
            {
                const TIN r = X[((int64_t)xp * e + y) * e + z];
                v           = rms ? (float)((double)r / rs) : (float)r;
            }
            dst[q] = v;
        }
    };
    stage(-1, 2);
    stage(0, 0);
    const int y = y0 + wid;
    for (int x = 0; x < e; x++)
    {
        stage(x + 1, (x + 1) % 3);
        __syncthreads();
        s8 b0, b1;
#pragma unroll
        for (int i = 0; i < 8; i++)
        {
            const int tap = 8 * g + i;
            float v0 = 0.f, v1 = 0.f;
            if (tap < kTaps)
            {
                const float* row = lds + ((x + 3 + tap_dx(tap)) % 3) * kP1 + (wid + 1 + tap_dy(tap)) * kLW;
                const int zz     = n + 1 + tap_dz(tap);
                v0               = row[zz];
                v1               = row[zz + 16];
            }
            b0[i] = (short)f2bf(v0);
            b1[i] = (short)f2bf(v1);
        }
        const f4 z4 = { 0.f, 0.f, 0.f, 0.f };
        const f4 acc0 = mfma32(wa, b0, z4), acc1 = mfma32(wa, b1, z4);
        if (y < e)
        {
#pragma unroll
            for (int t = 0; t < 2; t++)
            {
                const int z = z0 + 16 * t + n;
                if (z >= e) continue;
                const int64_t vox = (int64_t)s * e3 + ((int64_t)x * e + y) * e + z;
                const f4 a = t == 0 ? acc0 : acc1;
                s4 o;
#pragma unroll
                for (int i = 0; i < 4; i++)
                    o[i] = (short)f2bf(fmaxf(a[i] + bv4[i], 0.f));
                *(s4*)(out + vox * 16 + 4 * g) = o;
            }
        }
        __syncthreads();
    }
}

// The last layer (16 -> 1 channel) with the input staged in LDS: a block
// owns rows y0..y0+7, z0..z0+31 of one sample (one voxel per thread) and
// marches along x with a ring of 3 planes of (10 x 34) voxels; the 27 x 16
// weights are LDS broadcasts of BF16 channel pairs, applied with the BF16
// dot2 instruction (FP32 accumulation):
//   z = b + sum_tap sum_ci W[tap][ci] h[v + off(tap)][ci]   (tap-major)
// With R non-null it also finishes the network on the box: y = s x + z with
// x = (float)(R / rms), written as (T)(y * rms) to out; otherwise z to zout.
constexpr int kOY = 8, kOH = kOY + 2;
template <typename T>
__global__ void __launch_bounds__(256) k_conv_out_lds(const uint16_t* __restrict__ h, const uint16_t* __restrict__ wt,
    const float* __restrict__ prm, float* __restrict__ zout, const double* __restrict__ R,
    const double* __restrict__ rms, T* __restrict__ out, int e, int nyb, int nzb)
{
    constexpr int kPO = kOH * kLW * 16;
    __shared__ __align__(16) uint16_t lds[3 * kPO];
    __shared__ uint32_t w[kTaps * 8];  // channel pairs (2k, 2k+1) of W[tap], BF16
    for (int q = threadIdx.x; q < kTaps * 8; q += blockDim.x)
    {
        const int tap = q / 8, k = q % 8;  // row 0 of the padded [27][16][16] layout
        w[q] = (uint32_t)wt[tap * 256 + 2 * k] | ((uint32_t)wt[tap * 256 + 2 * k + 1] << 16);
    }
    const int b = blockIdx.x;
    const int s = b / (nyb * nzb), yb = (b / nzb) % nyb, zb = b % nzb;
    const int y0 = yb * kOY, z0 = zb * kLZ;
    const int64_t e3 = (int64_t)e * e * e;
    const uint16_t* X = h + (int64_t)s * e3 * 16;
    auto stage = [&](int xp, int slot) {
        uint4* dst = (uint4*)(lds + slot * kPO);
        for (int q = threadIdx.x; q < kOH * kLW * 2; q += blockDim.x)
        {
            const int vox = q >> 1, half = q & 1;
            const int yy = vox / kLW, zz = vox - (vox / kLW) * kLW;
            const int y = y0 - 1 + yy, z = z0 - 1 + zz;
            uint4 v = { 0u, 0u, 0u, 0u };
            if (xp >= 0 && xp < e && y >= 0 && y < e && z >= 0 && z < e)
                v = *(const uint4*)(X + (((int64_t)xp * e + y) * e + z) * 16 + 8 * half);
            dst[q] = v;
        }
    };
    stage(-1, 2);
    stage(0, 0);
    const int ry = threadIdx.x >> 5, rz = threadIdx.x & 31;
    const int y = y0 + ry, z = z0 + rz;
    const float bias = prm[kB3], sk = prm[kS];
    const double rs  = rms ? rms[s] : 1.;
    for (int x = 0; x < e; x++)
    {
        stage(x + 1, (x + 1) % 3);
        __syncthreads();
        float acc = bias;
        for (int tap = 0; tap < kTaps; tap++)
        {
            const uint16_t* src = lds + ((x + 3 + tap_dx(tap)) % 3) * kPO
                                  + ((ry + 1 + tap_dy(tap)) * kLW + (rz + 1 + tap_dz(tap))) * 16;
            const uint4 a = ((const uint4*)src)[0], bq = ((const uint4*)src)[1];
            const uint32_t u[8] = { a.x, a.y, a.z, a.w, bq.x, bq.y, bq.z, bq.w };
            const uint32_t* wr  = w + tap * 8;
#pragma unroll
            for (int k = 0; k < 8; k++)
                acc = __builtin_amdgcn_fdot2_f32_bf16(__builtin_bit_cast(bf2, u[k]), __builtin_bit_cast(bf2, wr[k]),
                    acc, false);
        }
        if (y < e && z < e)
        {
            const int64_t vox = (int64_t)s * e3 + ((int64_t)x * e + y) * e + z;
            if (R)
            {
                const float xv = (float)(R[vox] / rs);
                const float yv = sk * xv + acc;
                out[vox]       = (T)((double)yv * rs);
            }
            else
                zout[vox] = acc;
        }
        __syncthreads();
    }
}

// The last layer (16 -> 1 channel) on the vector units: one output voxel per
// thread, the 27 x 16 weights and the bias in LDS (FP32 values of the BF16
// weights), two 16-byte loads per tap.  z = b + sum_tap sum_ci W[tap][ci] h[v + off(tap)][ci]
// in FP32, as the MFMA path accumulates it.
__global__ void __launch_bounds__(256) k_conv_out(const uint16_t* __restrict__ h, const uint16_t* __restrict__ wt,
    const float* __restrict__ bias, float* __restrict__ out, int e, int64_t nvox)
{
    __shared__ float w[kTaps * 16];
    for (int q = threadIdx.x; q < kTaps * 16; q += blockDim.x)
        w[q] = bf2f(wt[(q / 16) * 256 + (q % 16)]);  // row 0 of the padded [27][16][16] layout
    __syncthreads();
    const int64_t v = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
    if (v >= nvox) return;
    const int64_t e3 = (int64_t)e * e * e;
    const int64_t s  = v / e3;
    int r            = (int)(v - s * e3);
    const int x = r / (e * e), y = (r / e) % e, z = r % e;
    float acc = bias[0];
    for (int tap = 0; tap < kTaps; tap++)
    {
        const int xi = x + tap_dx(tap), yi = y + tap_dy(tap), zi = z + tap_dz(tap);
        if (xi < 0 || xi >= e || yi < 0 || yi >= e || zi < 0 || zi >= e) continue;
        const uint4* p = (const uint4*)(h + (s * e3 + ((int64_t)xi * e + yi) * e + zi) * 16);
        const uint4 a = p[0], b = p[1];
        const uint32_t u[8] = { a.x, a.y, a.z, a.w, b.x, b.y, b.z, b.w };
        const float* wr = w + tap * 16;
#pragma unroll
        for (int k = 0; k < 8; k++)
        {
            acc += wr[2 * k] * bf2f((uint16_t)(u[k] & 0xFFFFu));
            acc += wr[2 * k + 1] * bf2f((uint16_t)(u[k] >> 16));
        }
    }
    out[v] = acc;
}

// Weight / bias gradient partials of one layer, per wave:
//   dW[tap][cout][cin] = sum_v dZ[v][cout] H[v + off(tap)][cin],  db = sum_v dZ
// dZ: BF16 [voxel][16] (COUT = 16) or FP32 scalar (COUT = 1); H: BF16
// [voxel][16] (CIN = 16) or FP32 scalar (CIN = 1).  A = dZ^T (m = cout,
// k = voxel), B = H shifted (k = voxel, n = cin, or n = tap for CIN = 1).
// part[w]: 27*16*16 + 16 floats in the [tap][cout][cin] layout.
template <int CIN, int COUT>
__global__ void __launch_bounds__(64) k_conv_dw(const void* __restrict__ dz, const void* __restrict__ hin, int e,
    int nzt, int64_t nt, float* __restrict__ part)
{
    const int w    = blockIdx.x;
    const int lane = threadIdx.x, n = lane & 15, g = lane >> 4;
    constexpr int NACC = CIN == 1 ? 2 : kTaps;
    f4 acc[NACC];
#pragma unroll
    for (int k = 0; k < NACC; k++)
        acc[k] = (f4){ 0.f, 0.f, 0.f, 0.f };
    float bsum = 0.f;
    for (int64_t t = w; t < nt; t += gridDim.x)
    {
        const TileIx T      = decode(t, e, nzt);
        const int64_t sbase = (int64_t)T.s * e * e * e;
        // A: dZ[v = z0 + 4g + i][cout = n]
        s4 a;
#pragma unroll
        for (int i = 0; i < 4; i++)
        {
            const int z = T.z0 + 4 * g + i;
            float v     = 0.f;
            if (z < e)
            {
                const int64_t vox = sbase + ((int64_t)T.x * e + T.y) * e + z;
                if (COUT == 16)
                    v = bf2f(((const uint16_t*)dz)[vox * 16 + n]);
                else
                    v = n == 0 ? ((const float*)dz)[vox] : 0.f;
            }
            bsum += v;
            a[i] = (short)f2bf(v);
        }
        if (CIN == 16)
        {
            const uint16_t* H = (const uint16_t*)hin;
#pragma unroll 3
            for (int tap = 0; tap < kTaps; tap++)
            {
                s4 b;
                const int xi = T.x + tap_dx(tap), yi = T.y + tap_dy(tap);
                const bool rowok = xi >= 0 && xi < e && yi >= 0 && yi < e;
#pragma unroll
                for (int i = 0; i < 4; i++)
                {
                    const int z = T.z0 + 4 * g + i, zi = z + tap_dz(tap);
                    uint16_t v  = 0;
                    if (rowok && z < e && zi >= 0 && zi < e)
                        v = H[(sbase + ((int64_t)xi * e + yi) * e + zi) * 16 + n];
                    b[i] = (short)v;
                }
                acc[tap] = __builtin_amdgcn_mfma_f32_16x16x16bf16_1k(a, b, acc[tap], 0, 0, 0);
            }
        }
        else
        {
            const float* X = (const float*)hin;
#pragma unroll
            for (int h = 0; h < 2; h++)
            {
                const int tap = 16 * h + n;
                s4 b;
#pragma unroll
                for (int i = 0; i < 4; i++)
                {
                    const int z = T.z0 + 4 * g + i;
                    float v     = 0.f;
                    if (tap < kTaps && z < e)
                    {
                        const int xi = T.x + tap_dx(tap), yi = T.y + tap_dy(tap), zi = z + tap_dz(tap);
                        if (xi >= 0 && xi < e && yi >= 0 && yi < e && zi >= 0 && zi < e)
                            v = X[sbase + ((int64_t)xi * e + yi) * e + zi];
                    }
                    b[i] = (short)f2bf(v);
                }
                acc[h] = __builtin_amdgcn_mfma_f32_16x16x16bf16_1k(a, b, acc[h], 0, 0, 0);
            }
        }
    }
    float* P = part + (size_t)w * (kTaps * 256 + 16);
    // D[m = 4g+i][n]
    if (CIN == 16)
    {
#pragma unroll
        for (int tap = 0; tap < kTaps; tap++)
#pragma unroll
            for (int i = 0; i < 4; i++)
                P[(tap * 16 + 4 * g + i) * 16 + n] = acc[tap][i];
    }
    else
    {
        // D[cout][tap]: stored at [tap][cout][0]
#pragma unroll
        for (int h = 0; h < 2; h++)
#pragma unroll
            for (int i = 0; i < 4; i++)
            {
                const int tap = 16 * h + n;
                if (tap < kTaps) P[(tap * 16 + 4 * g + i) * 16] = acc[h][i];
            }
    }
    // bias: lanes n, n+16, n+32, n+48 hold cout n's partial sums
    bsum += __shfl_xor(bsum, 16, 64);
    bsum += __shfl_xor(bsum, 32, 64);
    if (g == 0) P[kTaps * 256 + n] = bsum;
}

// sum the per-wave partials in a fixed order into grad[off..] for one layer
__global__ void k_dw_reduce(const float* part, int nw, int cin, int cout, float* gw, float* gb)
{
    const int q = blockIdx.x * blockDim.x + threadIdx.x;
    const int nwq = kTaps * cout * cin;
    if (q < nwq)
    {
        const int tap = q / (cout * cin), r = q - tap * cout * cin, co = r / cin, ci = r - co * cin;
        double s = 0.;
        for (int w = 0; w < nw; w++)
            s += part[(size_t)w * (kTaps * 256 + 16) + (tap * 16 + co) * 16 + ci];
        gw[q] = (float)s;
    }
    else if (q < nwq + cout)
    {
        const int co = q - nwq;
        double s     = 0.;
        for (int w = 0; w < nw; w++)
            s += part[(size_t)w * (kTaps * 256 + 16) + kTaps * 256 + co];
        gb[co] = (float)s;
    }
}

// loss pieces: y = s x + z; d = y - t; dy = 2 d / N; per block sums of d^2
// and dy*x (for ds), written to part[2*block]
__global__ void __launch_bounds__(256) k_loss(const float* x, const float* z, const float* t, const float* prm,
    int64_t n, float* y, float* dy, double* part)
{
    __shared__ double red[2][4];
    const float s   = prm[kS];
    const double sc = 2.0 / (double)n;
    double a = 0., b = 0.;
    for (int64_t i = blockIdx.x * (int64_t)blockDim.x + threadIdx.x; i < n; i += (int64_t)gridDim.x * blockDim.x)
    {
        const float yv = s * x[i] + z[i];
        if (y) y[i] = yv;
        if (t)
        {
            const double d = (double)yv - (double)t[i];
            const float g  = (float)(sc * d);
            dy[i]          = g;
            a += d * d;
            b += (double)g * (double)x[i];
        }
    }
    for (int o = 32; o > 0; o >>= 1)
    {
        a += __shfl_xor(a, o, 64);
        b += __shfl_xor(b, o, 64);
    }
    const int lane = threadIdx.x & 63, wid = threadIdx.x >> 6;
    if (lane == 0)
    {
        red[0][wid] = a;
        red[1][wid] = b;
    }
    __syncthreads();
    if (threadIdx.x == 0)
    {
        part[2 * blockIdx.x]     = red[0][0] + red[0][1] + red[0][2] + red[0][3];
        part[2 * blockIdx.x + 1] = red[1][0] + red[1][1] + red[1][2] + red[1][3];
    }
}

// BF16 weights for the forward and backward (flipped, transposed) passes
// from the FP32 master parameters
__global__ void k_derive(const float* p, uint16_t* f0, uint16_t* f1, uint16_t* f2, uint16_t* f3, uint16_t* r3,
    uint16_t* r2, uint16_t* r1)
{
    const int q = blockIdx.x * blockDim.x + threadIdx.x;
    // forward layer 0: [cout][32 taps]
    if (q < 16 * 32)
    {
        const int co = q / 32, tap = q % 32;
        f0[q]        = tap < kTaps ? f2bf(p[kW0 + tap * 16 + co]) : 0;
    }
    if (q < kTaps * 256)
    {
        const int tap = q / 256, co = (q / 16) % 16, ci = q % 16;
        const int ft  = kTaps - 1 - tap;
        f1[q]         = f2bf(p[kW1 + q]);
        f2[q]         = f2bf(p[kW2 + q]);
        // layer 3 forward: row 0 only
        f3[q] = co == 0 ? f2bf(p[kW3 + tap * 16 + ci]) : 0;
        // backward of layers 2, 1: W'[tap][co' = ci][ci' = co] = W[26 - tap][co][ci]
        r2[q] = f2bf(p[kW2 + (ft * 16 + ci) * 16 + co]);
        r1[q] = f2bf(p[kW1 + (ft * 16 + ci) * 16 + co]);
    }
    // backward of layer 3 (scalar input dZ3 -> 16 channels): [co' = cin][32 taps]
    if (q < 16 * 32)
    {
        const int co = q / 32, tap = q % 32;
        r3[q]        = tap < kTaps ? f2bf(p[kW3 + (kTaps - 1 - tap) * 16 + co]) : 0;
    }
}

__global__ void k_adam(float* p, const float* g, float* m, float* v, int n, float lr, float b1, float b2, float eps,
    float c1, float c2)
{
    const int q = blockIdx.x * blockDim.x + threadIdx.x;
    if (q >= n) return;
    const float gq = g[q];
    const float mq = b1 * m[q] + (1.f - b1) * gq;
    const float vq = b2 * v[q] + (1.f - b2) * gq * gq;
    m[q]           = mq;
    v[q]           = vq;
    p[q] -= lr * (mq / c1) / (sqrtf(vq / c2) + eps);
}

// per box: rms_j = sqrt(mean R_j^2) (0 -> 1); x = R/rms (FP32); t = KR/rms
__global__ void __launch_bounds__(256) k_box_rms(const double* R, int64_t m, double* rms)
{
    __shared__ double red[4];
    const int j = blockIdx.x;
    double a    = 0.;
    for (int64_t i = threadIdx.x; i < m; i += blockDim.x)
    {
        const double v = R[(int64_t)j * m + i];
        a += v * v;
    }
    for (int o = 32; o > 0; o >>= 1)
        a += __shfl_xor(a, o, 64);
    if ((threadIdx.x & 63) == 0) red[threadIdx.x >> 6] = a;
    __syncthreads();
    if (threadIdx.x == 0)
    {
        const double r = sqrt((red[0] + red[1] + red[2] + red[3]) / (double)m);
        rms[j]         = r > 0. ? r : 1.;
    }
}

__global__ void k_box_scale_to_f32(const double* R, const double* rms, int64_t m, int64_t n, float* x)
{
    for (int64_t i = blockIdx.x * (int64_t)blockDim.x + threadIdx.x; i < n; i += (int64_t)gridDim.x * blockDim.x)
        x[i] = (float)(R[i] / rms[i / m]);
}

template <typename T>
__global__ void k_box_scale_from_f32(const float* y, const double* rms, int64_t m, int64_t n, T* out)
{
    for (int64_t i = blockIdx.x * (int64_t)blockDim.x + threadIdx.x; i < n; i += (int64_t)gridDim.x * blockDim.x)
        out[i] = (T)((double)y[i] * rms[i / m]);
}

}  // namespace
}  // namespace mgx

struct mgx_cnn
{
    mgx_ctx* ctx  = nullptr;
    int edge      = 0;
    int max_batch = 0;
    int step      = 0;  // Adam steps taken
    float *p = nullptr, *g = nullptr, *m = nullptr, *v = nullptr;
    uint16_t* wbf = nullptr;  // f0 f1 f2 f3 r3 r2 r1
    uint16_t *f0, *f1, *f2, *f3, *r3, *r2, *r1;
    // activations for max_batch samples
    uint16_t *h1 = nullptr, *h2 = nullptr, *h3 = nullptr, *dza = nullptr, *dzb = nullptr;
    float *z = nullptr, *y = nullptr, *dy = nullptr, *xf = nullptr, *part = nullptr;
    double *lpart = nullptr, *rms = nullptr;
    double* host  = nullptr;  // pinned
    bool dirty    = true;     // BF16 weights need refreshing
};

namespace mgx
{
namespace
{

constexpr int kLossBlocks = 1024;

int64_t ntiles(int e, int nb) { return (int64_t)nb * e * e * ((e + 15) / 16); }

int derive(mgx_cnn* c)
{
    if (!c->dirty) return MGX_OK;
    hipLaunchKernelGGL(k_derive, dim3((kTaps * 256 + 255) / 256), dim3(256), 0, c->ctx->stream, (const float*)c->p,
        c->f0, c->f1, c->f2, c->f3, c->r3, c->r2, c->r1);
    MGX_CHECK_LAUNCH(c->ctx);
    c->dirty = false;
    return MGX_OK;
}

// forward of nb samples from FP32 x; z (pre-skip last layer) into c->z
// the network over nb samples: from FP32 x (training, z of the last layer to
// c->z), or from FP64 box residuals R with per-box rms (inference: the
// rescaled output rms * (s x + z) written to out in the type of T)
template <typename T>
int forward_any(mgx_cnn* c, const float* x, const double* R, const double* rms, int nb, T* out)
{
    mgx_ctx* ctx = c->ctx;
    int rc;
    if ((rc = derive(c))) return rc;
    const int e   = c->edge;
    const int nyb = (e + kLY - 1) / kLY, nzb = (e + kLZ - 1) / kLZ;
    const dim3 lgrid((unsigned)(nb * nyb * nzb));
    if (R)
        hipLaunchKernelGGL(k_conv1_lds<double>, lgrid, dim3(256), 0, ctx->stream, R, rms, (const uint16_t*)c->f0,
            (const float*)c->p + kB0, c->h1, e, nyb, nzb);
    else
        hipLaunchKernelGGL(k_conv1_lds<float>, lgrid, dim3(256), 0, ctx->stream, x, (const double*)nullptr,
            (const uint16_t*)c->f0, (const float*)c->p + kB0, c->h1, e, nyb, nzb);
    hipLaunchKernelGGL((k_conv16_lds<CONV_FWD_RELU>), lgrid, dim3(256), 0, ctx->stream, (const uint16_t*)c->h1,
        (const uint16_t*)c->f1, (const float*)c->p + kB1, c->h2, (const uint16_t*)nullptr, e, nyb, nzb);
    hipLaunchKernelGGL((k_conv16_lds<CONV_FWD_RELU>), lgrid, dim3(256), 0, ctx->stream, (const uint16_t*)c->h2,
        (const uint16_t*)c->f2, (const float*)c->p + kB2, c->h3, (const uint16_t*)nullptr, e, nyb, nzb);
    const int noy = (e + kOY - 1) / kOY;
    hipLaunchKernelGGL(k_conv_out_lds<T>, dim3((unsigned)(nb * noy * nzb)), dim3(256), 0, ctx->stream,
        (const uint16_t*)c->h3, (const uint16_t*)c->f3, (const float*)c->p, c->z, R, rms, out, e, noy, nzb);
    MGX_CHECK_LAUNCH(ctx);
    return MGX_OK;
}

int forward(mgx_cnn* c, const float* x, int nb)
{
    return forward_any<float>(c, x, nullptr, nullptr, nb, nullptr);
}

template <int CIN, int COUT>
int dw(mgx_cnn* c, const void* dz, const void* hin, int nb, int woff, int boff, int cin, int cout)
{
    mgx_ctx* ctx     = c->ctx;
    const int e      = c->edge, nzt = (e + 15) / 16;
    const int64_t nt = ntiles(e, nb);
    const int nw     = (int)std::min<int64_t>(kDwWaves, nt);
    hipLaunchKernelGGL((k_conv_dw<CIN, COUT>), dim3(nw), dim3(64), 0, ctx->stream, dz, hin, e, nzt, nt, c->part);
    const int nq = kTaps * cin * cout + cout;
    hipLaunchKernelGGL(k_dw_reduce, dim3((nq + 255) / 256), dim3(256), 0, ctx->stream, (const float*)c->part, nw,
        cin, cout, c->g + woff, c->g + boff);
    MGX_CHECK_LAUNCH(ctx);
    return MGX_OK;
}

// loss and gradient into c->g; loss (mean squared error) to *loss
int grad(mgx_cnn* c, const float* x, const float* t, int nb, double* loss)
{
    mgx_ctx* ctx = c->ctx;
    int rc;
    if ((rc = forward(c, x, nb))) return rc;
    const int e = c->edge, nzt = (e + 15) / 16;
    const int64_t n  = (int64_t)nb * e * e * e;
    const int64_t nt = ntiles(e, nb);
    hipLaunchKernelGGL(k_loss, dim3(kLossBlocks), dim3(256), 0, ctx->stream, x, (const float*)c->z, t,
        (const float*)c->p, n, (float*)nullptr, c->dy, c->lpart);
    MGX_CHECK_LAUNCH(ctx);
    // layer 3: dZ3 = dy (scalar)
    if ((rc = dw<16, 1>(c, c->dy, c->h3, nb, kW3, kB3, 16, 1))) return rc;
    const dim3 grid((unsigned)((nt + kWaves - 1) / kWaves)), blk(64 * kWaves);
    hipLaunchKernelGGL((k_conv<1, 16, CONV_BWD_MASK>), grid, blk, 0, ctx->stream, (const void*)c->dy, c->r3,
        (const float*)nullptr, (void*)c->dza, (const uint16_t*)c->h3, e, nzt, nt);
    MGX_CHECK_LAUNCH(ctx);
    // layer 2: dZ2 = dza
    if ((rc = dw<16, 16>(c, c->dza, c->h2, nb, kW2, kB2, 16, 16))) return rc;
    const int nyb = (e + kLY - 1) / kLY, nzb = (e + kLZ - 1) / kLZ;
    const dim3 lgrid((unsigned)(nb * nyb * nzb));
    hipLaunchKernelGGL((k_conv16_lds<CONV_BWD_MASK>), lgrid, dim3(256), 0, ctx->stream, (const uint16_t*)c->dza,
        (const uint16_t*)c->r2, (const float*)nullptr, c->dzb, (const uint16_t*)c->h2, e, nyb, nzb);
    MGX_CHECK_LAUNCH(ctx);
    // layer 1: dZ1 = dzb
    if ((rc = dw<16, 16>(c, c->dzb, c->h1, nb, kW1, kB1, 16, 16))) return rc;
    hipLaunchKernelGGL((k_conv16_lds<CONV_BWD_MASK>), lgrid, dim3(256), 0, ctx->stream, (const uint16_t*)c->dzb,
        (const uint16_t*)c->r1, (const float*)nullptr, c->dza, (const uint16_t*)c->h1, e, nyb, nzb);
    MGX_CHECK_LAUNCH(ctx);
    // layer 0: dZ0 = dza, input x
    if ((rc = dw<1, 16>(c, c->dza, x, nb, kW0, kB0, 1, 16))) return rc;
    // loss and ds
    MGX_CHECK_HIP(ctx, hipMemcpyAsync(c->host, c->lpart, 2 * kLossBlocks * sizeof(double), hipMemcpyDeviceToHost,
                           ctx->stream));
    MGX_CHECK_HIP(ctx, hipStreamSynchronize(ctx->stream));
    double a = 0., b = 0.;
    for (int k = 0; k < kLossBlocks; k++)
    {
        a += c->host[2 * k];
        b += c->host[2 * k + 1];
    }
    const float ds = (float)b;
    MGX_CHECK_HIP(ctx, hipMemcpyAsync(c->g + kS, &ds, sizeof(float), hipMemcpyHostToDevice, ctx->stream));
    MGX_CHECK_HIP(ctx, hipStreamSynchronize(ctx->stream));
    *loss = a / (double)n;
    return MGX_OK;
}

bool batch_ok(const mgx_cnn* c, int nb) { return nb >= 1 && nb <= c->max_batch; }

}  // namespace

int cnn_apply_boxes(mgx_cnn* c, int nbox, int edge, const double* R, void* out, int out_dtype);

}  // namespace mgx

using namespace mgx;

extern "C" {

int mgx_cnn_nparams(void) { return kNP; }

int mgx_cnn_create(mgx_ctx* ctx, int edge, int max_batch, mgx_cnn** out)
{
    MGX_REQUIRE(ctx, out && edge >= 1 && edge <= 128 && max_batch >= 1, "bad CNN geometry");
    mgx_cnn* c   = new mgx_cnn();
    c->ctx       = ctx;
    c->edge      = edge;
    c->max_batch = max_batch;
    const size_t nv = (size_t)max_batch * edge * edge * edge;
    const int nw    = kDwWaves;
    bool ok         = true;
    auto al         = [&](void** p, size_t bytes) { ok = ok && hipMalloc(p, std::max<size_t>(bytes, 256)) == hipSuccess; };
    al((void**)&c->p, kNP * sizeof(float));
    al((void**)&c->g, kNP * sizeof(float));
    al((void**)&c->m, kNP * sizeof(float));
    al((void**)&c->v, kNP * sizeof(float));
    al((void**)&c->wbf, (2 * 512 + 5 * kTaps * 256) * sizeof(uint16_t));
    al((void**)&c->h1, nv * 16 * sizeof(uint16_t));
    al((void**)&c->h2, nv * 16 * sizeof(uint16_t));
    al((void**)&c->h3, nv * 16 * sizeof(uint16_t));
    al((void**)&c->dza, nv * 16 * sizeof(uint16_t));
    al((void**)&c->dzb, nv * 16 * sizeof(uint16_t));
    al((void**)&c->z, nv * sizeof(float));
    al((void**)&c->y, nv * sizeof(float));
    al((void**)&c->dy, nv * sizeof(float));
    al((void**)&c->xf, nv * sizeof(float));
    al((void**)&c->part, (size_t)nw * (kTaps * 256 + 16) * sizeof(float));
    al((void**)&c->lpart, 2 * kLossBlocks * sizeof(double));
    al((void**)&c->rms, (size_t)max_batch * sizeof(double));
    ok = ok && hipHostMalloc((void**)&c->host, 2 * kLossBlocks * sizeof(double), hipHostMallocDefault) == hipSuccess;
    if (!ok)
    {
        mgx_cnn_destroy(c);
        return set_err(ctx, MGX_ENOMEM, "CNN allocation failed");
    }
    c->f0 = c->wbf;
    c->r3 = c->f0 + 512;
    c->f1 = c->r3 + 512;
    c->f2 = c->f1 + kTaps * 256;
    c->f3 = c->f2 + kTaps * 256;
    c->r2 = c->f3 + kTaps * 256;
    c->r1 = c->r2 + kTaps * 256;
    hipMemsetAsync(c->p, 0, kNP * sizeof(float), ctx->stream);
    hipMemsetAsync(c->m, 0, kNP * sizeof(float), ctx->stream);
    hipMemsetAsync(c->v, 0, kNP * sizeof(float), ctx->stream);
    if (hipStreamSynchronize(ctx->stream) != hipSuccess)
    {
        mgx_cnn_destroy(c);
        return set_err(ctx, MGX_EHIP, "CNN init failed");
    }
    *out = c;
    return MGX_OK;
}

int mgx_cnn_destroy(mgx_cnn* c)
{
    if (!c) return MGX_OK;
    void* bufs[] = { c->p, c->g, c->m, c->v, c->wbf, c->h1, c->h2, c->h3, c->dza, c->dzb, c->z, c->y, c->dy, c->xf,
        c->part, c->lpart, c->rms };
    for (void* b : bufs)
        if (b) (void)hipFree(b);
    if (c->host) (void)hipHostFree(c->host);
    delete c;
    return MGX_OK;
}

int mgx_cnn_set_params(mgx_cnn* c, const float* host)
{
    if (!c || !host) return MGX_EINVAL;
    MGX_CHECK_HIP(c->ctx, hipMemcpy(c->p, host, kNP * sizeof(float), hipMemcpyHostToDevice));
    MGX_CHECK_HIP(c->ctx, hipMemsetAsync(c->m, 0, kNP * sizeof(float), c->ctx->stream));
    MGX_CHECK_HIP(c->ctx, hipMemsetAsync(c->v, 0, kNP * sizeof(float), c->ctx->stream));
    MGX_CHECK_HIP(c->ctx, hipStreamSynchronize(c->ctx->stream));
    c->step  = 0;
    c->dirty = true;
    return MGX_OK;
}

int mgx_cnn_get_params(mgx_cnn* c, float* host)
{
    if (!c || !host) return MGX_EINVAL;
    MGX_CHECK_HIP(c->ctx, hipMemcpy(host, c->p, kNP * sizeof(float), hipMemcpyDeviceToHost));
    return MGX_OK;
}

int mgx_cnn_forward(mgx_cnn* c, const float* x, int nb, float* y)
{
    if (!c || !x || !y) return MGX_EINVAL;
    MGX_REQUIRE(c->ctx, batch_ok(c, nb), "batch larger than the CNN was created for");
    int rc = forward(c, x, nb);
    if (rc) return rc;
    const int64_t n = (int64_t)nb * c->edge * c->edge * c->edge;
    hipLaunchKernelGGL(k_loss, dim3(kLossBlocks), dim3(256), 0, c->ctx->stream, x, (const float*)c->z,
        (const float*)nullptr, (const float*)c->p, n, y, (float*)nullptr, c->lpart);
    MGX_CHECK_LAUNCH(c->ctx);
    return sync(c->ctx);
}

int mgx_cnn_grad(mgx_cnn* c, const float* x, const float* t, int nb, double* loss, float* grad_host)
{
    if (!c || !x || !t || !loss) return MGX_EINVAL;
    MGX_REQUIRE(c->ctx, batch_ok(c, nb), "batch larger than the CNN was created for");
    int rc = grad(c, x, t, nb, loss);
    if (rc) return rc;
    if (grad_host) MGX_CHECK_HIP(c->ctx, hipMemcpy(grad_host, c->g, kNP * sizeof(float), hipMemcpyDeviceToHost));
    return MGX_OK;
}

int mgx_cnn_train_step(mgx_cnn* c, const float* x, const float* t, int nb, float lr, double* loss)
{
    if (!c || !x || !t || !loss) return MGX_EINVAL;
    MGX_REQUIRE(c->ctx, batch_ok(c, nb), "batch larger than the CNN was created for");
    MGX_REQUIRE(c->ctx, lr > 0.f, "learning rate must be positive");
    int rc = grad(c, x, t, nb, loss);
    if (rc) return rc;
    c->step++;
    const float b1 = 0.9f, b2 = 0.999f;
    const float c1 = 1.f - powf(b1, (float)c->step), c2 = 1.f - powf(b2, (float)c->step);
    hipLaunchKernelGGL(k_adam, dim3((kNP + 255) / 256), dim3(256), 0, c->ctx->stream, c->p, (const float*)c->g, c->m,
        c->v, kNP, lr, b1, b2, 1e-8f, c1, c2);
    MGX_CHECK_LAUNCH(c->ctx);
    c->dirty = true;
    return MGX_OK;
}

// out_j = rms_j * net(R_j / rms_j) for nbox boxes of edge e packed one after
// another (FP64 in and out, device); processed in batches of max_batch
int mgx_cnn_apply(mgx_cnn* c, int nbox, const double* R, double* out)
{
    if (!c || !R || !out || nbox < 1) return MGX_EINVAL;
    int rc = cnn_apply_boxes(c, nbox, c->edge, R, out, MGX_F64);
    return rc ? rc : sync(c->ctx);
}

// training pairs: x = R_j / rms_j, t = KR_j / rms_j (FP32, device)
int mgx_cnn_make_pairs(mgx_cnn* c, int nbox, const double* R, const double* KR, float* x, float* t)
{
    if (!c || !R || !KR || !x || !t || nbox < 1) return MGX_EINVAL;
    mgx_ctx* ctx = c->ctx;
    const int64_t m = (int64_t)c->edge * c->edge * c->edge;
    double* rms     = nullptr;
    MGX_CHECK_HIP(ctx, hipMalloc((void**)&rms, (size_t)nbox * sizeof(double)));
    hipLaunchKernelGGL(k_box_rms, dim3(nbox), dim3(256), 0, ctx->stream, R, m, rms);
    hipLaunchKernelGGL(k_box_scale_to_f32, dim3(4096), dim3(256), 0, ctx->stream, R, (const double*)rms, m,
        m * nbox, x);
    hipLaunchKernelGGL(k_box_scale_to_f32, dim3(4096), dim3(256), 0, ctx->stream, KR, (const double*)rms, m,
        m * nbox, t);
    hipError_t e1 = hipGetLastError();
    hipError_t e2 = hipStreamSynchronize(ctx->stream);
    (void)hipFree(rms);
    if (e1 != hipSuccess || e2 != hipSuccess) return set_err(ctx, MGX_EHIP, "make_pairs failed");
    return MGX_OK;
}

}  // extern "C"

namespace mgx
{
// out (type out_dtype) = rms_j * net(R_j / rms_j) per box
int cnn_apply_boxes(mgx_cnn* c, int nbox, int edge, const double* R, void* out, int out_dtype)
{
    mgx_ctx* ctx = c->ctx;
    MGX_REQUIRE(ctx, edge == c->edge, "box edge differs from the CNN's");
    const int64_t m = (int64_t)edge * edge * edge;
    int rc;
    for (int b0 = 0; b0 < nbox; b0 += c->max_batch)
    {
        const int nb     = std::min(c->max_batch, nbox - b0);
        const double* Rb = R + (int64_t)b0 * m;
        hipLaunchKernelGGL(k_box_rms, dim3(nb), dim3(256), 0, ctx->stream, Rb, m, c->rms);
        MGX_CHECK_LAUNCH(ctx);
        rc = out_dtype == MGX_F32
                 ? forward_any<float>(c, nullptr, Rb, c->rms, nb, (float*)out + (int64_t)b0 * m)
                 : forward_any<double>(c, nullptr, Rb, c->rms, nb, (double*)out + (int64_t)b0 * m);
        if (rc) return rc;
    }
    return MGX_OK;
}
}  // namespace mgx
