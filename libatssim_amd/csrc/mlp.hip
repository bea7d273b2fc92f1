// mlp.hip -- learned warm start for the VBD solve: a per-vertex MLP
// (48 -> 128 -> 128 -> 128 -> 3, SiLU) with BF16 weights on MFMA and FP32
// accumulation, hand-written forward, backward and Adam kernels.  No
// reference counterpart (the reference starts every frame from the inertial
// guess, Core::predict_position xpbd_constraints.h:321-384).
//
// Matrix-core mapping (gfx950 v_mfma_f32_16x16x32_bf16, wave64): a wave owns
// 16 vertices (rows); for a 16-wide column block n of a layer it accumulates
// K/32 MFMAs.  Lane l holds A[l%16][8(l/16) .. 8(l/16)+7] (activations, row
// major in LDS), B[8(l/16) .. +7][l%16] (weights stored transposed so the 8 k
// values are contiguous) and C[4(l/16) + i][l%16], i = 0..3.
//
// Layout of the FP32 master parameters (row-major [in][out]): W1 (64 x 128,
// rows 48..63 zero), b1, W2 (128 x 128), b2, W3, b3, W4 (128 x 16, columns
// 3..15 zero), b4 (16).
#include <hip/hip_runtime.h>

#include <algorithm>
#include <cmath>
#include <cstring>
#include <string>
#include <vector>

#include "../../include/libatssim_hip.h"
#include "host/errors.h"
#include "vbd_context.h"

using namespace lats;

namespace {

constexpr int kIn = 48, kInP = 64, kHid = 128, kOut = 3, kOutP = 16, kOutK = 32;
constexpr int kTile = 64;   // rows per block of the backward pass (4 waves x 16)
constexpr int kT = 256;     // threads per block (backward and small kernels)
constexpr int kTF = 512;    // threads per block of the forward pass (8 waves share the LDS weights)
constexpr int kTileF = 128; // rows per forward tile (8 waves x 16)

constexpr int OFF_W1 = 0;
constexpr int OFF_B1 = OFF_W1 + kInP * kHid;
constexpr int OFF_W2 = OFF_B1 + kHid;
constexpr int OFF_B2 = OFF_W2 + kHid * kHid;
constexpr int OFF_W3 = OFF_B2 + kHid;
constexpr int OFF_B3 = OFF_W3 + kHid * kHid;
constexpr int OFF_W4 = OFF_B3 + kHid;
constexpr int OFF_B4 = OFF_W4 + kHid * kOutP;
constexpr int kNParam = OFF_B4 + kOutP;

// bf16 transposed weights for the forward B operand ([out][in]); each row
// is padded by 8 elements (16 bytes) so that the 16 rows one MFMA B operand
// reads from LDS start in different banks
constexpr int kPad = 8;
constexpr int kS1 = kInP + kPad;   // row stride of WT1
constexpr int kSH = kHid + kPad;   // row stride of WT2, WT3, WT4
constexpr int WT1 = 0;
constexpr int WT2 = WT1 + kHid * kS1;
constexpr int WT3 = WT2 + kHid * kSH;
constexpr int WT4 = WT3 + kHid * kSH;
constexpr int kNWT = WT4 + kOutP * kSH;
// bf16 row-major weights for dH = dZ W^T ([in][out], W4 padded to 32 outputs)
constexpr int WB2 = 0;
constexpr int WB3 = WB2 + kHid * kHid;
constexpr int WB4 = WB3 + kHid * kHid;
constexpr int kNWB = WB4 + kHid * kOutK;

constexpr int kActStride = kHid + 8;   // padded LDS row (bf16 elements)

typedef __bf16 bf16x8 __attribute__((ext_vector_type(8)));
typedef float f32x4 __attribute__((ext_vector_type(4)));

__host__ __device__ inline uint16_t f2bf(float f) {
    // round to nearest even (finite inputs)
    uint32_t u;
    std::memcpy(&u, &f, 4);
    u += 0x7fffu + ((u >> 16) & 1u);
    return (uint16_t)(u >> 16);
}
__host__ __device__ inline float bf2f(uint16_t b) {
    const uint32_t u = (uint32_t)b << 16;
    float f;
    std::memcpy(&f, &u, 4);
    return f;
}

__device__ inline bf16x8 ld_bf8(const uint16_t* p) {
    const uint4 v = *reinterpret_cast<const uint4*>(p);
    return __builtin_bit_cast(bf16x8, v);
}
__device__ inline f32x4 mfma(bf16x8 a, bf16x8 b, f32x4 c) {
    return __builtin_amdgcn_mfma_f32_16x16x32_bf16(a, b, c, 0, 0, 0);
}
// Row r of a batch reads recorded row map(r): identity when seed == 0, else a
// splitmix64 draw over [0, n) (sampling with replacement, same in the forward
// and the backward pass of one step).
__host__ __device__ inline uint64_t mix64(uint64_t z) {
    z += 0x9E3779B97F4A7C15ull;
    z = (z ^ (z >> 30)) * 0xBF58476D1CE4E5B9ull;
    z = (z ^ (z >> 27)) * 0x94D049BB133111EBull;
    return z ^ (z >> 31);
}
struct RowMap {
    uint64_t n = 0, seed = 0;
    __device__ uint64_t operator()(uint32_t r) const { return seed ? mix64(seed + r) % n : (uint64_t)r; }
};

// logistic sigmoid with the hardware exp and reciprocal (a few ulp; the
// activations are rounded to BF16 right after)
__device__ inline float sigmoid(float z) { return __builtin_amdgcn_rcpf(1.f + __expf(-z)); }
__device__ inline float silu(float z) { return z * sigmoid(z); }
__device__ inline float dsilu(float z) {
    const float s = sigmoid(z);
    return s * (1.f + z * (1.f - s));
}

// ---------------------------------------------------------------------------
// forward
// ---------------------------------------------------------------------------
struct FwdArgs {
    const uint16_t* X;     // [rows][64] bf16, row r of the batch at map(r)
    RowMap map;
    uint32_t M;
    const uint16_t* WT;    // transposed bf16 weights
    const float* P;        // fp32 parameters (biases)
    float* Y;              // [M][4]
    float* Z;              // [3][M][128] pre-activations (training) or null
    uint16_t* H;           // [3][M][128] bf16 activations (training) or null
};

// one dense layer for the wave's 16 rows: out[nb] = bias + A (16 x K) W^T
// (W^T rows of stride K + kPad in LDS, biases in LDS).  The NB column blocks
// are independent MFMA chains; each k step loads all NB B operands first.
template <int K, int NB>
__device__ __forceinline__ void layer_fwd(const bf16x8* a, const uint16_t* wt, const float* sbias, uint32_t l,
                                          f32x4* out) {
    const uint32_t kq = 8 * (l / 16), col = l % 16;
#pragma unroll
    for (int nb = 0; nb < NB; nb++) {
        const float bv = sbias[nb * 16 + col];
        out[nb] = f32x4{bv, bv, bv, bv};
    }
#pragma unroll
    for (int kb = 0; kb < K / 32; kb++) {
        bf16x8 b[NB];
#pragma unroll
        for (int nb = 0; nb < NB; nb++) b[nb] = ld_bf8(wt + (nb * 16 + col) * (K + kPad) + kb * 32 + kq);
#pragma unroll
        for (int nb = 0; nb < NB; nb++) out[nb] = mfma(a[kb], b[nb], out[nb]);
    }
}

__global__ __launch_bounds__(kTF) void k_mlp_forward(FwdArgs a) {
    extern __shared__ uint4 smem_raw[];
    uint16_t* sWT = reinterpret_cast<uint16_t*>(smem_raw);
    uint16_t* sAct = sWT + kNWT;                                  // [8][16][kActStride]
    float* sBias = reinterpret_cast<float*>(sAct + (kTF / 64) * 16 * kActStride);   // b1, b2, b3, b4
    const uint32_t t = threadIdx.x, w = t / 64, l = t % 64;
    for (uint32_t i = t; i < kNWT / 8; i += kTF)
        reinterpret_cast<uint4*>(sWT)[i] = reinterpret_cast<const uint4*>(a.WT)[i];
    for (uint32_t i = t; i < 3 * kHid + kOutP; i += kTF) {
        const uint32_t layer = i / kHid, j = i % kHid;
        const int off = layer == 0 ? OFF_B1 : layer == 1 ? OFF_B2 : layer == 2 ? OFF_B3 : OFF_B4;
        sBias[i] = a.P[off + j];
    }
    __syncthreads();
    uint16_t* act = sAct + w * 16 * kActStride;
    const uint32_t kq = 8 * (l / 16), col = l % 16;
    for (uint32_t tile = blockIdx.x; tile * kTileF < a.M; tile += gridDim.x) {
        const uint32_t r0 = tile * kTileF + w * 16;
        const uint32_t ra = r0 + col;                              // A-operand row of this lane
        bf16x8 ain[4];
        for (int kb = 0; kb < 2; kb++) {
            if (ra < a.M) ain[kb] = ld_bf8(a.X + a.map(ra) * kInP + kb * 32 + kq);
            else ain[kb] = __builtin_bit_cast(bf16x8, make_uint4(0, 0, 0, 0));
        }
        f32x4 acc[8];
        for (int layer = 0; layer < 3; layer++) {
            if (layer == 0) layer_fwd<kInP, 8>(ain, sWT + WT1, sBias, l, acc);
            else layer_fwd<kHid, 8>(ain, sWT + (layer == 1 ? WT2 : WT3), sBias + layer * kHid, l, acc);
            __syncthreads();   // every wave has read its previous activations
            for (int nb = 0; nb < 8; nb++)
                for (int i = 0; i < 4; i++) {
                    const uint32_t rl = 4 * (l / 16) + i, c = nb * 16 + col, row = r0 + rl;
                    const float z = acc[nb][i];
                    const uint16_t hb = f2bf(silu(z));
                    act[rl * kActStride + c] = hb;
                    if (a.Z && row < a.M) {
                        a.Z[((size_t)layer * a.M + row) * kHid + c] = z;
                        a.H[((size_t)layer * a.M + row) * kHid + c] = hb;
                    }
                }
            __syncthreads();
            for (int kb = 0; kb < 4; kb++) ain[kb] = ld_bf8(act + col * kActStride + kb * 32 + kq);
        }
        f32x4 y[1];
        layer_fwd<kHid, 1>(ain, sWT + WT4, sBias + 3 * kHid, l, y);
        for (int i = 0; i < 4; i++) {
            const uint32_t row = r0 + 4 * (l / 16) + i;
            if (row < a.M && col < 4) a.Y[(size_t)row * 4 + col] = y[0][i];
        }
        __syncthreads();
    }
}

// ---------------------------------------------------------------------------
// backward: loss = mean over rows and the 3 outputs of (Y - T)^2
// ---------------------------------------------------------------------------
struct BwdArgs {
    const uint16_t* X;
    RowMap map;            // as in the forward pass (X and T rows)
    uint32_t M;
    const float* Y;
    const float* T;        // [rows][4] targets
    const float* Z;        // [3][M][128]
    const uint16_t* H;     // [3][M][128]
    const uint16_t* WB;    // row-major bf16 weights
    float* part;           // [gridDim.x][kNParam] per-block gradient sums
    double* loss_part;     // [gridDim.x]
    float inv_count;       // 1 / (3 M_total)
};

// dH (16 rows x 128) = dZ (16 x K) W^T with W row-major [128 in][K out] bf16
template <int K>
__device__ inline void layer_bwd(const bf16x8* a, const uint16_t* wb, uint32_t l, f32x4* out) {
    const uint32_t kq = 8 * (l / 16), col = l % 16;
    for (int nb = 0; nb < 8; nb++) {
        f32x4 acc = {0.f, 0.f, 0.f, 0.f};
        for (int kb = 0; kb < K / 32; kb++) acc = mfma(a[kb], ld_bf8(wb + (nb * 16 + col) * K + kb * 32 + kq), acc);
        out[nb] = acc;
    }
}

// Weight gradients dW = A^T dZ on the matrix cores: the tile's rows are the
// reduction dimension.  For a 16 x 16 block (input block ib, output block ob)
// of dW, lane l feeds A[m = l%16][k = 8(l/16)..+7] = input[row k][ib*16 + m]
// and B[k][n = l%16] = dZ[row k][ob*16 + n], both read as 8 consecutive rows
// of one column from the transposed BF16 tiles in LDS, and accumulates
// C[4(l/16) + j][l%16] = dW[ib*16 + 4(l/16) + j][ob*16 + l%16].  Wave w owns
// blocks w, w + 4, ... of every layer; its accumulators stay in registers
// over all of the block's tiles.
constexpr int kTS = kTile + 8;   // row stride of the transposed BF16 tiles (padded)

template <int NI, int NO>
__device__ __forceinline__ void dw_accumulate(const uint16_t* sinT, const uint16_t* sdzT, uint32_t w, uint32_t l,
                                              f32x4* acc) {
    constexpr int NBW = NI * NO / 4;   // blocks per wave
    const uint32_t kq = 8 * (l / 16), col = l % 16;
#pragma unroll
    for (int q = 0; q < NBW; q++) {
        const uint32_t b = w + 4 * q, ib = b / NO, ob = b % NO;
#pragma unroll
        for (int ks = 0; ks < kTile / 32; ks++) {
            const bf16x8 av = ld_bf8(sinT + (ib * 16 + col) * kTS + ks * 32 + kq);
            const bf16x8 bv = ld_bf8(sdzT + (ob * 16 + col) * kTS + ks * 32 + kq);
            acc[q] = mfma(av, bv, acc[q]);
        }
    }
}

template <int NI, int NO>
__device__ __forceinline__ void dw_store(float* gp, int offW, uint32_t w, uint32_t l, const f32x4* acc) {
    constexpr int NBW = NI * NO / 4;
    const uint32_t col = l % 16;
#pragma unroll
    for (int q = 0; q < NBW; q++) {
        const uint32_t b = w + 4 * q, ib = b / NO, ob = b % NO;
#pragma unroll
        for (int jj = 0; jj < 4; jj++) gp[offW + (ib * 16 + 4 * (l / 16) + jj) * (NO * 16) + ob * 16 + col] = acc[q][jj];
    }
}

__global__ __launch_bounds__(kT) __attribute__((amdgpu_waves_per_eu(1, 1))) void k_mlp_backward(BwdArgs a) {
    // LDS: dZ of the tile (fp32 [64][128], bias sums), its bf16 copy row-major
    // (A operand of dH = dZ W^T) and transposed (B operand of dW), and the
    // layer input transposed in bf16 (A operand of dW)
    extern __shared__ uint4 smem_raw[];
    float* sdz = reinterpret_cast<float*>(smem_raw);                          // [64][128]
    uint16_t* sdzb = reinterpret_cast<uint16_t*>(sdz + kTile * kHid);        // [64][kActStride]
    uint16_t* sdzT = sdzb + kTile * kActStride;                               // [128][kTS]
    uint16_t* sinT = sdzT + kHid * kTS;                                       // [128][kTS]
    double* sloss = reinterpret_cast<double*>(sinT + kHid * kTS);
    const uint32_t t = threadIdx.x, w = t / 64, l = t % 64;
    const uint32_t kq = 8 * (l / 16), col = l % 16;
    f32x4 acc1[4 * 8 / 4], acc2[8 * 8 / 4], acc3[8 * 8 / 4], acc4[8 * 1 / 4];
    for (auto& v : acc1) v = f32x4{0.f, 0.f, 0.f, 0.f};
    for (auto& v : acc2) v = f32x4{0.f, 0.f, 0.f, 0.f};
    for (auto& v : acc3) v = f32x4{0.f, 0.f, 0.f, 0.f};
    for (auto& v : acc4) v = f32x4{0.f, 0.f, 0.f, 0.f};
    float bsum[4] = {0.f, 0.f, 0.f, 0.f};   // thread t < 128: bias gradient of output t, layers 1..4
    double loss = 0.0;
    // dZ of the previous layer from this wave's dH rows: fp32, bf16 row-major, bf16 transposed
    auto store_dz = [&](const f32x4* dh, int prev, uint32_t base) {
        for (int nb = 0; nb < 8; nb++)
            for (int i = 0; i < 4; i++) {
                const uint32_t rl = w * 16 + 4 * (l / 16) + i, c = nb * 16 + col, row = base + rl;
                float dz = 0.f;
                if (row < a.M) dz = dh[nb][i] * dsilu(a.Z[((size_t)prev * a.M + row) * kHid + c]);
                const uint16_t b = f2bf(dz);
                sdz[rl * kHid + c] = dz;
                sdzb[rl * kActStride + c] = b;
                sdzT[c * kTS + rl] = b;
            }
    };
    // the input of layer `prev + 1` (H of layer prev, or X), transposed; a
    // thread reads 8 consecutive columns of one row with one 16-byte load,
    // consecutive threads take consecutive rows (so their transposed 2-byte
    // stores land in consecutive LDS words)
    auto load_input = [&](int prev, uint32_t base) {
        const uint32_t nc8 = (prev < 0 ? (uint32_t)kInP : (uint32_t)kHid) / 8;
        for (uint32_t e = t; e < kTile * nc8; e += kT) {
            const uint32_t r = e % kTile, c0 = 8 * (e / kTile), row = base + r;
            uint4 v = make_uint4(0, 0, 0, 0);
            if (row < a.M)
                v = *reinterpret_cast<const uint4*>(prev < 0 ? a.X + a.map(row) * kInP + c0
                                                             : a.H + ((size_t)prev * a.M + row) * kHid + c0);
            const uint32_t w4[4] = {v.x, v.y, v.z, v.w};
#pragma unroll
            for (int q = 0; q < 4; q++) {
                sinT[(c0 + 2 * q) * kTS + r] = (uint16_t)(w4[q] & 0xffffu);
                sinT[(c0 + 2 * q + 1) * kTS + r] = (uint16_t)(w4[q] >> 16);
            }
        }
    };
    // column sums of the fp32 dZ tile in a fixed order: 8 interleaved partial
    // sums (independent chains), combined pairwise
    auto bias_sum = [&](int layer, uint32_t K) {
        if (t < K) {
            float p[8];
#pragma unroll
            for (int q = 0; q < 8; q++) p[q] = sdz[q * kHid + t];
#pragma unroll
            for (uint32_t r = 8; r < (uint32_t)kTile; r += 8)
#pragma unroll
                for (int q = 0; q < 8; q++) p[q] = p[q] + sdz[(r + q) * kHid + t];
            const float s = ((p[0] + p[1]) + (p[2] + p[3])) + ((p[4] + p[5]) + (p[6] + p[7]));
            bsum[layer] = bsum[layer] + s;
        }
    };
    for (uint32_t tile = blockIdx.x; tile * kTile < a.M; tile += gridDim.x) {
        const uint32_t base = tile * kTile;
        // dY (fp32, 64 x 16) -> sdz[:, 0:16]; bf16 copies (row-major padded to 32 columns, transposed)
        for (uint32_t e = t; e < kTile * kOutK; e += kT) {
            const uint32_t r = e / kOutK, c = e % kOutK, row = base + r;
            float d = 0.f;
            if (row < a.M && c < (uint32_t)kOut) {
                const float diff = a.Y[(size_t)row * 4 + c] - a.T[a.map(row) * 4 + c];
                loss += (double)(diff * diff);
                d = 2.f * a.inv_count * diff;
            }
            const uint16_t b = f2bf(d);
            if (c < (uint32_t)kOutP) {
                sdz[r * kHid + c] = d;
                sdzT[c * kTS + r] = b;
            }
            sdzb[r * kActStride + c] = b;
        }
        load_input(2, base);   // H3, the input of W4
        __syncthreads();
        // layer 4 (W4: 128 x 16)
        dw_accumulate<8, 1>(sinT, sdzT, w, l, acc4);
        bias_sum(3, kOutP);
        {
            bf16x8 adz[1];
            adz[0] = ld_bf8(sdzb + (w * 16 + col) * kActStride + kq);
            f32x4 dh[8];
            layer_bwd<kOutK>(adz, a.WB + WB4, l, dh);
            __syncthreads();   // all reads of this layer's LDS tiles are done
            store_dz(dh, 2, base);
            load_input(1, base);
            __syncthreads();
        }
        // layer 3 (W3)
        dw_accumulate<8, 8>(sinT, sdzT, w, l, acc3);
        bias_sum(2, kHid);
        {
            bf16x8 adz[4];
            for (int kb = 0; kb < 4; kb++) adz[kb] = ld_bf8(sdzb + (w * 16 + col) * kActStride + kb * 32 + kq);
            f32x4 dh[8];
            layer_bwd<kHid>(adz, a.WB + WB3, l, dh);
            __syncthreads();
            store_dz(dh, 1, base);
            load_input(0, base);
            __syncthreads();
        }
        // layer 2 (W2)
        dw_accumulate<8, 8>(sinT, sdzT, w, l, acc2);
        bias_sum(1, kHid);
        {
            bf16x8 adz[4];
            for (int kb = 0; kb < 4; kb++) adz[kb] = ld_bf8(sdzb + (w * 16 + col) * kActStride + kb * 32 + kq);
            f32x4 dh[8];
            layer_bwd<kHid>(adz, a.WB + WB2, l, dh);
            __syncthreads();
            store_dz(dh, 0, base);
            load_input(-1, base);   // X
            __syncthreads();
        }
        // layer 1 (W1: 64 x 128)
        dw_accumulate<4, 8>(sinT, sdzT, w, l, acc1);
        bias_sum(0, kHid);
        __syncthreads();   // before the next tile overwrites the LDS tiles
    }
    // every entry of this block's partial is written exactly once
    float* gp = a.part + (size_t)blockIdx.x * kNParam;
    dw_store<4, 8>(gp, OFF_W1, w, l, acc1);
    dw_store<8, 8>(gp, OFF_W2, w, l, acc2);
    dw_store<8, 8>(gp, OFF_W3, w, l, acc3);
    dw_store<8, 1>(gp, OFF_W4, w, l, acc4);
    if (t < (uint32_t)kHid) {
        gp[OFF_B1 + t] = bsum[0];
        gp[OFF_B2 + t] = bsum[1];
        gp[OFF_B3 + t] = bsum[2];
    }
    if (t < (uint32_t)kOutP) gp[OFF_B4 + t] = bsum[3];
    // block loss
    for (int off = 32; off > 0; off >>= 1) loss += __shfl_xor(loss, off, 64);
    if (l == 0) sloss[w] = loss;
    __syncthreads();
    if (t == 0) a.loss_part[blockIdx.x] = (sloss[0] + sloss[1]) + (sloss[2] + sloss[3]);
}

// grad[i] = sum over blocks of part[b][i] in a fixed order: 8 interleaved
// partial sums (b = q, q + 8, ...), combined pairwise
__global__ void k_mlp_reduce(const float* part, uint32_t nparts, float* grad) {
    const uint32_t i = blockIdx.x * blockDim.x + threadIdx.x;
    if (i >= (uint32_t)kNParam) return;
    float p[8] = {0.f, 0.f, 0.f, 0.f, 0.f, 0.f, 0.f, 0.f};
    uint32_t b = 0;
    for (; b + 8 <= nparts; b += 8)
#pragma unroll
        for (int q = 0; q < 8; q++) p[q] = p[q] + part[(size_t)(b + q) * kNParam + i];
    for (int q = 0; b + q < nparts; q++) p[q] = p[q] + part[(size_t)(b + q) * kNParam + i];
    grad[i] = ((p[0] + p[1]) + (p[2] + p[3])) + ((p[4] + p[5]) + (p[6] + p[7]));
}

// Adam (Kingma & Ba 2015) on the FP32 master parameters
__global__ void k_mlp_adam(float* p, float* m, float* v, const float* g, float lr, float b1, float b2, float eps,
                           float c1, float c2) {
    const uint32_t i = blockIdx.x * blockDim.x + threadIdx.x;
    if (i >= (uint32_t)kNParam) return;
    const float gi = g[i];
    const float mi = b1 * m[i] + (1.f - b1) * gi;
    const float vi = b2 * v[i] + (1.f - b2) * gi * gi;
    m[i] = mi;
    v[i] = vi;
    p[i] = p[i] - lr * (mi / c1) / (sqrtf(vi / c2) + eps);
}

// bf16 weight copies (transposed and row-padded for the forward, row-major
// for dH); one thread per element of the largest array
constexpr int kPackThreads = kHid * kSH;
__global__ void k_mlp_pack(const float* P, uint16_t* WT, uint16_t* WB) {
    const uint32_t i = blockIdx.x * blockDim.x + threadIdx.x;
    if (i < (uint32_t)(kHid * kS1)) {                        // WT1[n][k] = W1[k][n]
        const uint32_t n = i / kS1, k = i % kS1;
        WT[WT1 + i] = k < (uint32_t)kInP ? f2bf(P[OFF_W1 + k * kHid + n]) : (uint16_t)0;
    }
    if (i < (uint32_t)(kHid * kSH)) {                        // WT2, WT3
        const uint32_t n = i / kSH, k = i % kSH;
        WT[WT2 + i] = k < (uint32_t)kHid ? f2bf(P[OFF_W2 + k * kHid + n]) : (uint16_t)0;
        WT[WT3 + i] = k < (uint32_t)kHid ? f2bf(P[OFF_W3 + k * kHid + n]) : (uint16_t)0;
    }
    if (i < (uint32_t)(kOutP * kSH)) {                       // WT4[n][k] = W4[k][n]
        const uint32_t n = i / kSH, k = i % kSH;
        WT[WT4 + i] = k < (uint32_t)kHid ? f2bf(P[OFF_W4 + k * kOutP + n]) : (uint16_t)0;
    }
    if (i < (uint32_t)(kHid * kHid)) {
        WB[WB2 + i] = f2bf(P[OFF_W2 + i]);
        WB[WB3 + i] = f2bf(P[OFF_W3 + i]);
    }
    if (i < (uint32_t)(kHid * kOutK)) {                      // WB4[h][o] = W4[h][o], o < 16
        const uint32_t h = i / kOutK, o = i % kOutK;
        WB[WB4 + i] = o < (uint32_t)kOutP ? f2bf(P[OFF_W4 + h * kOutP + o]) : (uint16_t)0;
    }
}

// ---------------------------------------------------------------------------
// features of one vertex (slot s) at the start of a frame (after predict)
// ---------------------------------------------------------------------------
struct FeatArgs {
    VbdArgs a;                 // solver arrays (x = predicted positions of the frame)
    const ColourInfo* colours; // per-colour slot range and ELL bases
    const uint32_t* colour;    // colour id per slot
    const float* corr1;        // previous frame's solver correction x_end - x_pred (3 per slot)
    const float* corr2;        // the one before
    const float* scale;        // 48 feature scales (divide)
    uint16_t* X;               // [V][64] bf16 out
    uint32_t nv;
    float inv_s0;              // 1 / mean rest length
};

// 48 features of one vertex at the start of a frame (after predict); every
// displacement is in units of the mean rest length s0, every force in the
// displacement it causes over one step (h^2 / m).
__global__ __launch_bounds__(kThreads) void k_mlp_features(FeatArgs f) {
    const uint32_t s = blockIdx.x * kThreads + threadIdx.x;
    if (s >= f.nv) return;
    const VbdArgs& a = f.a;
    const F3 xp = ld3(a.x, s), x0 = ld3(a.x0, s), v0 = ld3(a.v, s);
    const uint32_t meta = a.meta[s];
    const uint32_t edeg = meta & 0xffu, bdeg = (meta >> 8) & 0xffu;
    const F3 c1 = ld3(f.corr1, s), c2 = ld3(f.corr2, s);
    const uint32_t colid = f.colour[s];
    const ColourInfo co = f.colours[colid];
    const uint32_t i = s - co.start, n = co.count;
    // 1-ring means over spring neighbours, spring force and stiffness at x_pred
    F3 rc1 = F3{0, 0, 0}, rc2 = F3{0, 0, 0}, rv = F3{0, 0, 0}, rx = F3{0, 0, 0}, fs = F3{0, 0, 0};
    float strain = 0.f, strain_max = 0.f;
    Sym3 S = sym_zero();
    for (uint32_t j = 0; j < edeg; j++) {
        const uint2 e = a.ell_e[co.eoff + j * n + i];
        const F3 xo = ld3(a.x, e.x);
        rc1 = add3(rc1, ld3(f.corr1, e.x));
        rc2 = add3(rc2, ld3(f.corr2, e.x));
        rv = add3(rv, ld3(a.v, e.x));
        rx = add3(rx, xo);
        const F3 d = sub3(xo, xp);
        const float L = __uint_as_float(e.y);
        const float st = (sqrtf(dot3(d, d)) - L) / L;
        strain += st;
        strain_max = fmaxf(strain_max, fabsf(st));
        stretch_term(xp, xo, L, a.ks, S, fs);
    }
    float hb = 0.f;
    F3 fb = F3{0, 0, 0};
    for (uint32_t j = 0; j < bdeg; j++) {
        const uint32_t k = co.boff + j * n + i;
        const float4 q = a.ell_bq[k];
        const PairIds id = a.ell_bi[k];
        const uint32_t o = id.a >> 30;
        F3 p0, p1, p2, p3;
        pair_points(o, xp, ld3(a.x, id.a & 0x3fffffffu), ld3(a.x, id.b), ld3(a.x, id.c), p0, p1, p2, p3);
        bending_term(q, o, p0, p1, p2, p3, a.kb, hb, fb);
    }
    const float invd = edeg ? 1.f / (float)edeg : 0.f;
    const float m = a.mass[s], h = a.h, is0 = f.inv_s0;
    const float h2m = h * h / fmaxf(m, 1e-12f);
    const F3 ext = F3{m * a.g[0] + a.fp->ext_cur[0], m * a.g[1] + a.fp->ext_cur[1], m * a.g[2] + a.fp->ext_cur[2]};
    float fx[kIn];
    int k = 0;
    auto put3 = [&](F3 p) {
        fx[k++] = p.x;
        fx[k++] = p.y;
        fx[k++] = p.z;
    };
    put3(scale3(h * is0, v0));                          // 0-2   velocity x one step
    put3(scale3(is0, c1));                              // 3-5   own correction, last frame
    put3(scale3(is0, c2));                              // 6-8   own correction, two frames back
    put3(scale3(invd * is0, rc1));                      // 9-11  1-ring mean correction, last frame
    put3(scale3(invd * is0, rc2));                      // 12-14 1-ring mean correction, two frames back
    put3(scale3(invd * h * is0, rv));                   // 15-17 1-ring mean velocity x one step
    put3(scale3(is0, sub3(scale3(invd, rx), xp)));      // 18-20 umbrella Laplacian of x_pred
    put3(scale3(h2m * is0, fs));                        // 21-23 spring force at x_pred
    put3(scale3(h2m * is0, fb));                        // 24-26 bending force at x_pred
    put3(scale3(h2m * is0, ext));                       // 27-29 gravity + external force
    fx[k++] = (float)((meta >> 16) & 1u);               // 30    pinned
    for (uint32_t c = 0; c < 8; c++) fx[k++] = (colid == c) ? 1.f : 0.f;   // 31-38 colour one-hot
    fx[k++] = strain * invd;                            // 39    mean spring strain
    fx[k++] = strain_max;                               // 40    max |strain|
    fx[k++] = (float)edeg * 0.125f;                     // 41    spring degree
    fx[k++] = (float)bdeg * 0.0625f;                    // 42    bending degree
    fx[k++] = (S.xx + S.yy + S.zz) * h2m;               // 43    spring stiffness / inertia
    fx[k++] = hb * h2m;                                 // 44    bending stiffness / inertia
    // 45-47 obstacle contacts: sum over the vertex's contacts of the depth
    // of x_pred inside the contact band times the face normal
    F3 pc = F3{0.f, 0.f, 0.f};
    if (a.ccount) {
        const uint32_t nc = a.ccount[s];
        for (uint32_t q = 0; q < nc; q++) {
            const float4 xs = a.ct_xs[q * a.nslots + s];
            const float4 nn = a.ct_n[q * a.nslots + s];
            const float proj = dot3(sub3(xp, F3{xs.x, xs.y, xs.z}), F3{nn.x, nn.y, nn.z});
            if (proj < a.thickness) pc = add3(pc, scale3(a.thickness - proj, F3{nn.x, nn.y, nn.z}));
        }
    }
    put3(scale3(is0, pc));                              // 45-47 contact push
    (void)x0;
    uint16_t* out = f.X + (size_t)s * kInP;
    for (int c = 0; c < kIn; c++) out[c] = f2bf(fx[c] / f.scale[c]);
    for (int c = kIn; c < kInP; c++) out[c] = 0;
}

// One smoothing pass of the predicted correction over the spring 1-ring:
// out[s] = (y[s] + mean over unpinned spring neighbours of y) / 2 (pinned
// neighbours count as zero, their correction is the inertial one).
__global__ __launch_bounds__(kThreads) void k_mlp_smooth(const VbdArgs a, const ColourInfo* colours,
                                                         const uint32_t* colour, const float* y, float* out,
                                                         uint32_t nv) {
    const uint32_t s = blockIdx.x * kThreads + threadIdx.x;
    if (s >= nv) return;
    const uint32_t edeg = a.meta[s] & 0xffu;
    const ColourInfo co = colours[colour[s]];
    const uint32_t i = s - co.start, n = co.count;
    float m0 = 0.f, m1 = 0.f, m2 = 0.f;
    for (uint32_t j = 0; j < edeg; j++) {
        const uint32_t o = a.ell_e[co.eoff + j * n + i].x;
        if ((a.meta[o] >> 16) & 1u) continue;
        m0 += y[(size_t)o * 4];
        m1 += y[(size_t)o * 4 + 1];
        m2 += y[(size_t)o * 4 + 2];
    }
    const float w = edeg ? 0.5f / (float)edeg : 0.f;
    const float self = edeg ? 0.5f : 1.f;
    out[(size_t)s * 4] = self * y[(size_t)s * 4] + w * m0;
    out[(size_t)s * 4 + 1] = self * y[(size_t)s * 4 + 1] + w * m1;
    out[(size_t)s * 4 + 2] = self * y[(size_t)s * 4 + 2] + w * m2;
    out[(size_t)s * 4 + 3] = 0.f;
}

// x = x_pred + alpha s0 y (pinned vertices keep the inertial guess)
__global__ void k_mlp_apply_scaled(float* x, const float* xpred, const float* y, uint32_t nv, float alpha_s0,
                                   const uint32_t* meta) {
    const uint32_t s = blockIdx.x * blockDim.x + threadIdx.x;
    if (s >= nv) return;
    const bool pinned = (meta[s] >> 16) & 1u;
    for (int c = 0; c < 3; c++)
        x[3 * s + c] = pinned ? xpred[3 * s + c] : xpred[3 * s + c] + alpha_s0 * y[(size_t)s * 4 + c];
}

// targets of a frame: (x_end - x_pred) / s0
__global__ void k_mlp_target(const float* x, const float* xpred, float* T, uint32_t nv, float inv_s0) {
    const uint32_t s = blockIdx.x * blockDim.x + threadIdx.x;
    if (s >= nv) return;
    for (int c = 0; c < 3; c++) T[(size_t)s * 4 + c] = (x[3 * s + c] - xpred[3 * s + c]) * inv_s0;
    T[(size_t)s * 4 + 3] = 0.f;
}

// history: corr2 <- corr1, corr1 <- x_end - x_pred
__global__ void k_mlp_history(const float* x, const float* xpred, float* c1, float* c2, uint32_t n3) {
    const uint32_t i = blockIdx.x * blockDim.x + threadIdx.x;
    if (i >= n3) return;
    c2[i] = c1[i];
    c1[i] = x[i] - xpred[i];
}

int fail(int code, const std::string& msg) {
    lats::set_error(msg);
    return code;
}

#define HIPCHK(call)                                                                            \
    do {                                                                                        \
        hipError_t e_ = (call);                                                                 \
        if (e_ != hipSuccess)                                                                   \
            return fail(LATS_ERR_HIP, std::string(#call) + ": " + hipGetErrorString(e_));       \
    } while (0)


// per-block sums of squares of the 48 recorded features (FP64)
__global__ __launch_bounds__(kT) void k_mlp_feat_sumsq(const uint16_t* X, uint64_t rows, double* part) {
    __shared__ double acc[kT / 64][kIn];
    const uint32_t t = threadIdx.x, w = t / 64, l = t % 64;
    double s = 0.0;
    // lane l < 48 of every wave owns feature l; waves stride over rows
    if (l < (uint32_t)kIn)
        for (uint64_t r = (uint64_t)blockIdx.x * (kT / 64) + w; r < rows; r += (uint64_t)gridDim.x * (kT / 64)) {
            const double v = (double)bf2f(X[r * kInP + l]);
            s += v * v;
        }
    if (l < (uint32_t)kIn) acc[w][l] = s;
    __syncthreads();
    if (t < (uint32_t)kIn) part[(size_t)blockIdx.x * kIn + t] = (acc[0][t] + acc[1][t]) + (acc[2][t] + acc[3][t]);
}

// X[r][c] *= mul[c]
__global__ void k_mlp_rescale(uint16_t* X, uint64_t rows, const float* mul) {
    const uint64_t i = (uint64_t)blockIdx.x * blockDim.x + threadIdx.x;
    if (i >= rows * kIn) return;
    const uint64_t r = i / kIn, c = i % kIn;
    X[r * kInP + c] = f2bf(bf2f(X[r * kInP + c]) * mul[c]);
}

constexpr size_t kFwdLds =
    (size_t)kNWT * 2 + (size_t)(kTF / 64) * 16 * kActStride * 2 + (3 * kHid + kOutP) * sizeof(float);
constexpr size_t kBwdLds = (size_t)kTile * kHid * 4 + (size_t)kTile * kActStride * 2 + (size_t)2 * kHid * (kTile + 8) * 2 +
                           4 * sizeof(double);
constexpr uint32_t kGrid = 256;   // blocks of the training kernels (one partial gradient each)

}  // namespace

// ===========================================================================
struct lats_mlp {
    int device = 0;
    hipStream_t stream = nullptr;              // standalone work (training, host checks)
    DBuf<float> P, Mom, Vel, G, part, scale;   // parameters, Adam moments, gradient, partials, feature scales
    DBuf<uint16_t> WT, WB;                     // bf16 weight copies
    DBuf<double> loss_part;
    uint32_t step = 0;                         // Adam step counter
    // bound solver context and its per-vertex work buffers
    lats_vbd* bound = nullptr;
    uint64_t bound_uid = 0;                    // uid of the bound context (pointers can be reused)
    float s0 = 0.01f;
    DBuf<uint16_t> X;
    DBuf<float> Y, Y2, T, xpred, xfull, xhalf, corr1, corr2;
    hipStream_t side = nullptr;                // guess energies, beside the first sweeps
    hipEvent_t ev_applied = nullptr;           // the guesses are in place
    hipEvent_t ev_guess = nullptr;             // the guesses' energies are in host memory
    uint32_t smooth_passes = 0;                // 1-ring smoothing passes of the prediction
    int half_step = 0;                         // retry at half the prediction before falling back
    float stop_rel = 0.f;                      // stopping rule without a target (0: off)
    DBuf<uint32_t> colour;
    // recorded training rows
    DBuf<uint16_t> dsX;
    DBuf<float> dsT;
    uint64_t ds_rows = 0, ds_cap = 0;
    // training step buffers
    DBuf<float> tY, tZ;
    DBuf<uint16_t> tH;
    uint32_t t_cap = 0;
    ~lats_mlp() {
        if (side) (void)hipStreamSynchronize(side);
        if (ev_guess) (void)hipEventDestroy(ev_guess);
        if (ev_applied) (void)hipEventDestroy(ev_applied);
        if (side) (void)hipStreamDestroy(side);
        if (stream) (void)hipStreamDestroy(stream);
    }
};

namespace {

void enqueue_pack(lats_mlp* m, hipStream_t st) {
    hipLaunchKernelGGL(k_mlp_pack, dim3((kPackThreads + 255) / 256), dim3(256), 0, st, m->P.p, m->WT.p, m->WB.p);
}

uint32_t fwd_grid(uint32_t M) { return std::max(1u, std::min((M + kTileF - 1) / kTileF, kGrid)); }
uint32_t bwd_grid(uint32_t M) { return std::max(1u, std::min((M + kTile - 1) / kTile, kGrid)); }

void enqueue_forward(lats_mlp* m, const uint16_t* X, RowMap map, uint32_t M, float* Y, float* Z, uint16_t* H,
                     hipStream_t st) {
    FwdArgs a;
    a.X = X;
    a.map = map;
    a.M = M;
    a.WT = m->WT.p;
    a.P = m->P.p;
    a.Y = Y;
    a.Z = Z;
    a.H = H;
    hipLaunchKernelGGL(k_mlp_forward, dim3(fwd_grid(M)), dim3(kTF), kFwdLds, st, a);
}

// forward (keeping Z, H) + backward + fixed-order reduction into m->G
void enqueue_gradient(lats_mlp* m, const uint16_t* X, const float* T, RowMap map, uint32_t M, float* Y, float* Z,
                      uint16_t* H, hipStream_t st) {
    enqueue_forward(m, X, map, M, Y, Z, H, st);
    BwdArgs b;
    b.X = X;
    b.map = map;
    b.M = M;
    b.Y = Y;
    b.T = T;
    b.Z = Z;
    b.H = H;
    b.WB = m->WB.p;
    b.part = m->part.p;
    b.loss_part = m->loss_part.p;
    b.inv_count = 1.f / (3.f * (float)M);
    const uint32_t g = bwd_grid(M);
    hipLaunchKernelGGL(k_mlp_backward, dim3(g), dim3(kT), kBwdLds, st, b);
    hipLaunchKernelGGL(k_mlp_reduce, dim3((kNParam + 255) / 256), dim3(256), 0, st, m->part.p, g, m->G.p);
}

void enqueue_features(lats_mlp* m, lats_vbd* c) {
    FeatArgs f;
    f.a = c->args(c->x.p);
    f.colours = c->colours_d.p;
    f.colour = m->colour.p;
    f.corr1 = m->corr1.p;
    f.corr2 = m->corr2.p;
    f.scale = m->scale.p;
    f.X = m->X.p;
    f.nv = c->V;
    f.inv_s0 = 1.f / m->s0;
    hipLaunchKernelGGL(k_mlp_features, dim3(nblocks(c->V)), dim3(kThreads), 0, c->stream, f);
}

int ensure_train_buffers(lats_mlp* m, uint32_t M) {
    if (m->t_cap >= M) return LATS_OK;
    HIPCHK(m->tY.alloc((size_t)M * 4));
    HIPCHK(m->tZ.alloc((size_t)3 * M * kHid));
    HIPCHK(m->tH.alloc((size_t)3 * M * kHid));
    m->t_cap = M;
    return LATS_OK;
}

int check_bound(lats_mlp* m, lats_vbd* c) {
    if (!m || !c) return fail(LATS_ERR_INVALID_ARGUMENT, "null argument");
    if (m->bound != c || m->bound_uid != c->uid)
        return fail(LATS_ERR_INVALID_ARGUMENT, "solver context is not bound to this network (lats_mlp_bind)");
    return LATS_OK;
}

std::vector<uint16_t> to_bf16_rows(const float* X, uint32_t M) {
    std::vector<uint16_t> xb((size_t)M * kInP, 0);
    for (uint32_t r = 0; r < M; r++)
        for (int c = 0; c < kIn; c++) xb[(size_t)r * kInP + c] = f2bf(X[(size_t)r * kIn + c]);
    return xb;
}

// The copies are enqueued on dst's stream, ahead of dst's next frame: a
// synchronous hipMemcpy between device buffers may return before the copy
// is done, and the null stream it runs on does not order dst's
// non-blocking stream (its predict could read a half-copied state).
int copy_frame_state(lats_vbd* dst, lats_vbd* src) {
    const size_t nf = (size_t)3 * src->V * 4;
    HIPCHK(hipStreamSynchronize(src->stream));
    HIPCHK(hipMemcpyAsync(dst->x.p, src->x.p, nf, hipMemcpyDeviceToDevice, dst->stream));
    HIPCHK(hipMemcpyAsync(dst->x0.p, src->x0.p, nf, hipMemcpyDeviceToDevice, dst->stream));
    HIPCHK(hipMemcpyAsync(dst->v.p, src->v.p, nf, hipMemcpyDeviceToDevice, dst->stream));
    HIPCHK(hipMemcpyAsync(dst->fp.p, src->fp.p, sizeof(FrameParams), hipMemcpyDeviceToDevice, dst->stream));
    HIPCHK(hipStreamSynchronize(dst->stream));
    return LATS_OK;
}

// One frame of the bound context c: predict; learned guess (use_model) kept
// only if its energy is <= the inertial guess's; at least first_check colour
// sweeps from a kept guess, then sweeps until the energy is <= target
// (checked every check_every); history; update velocity.
// Without a target (target <= 0) and with the stopping rule on
// (m->stop_rel > 0): from a kept guess, sweeps until the energy decrease per
// sweep since the previous check is <= stop_rel * |energy| (the first check,
// after first_check sweeps, compares with the guess's energy; a rise counts
// as no decrease); from the inertial guess, the plain iteration count (the
// plain frame).  No reference solve is involved.
// With ref: the target is the energy ref reaches after the full iteration
// count from the same frame start, and the frame is recorded as training
// rows (features of c, correction of ref's solve).
int warm_frame(lats_mlp* m, lats_vbd* c, lats_vbd* ref, float target, uint32_t max_sweeps, uint32_t first_check,
               uint32_t check_every, int use_model, float* out) {
    if (!check_every) check_every = 1;
    const uint32_t V = c->V;
    int rc;
    if (ref) {
        if (m->ds_rows + V > m->ds_cap) {
            rc = lats_mlp_reserve_rows(m, std::max<uint64_t>(m->ds_rows + V, 2 * m->ds_cap));
            if (rc) return rc;
        }
        rc = copy_frame_state(ref, c);
        if (rc) return rc;
        rc = lats_vbd_frame_begin(ref);
        if (!rc) rc = lats_vbd_iterate(ref, ref->prm.constraint_iter_count);
        if (!rc) rc = lats_vbd_energy_now(ref, &target);
        if (rc) return rc;
    }
    HIPCHK(hipStreamSynchronize(m->stream));
    rc = lats_vbd_frame_begin(c);
    if (rc) return rc;
    HIPCHK(hipMemcpyAsync(m->xpred.p, c->x.p, (size_t)3 * V * 4, hipMemcpyDeviceToDevice, c->stream));
    float e_pred = 0.f, e_warm = 0.f;
    bool kept = false;
    uint32_t spec = 0;   // sweeps already run from the kept learned guess
    if (use_model || ref) enqueue_features(m, c);
    if (ref) {
        HIPCHK(hipMemcpyAsync(m->dsX.p + m->ds_rows * kInP, m->X.p, (size_t)V * kInP * 2, hipMemcpyDeviceToDevice,
                              c->stream));
        hipLaunchKernelGGL(k_mlp_target, dim3(nblocks(V)), dim3(kThreads), 0, c->stream, ref->x.p, m->xpred.p,
                           m->dsT.p + m->ds_rows * 4, V, 1.f / m->s0);
        HIPCHK(hipGetLastError());
    }
    const bool rule = !(target > 0.f) && !ref && m->stop_rel > 0.f && use_model;
    const bool checks = target > 0.f || rule;
    if (use_model) {
        // learned guess in x (and half of it in xhalf), then the energies of
        // the inertial guess, the learned guess and its half with one host
        // round trip
        enqueue_forward(m, m->X.p, RowMap{}, V, m->Y.p, nullptr, nullptr, c->stream);
        float* y = m->Y.p;
        for (uint32_t p = 0; p < m->smooth_passes; p++) {
            float* o = (y == m->Y.p) ? m->Y2.p : m->Y.p;
            hipLaunchKernelGGL(k_mlp_smooth, dim3(nblocks(V)), dim3(kThreads), 0, c->stream, c->args(c->x.p),
                               c->colours_d.p, m->colour.p, y, o, V);
            y = o;
        }
        // the learned guess goes to x (the sweeps start from it) and to
        // xfull (its energy is evaluated beside the sweeps)
        hipLaunchKernelGGL(k_mlp_apply_scaled, dim3(nblocks(V)), dim3(kThreads), 0, c->stream, c->x.p, m->xpred.p, y,
                           V, m->s0, c->meta.p);
        hipLaunchKernelGGL(k_mlp_apply_scaled, dim3(nblocks(V)), dim3(kThreads), 0, c->stream, m->xfull.p,
                           m->xpred.p, y, V, m->s0, c->meta.p);
        if (m->half_step)
            hipLaunchKernelGGL(k_mlp_apply_scaled, dim3(nblocks(V)), dim3(kThreads), 0, c->stream, m->xhalf.p,
                               m->xpred.p, y, V, 0.5f * m->s0, c->meta.p);
        HIPCHK(hipGetLastError());
        // The three energies run on the side stream while the first sweeps
        // from the learned guess run on the context's stream (the guess is
        // kept almost always; a rejected guess discards them below).  The
        // energies read xpred, xfull and xhalf and the frame start, none of
        // which the sweeps write.
        HIPCHK(hipEventRecord(m->ev_applied, c->stream));
        HIPCHK(hipStreamWaitEvent(m->side, m->ev_applied, 0));
        const float* xs[3] = {m->xpred.p, m->xfull.p, m->xhalf.p};
        const uint32_t ne = m->half_step ? 3u : 2u;
        rc = lats::energies_enqueue(c, xs, ne, m->side);
        if (rc) return rc;
        HIPCHK(hipEventRecord(m->ev_guess, m->side));
        spec = std::min(std::max(first_check, 1u), max_sweeps);
        rc = lats_vbd_iterate(c, spec);
        if (rc) return rc;
        HIPCHK(hipEventSynchronize(m->ev_guess));
        float en[3] = {0.f, 0.f, 0.f};
        for (uint32_t k = 0; k < ne; k++) en[k] = c->host_energy[k];
        e_pred = en[0];
        e_warm = en[1];
        kept = e_warm <= e_pred;
        if (!kept) spec = 0;   // restart below from the half-length or the inertial guess
        if (!kept && m->half_step && en[2] <= e_pred) {
            kept = true;
            e_warm = en[2];
            HIPCHK(hipMemcpyAsync(c->x.p, m->xhalf.p, (size_t)3 * V * 4, hipMemcpyDeviceToDevice, c->stream));
        }
        if (!kept)
            HIPCHK(hipMemcpyAsync(c->x.p, m->xpred.p, (size_t)3 * V * 4, hipMemcpyDeviceToDevice, c->stream));
    } else if (checks) {
        rc = lats_vbd_energy_now(c, &e_pred);
        if (rc) return rc;
        e_warm = e_pred;
    }
    float e = kept ? e_warm : e_pred;
    float e_prev = e;             // energy at the previous check (the rule)
    uint32_t s_prev = 0;          // sweeps at the previous check
    uint32_t sweeps = spec;
    if (spec && checks) {
        rc = lats_vbd_energy_now(c, &e);
        if (rc) return rc;
    }
    // From the learned guess the first check comes after first_check sweeps
    // (a few sweeps always smooth the guess's high-frequency error), then
    // every check_every.  From the inertial guess the frame is the plain
    // solve, so its first check comes after the plain iteration count.
    const uint32_t first = kept ? std::max(first_check, 1u) : std::max(c->prm.constraint_iter_count, 1u);
    auto done = [&]() {
        if (!rule) return checks && e <= target && sweeps >= (kept ? first_check : 0u);
        if (!kept) return sweeps >= c->prm.constraint_iter_count;   // the plain frame
        if (sweeps < std::max(first_check, 1u) || sweeps == s_prev) return false;
        return (e_prev - e) / (float)(sweeps - s_prev) <= m->stop_rel * std::fabs(e);
    };
    while (sweeps < max_sweeps && !done()) {
        const uint32_t k = std::min(sweeps == 0 ? first : check_every, max_sweeps - sweeps);
        rc = lats_vbd_iterate(c, k);
        if (rc) return rc;
        if (checks) {
            e_prev = e;
            s_prev = sweeps;
        }
        sweeps += k;
        if (checks) {
            rc = lats_vbd_energy_now(c, &e);
            if (rc) return rc;
        }
    }
    if (!checks) {
        rc = lats_vbd_energy_now(c, &e);
        if (rc) return rc;
    }
    hipLaunchKernelGGL(k_mlp_history, dim3(nblocks(3 * (uint64_t)V)), dim3(kThreads), 0, c->stream, c->x.p,
                       m->xpred.p, m->corr1.p, m->corr2.p, 3 * V);
    rc = lats_vbd_frame_end(c);
    if (rc) return rc;
    HIPCHK(hipGetLastError());
    HIPCHK(hipStreamSynchronize(c->stream));
    if (ref) m->ds_rows += V;
    if (out) {
        out[0] = (float)sweeps;
        out[1] = kept ? 1.f : 0.f;
        out[2] = e_pred;
        out[3] = e_warm;
        out[4] = e;
        out[5] = target;
    }
    return LATS_OK;
}

}  // namespace

extern "C" {

int lats_mlp_create(uint64_t seed, lats_mlp** out) {
    if (!out) return fail(LATS_ERR_INVALID_ARGUMENT, "null argument");
    *out = nullptr;
    int ndev = 0;
    if (hipGetDeviceCount(&ndev) != hipSuccess || ndev == 0) return fail(LATS_ERR_NO_DEVICE, "no HIP device");
    auto m = new lats_mlp();
    (void)hipGetDevice(&m->device);
    // uniform He initialisation from splitmix64(seed); padded rows / columns
    // zero; the output layer starts 100x smaller (prediction ~ inertial guess)
    std::vector<float> p(kNParam, 0.f);
    uint64_t st = seed;
    auto u = [&]() { return (double)(mix64(st++) >> 11) * (1.0 / 9007199254740992.0); };
    auto fill = [&](int off, int rows, int cols, int used_rows, int used_cols, double gain) {
        const double bound = gain * std::sqrt(6.0 / (double)used_rows);
        for (int r = 0; r < rows; r++)
            for (int c = 0; c < cols; c++)
                p[off + r * cols + c] = (r < used_rows && c < used_cols) ? (float)((2.0 * u() - 1.0) * bound) : 0.f;
    };
    fill(OFF_W1, kInP, kHid, kIn, kHid, 1.0);
    fill(OFF_W2, kHid, kHid, kHid, kHid, 1.0);
    fill(OFF_W3, kHid, kHid, kHid, kHid, 1.0);
    fill(OFF_W4, kHid, kOutP, kHid, kOut, 0.01);
    auto bad = [&](hipError_t e) {
        delete m;
        return fail(LATS_ERR_HIP, hipGetErrorString(e));
    };
    hipError_t e;
    if ((e = hipStreamCreateWithFlags(&m->stream, hipStreamNonBlocking)) != hipSuccess) return bad(e);
    if ((e = hipEventCreateWithFlags(&m->ev_guess, hipEventDisableTiming)) != hipSuccess) return bad(e);
    if ((e = hipEventCreateWithFlags(&m->ev_applied, hipEventDisableTiming)) != hipSuccess) return bad(e);
    if ((e = hipStreamCreateWithFlags(&m->side, hipStreamNonBlocking)) != hipSuccess) return bad(e);
    if ((e = m->P.upload(p)) != hipSuccess) return bad(e);
    if ((e = m->Mom.upload(std::vector<float>(kNParam, 0.f))) != hipSuccess) return bad(e);
    if ((e = m->Vel.upload(std::vector<float>(kNParam, 0.f))) != hipSuccess) return bad(e);
    if ((e = m->G.alloc(kNParam)) != hipSuccess) return bad(e);
    if ((e = m->part.alloc((size_t)kGrid * kNParam)) != hipSuccess) return bad(e);
    if ((e = m->loss_part.alloc(kGrid)) != hipSuccess) return bad(e);
    if ((e = m->scale.upload(std::vector<float>(kIn, 1.f))) != hipSuccess) return bad(e);
    if ((e = m->WT.alloc(kNWT)) != hipSuccess) return bad(e);
    if ((e = m->WB.alloc(kNWB)) != hipSuccess) return bad(e);
    if ((e = hipFuncSetAttribute((const void*)k_mlp_forward, hipFuncAttributeMaxDynamicSharedMemorySize,
                                 (int)kFwdLds)) != hipSuccess)
        return bad(e);
    if ((e = hipFuncSetAttribute((const void*)k_mlp_backward, hipFuncAttributeMaxDynamicSharedMemorySize,
                                 (int)kBwdLds)) != hipSuccess)
        return bad(e);
    enqueue_pack(m, m->stream);
    if ((e = hipStreamSynchronize(m->stream)) != hipSuccess) return bad(e);
    *out = m;
    return LATS_OK;
}

int lats_mlp_destroy(lats_mlp* m) {
    if (!m) return LATS_OK;
    (void)hipStreamSynchronize(m->stream);
    delete m;
    return LATS_OK;
}

int lats_mlp_num_params(uint32_t* n) {
    if (!n) return fail(LATS_ERR_INVALID_ARGUMENT, "null argument");
    *n = kNParam;
    return LATS_OK;
}

int lats_mlp_get_params(lats_mlp* m, float* p) {
    if (!m || !p) return fail(LATS_ERR_INVALID_ARGUMENT, "null argument");
    HIPCHK(hipStreamSynchronize(m->stream));
    HIPCHK(hipMemcpy(p, m->P.p, kNParam * sizeof(float), hipMemcpyDeviceToHost));
    return LATS_OK;
}

int lats_mlp_set_params(lats_mlp* m, const float* p) {
    if (!m || !p) return fail(LATS_ERR_INVALID_ARGUMENT, "null argument");
    HIPCHK(hipMemcpyAsync(m->P.p, p, kNParam * sizeof(float), hipMemcpyHostToDevice, m->stream));
    enqueue_pack(m, m->stream);
    HIPCHK(hipStreamSynchronize(m->stream));
    return LATS_OK;
}

int lats_mlp_get_feature_scale(lats_mlp* m, float* scale48) {
    if (!m || !scale48) return fail(LATS_ERR_INVALID_ARGUMENT, "null argument");
    HIPCHK(hipMemcpy(scale48, m->scale.p, kIn * sizeof(float), hipMemcpyDeviceToHost));
    return LATS_OK;
}

int lats_mlp_set_feature_scale(lats_mlp* m, const float* scale48) {
    if (!m || !scale48) return fail(LATS_ERR_INVALID_ARGUMENT, "null argument");
    for (int c = 0; c < kIn; c++)
        if (!(scale48[c] > 0.f) || !std::isfinite(scale48[c]))
            return fail(LATS_ERR_INVALID_ARGUMENT, "feature scales must be finite and positive");
    HIPCHK(hipMemcpyAsync(m->scale.p, scale48, kIn * sizeof(float), hipMemcpyHostToDevice, m->stream));
    HIPCHK(hipStreamSynchronize(m->stream));
    return LATS_OK;
}

int lats_mlp_forward_host(lats_mlp* m, const float* X, uint32_t M, float* Y) {
    if (!m || !X || !Y) return fail(LATS_ERR_INVALID_ARGUMENT, "null argument");
    if (M == 0) return LATS_OK;
    DBuf<uint16_t> dX;
    DBuf<float> dY;
    HIPCHK(dX.upload(to_bf16_rows(X, M)));
    HIPCHK(dY.alloc((size_t)M * 4));
    enqueue_forward(m, dX.p, RowMap{}, M, dY.p, nullptr, nullptr, m->stream);
    HIPCHK(hipGetLastError());
    std::vector<float> y((size_t)M * 4);
    HIPCHK(hipMemcpyAsync(y.data(), dY.p, y.size() * 4, hipMemcpyDeviceToHost, m->stream));
    HIPCHK(hipStreamSynchronize(m->stream));
    for (uint32_t r = 0; r < M; r++)
        for (int c = 0; c < kOut; c++) Y[(size_t)r * kOut + c] = y[(size_t)r * 4 + c];
    return LATS_OK;
}

int lats_mlp_gradient_host(lats_mlp* m, const float* X, const float* T, uint32_t M, float* grad, double* loss) {
    if (!m || !X || !T || !grad) return fail(LATS_ERR_INVALID_ARGUMENT, "null argument");
    if (M == 0) return fail(LATS_ERR_INVALID_ARGUMENT, "no rows");
    std::vector<float> tb((size_t)M * 4, 0.f);
    for (uint32_t r = 0; r < M; r++)
        for (int c = 0; c < kOut; c++) tb[(size_t)r * 4 + c] = T[(size_t)r * kOut + c];
    DBuf<uint16_t> dX;
    DBuf<float> dT;
    HIPCHK(dX.upload(to_bf16_rows(X, M)));
    HIPCHK(dT.upload(tb));
    int rc = ensure_train_buffers(m, M);
    if (rc) return rc;
    enqueue_gradient(m, dX.p, dT.p, RowMap{}, M, m->tY.p, m->tZ.p, m->tH.p, m->stream);
    HIPCHK(hipGetLastError());
    const uint32_t g = bwd_grid(M);
    std::vector<double> lp(g);
    HIPCHK(hipMemcpyAsync(grad, m->G.p, kNParam * sizeof(float), hipMemcpyDeviceToHost, m->stream));
    HIPCHK(hipMemcpyAsync(lp.data(), m->loss_part.p, g * sizeof(double), hipMemcpyDeviceToHost, m->stream));
    HIPCHK(hipStreamSynchronize(m->stream));
    if (loss) {
        double s = 0.0;
        for (double v : lp) s += v;
        *loss = s / (3.0 * M);
    }
    return LATS_OK;
}

int lats_mlp_bind(lats_mlp* m, lats_vbd* c) {
    if (!m || !c) return fail(LATS_ERR_INVALID_ARGUMENT, "null argument");
    HIPCHK(hipStreamSynchronize(c->stream));
    const uint32_t V = c->V;
    HIPCHK(m->X.alloc((size_t)V * kInP));
    HIPCHK(m->Y.alloc((size_t)V * 4));
    HIPCHK(m->Y2.alloc((size_t)V * 4));
    HIPCHK(m->T.alloc((size_t)V * 4));
    HIPCHK(m->xpred.alloc((size_t)3 * V));
    HIPCHK(m->xhalf.alloc((size_t)3 * V));
    HIPCHK(m->xfull.alloc((size_t)3 * V));
    HIPCHK(m->corr1.upload(std::vector<float>((size_t)3 * V, 0.f)));
    HIPCHK(m->corr2.upload(std::vector<float>((size_t)3 * V, 0.f)));
    std::vector<uint32_t> col(V, 0);
    for (uint32_t k = 0; k < c->C; k++)
        for (uint32_t i = 0; i < c->ccount[k]; i++) col[c->cstart[k] + i] = k;
    HIPCHK(m->colour.upload(col));
    // length unit of the features and of the prediction: mean rest length
    std::vector<float> rest(c->E);
    if (c->E) HIPCHK(hipMemcpy(rest.data(), c->rest.p, c->E * sizeof(float), hipMemcpyDeviceToHost));
    double sum = 0.0;
    for (float r : rest) sum += r;
    m->s0 = (c->E && sum > 0.0) ? (float)(sum / c->E) : 0.01f;
    m->bound = c;
    m->bound_uid = c->uid;
    return LATS_OK;
}

int lats_mlp_set_guess_options(lats_mlp* m, uint32_t smooth_passes, int half_step) {
    if (!m) return fail(LATS_ERR_INVALID_ARGUMENT, "null argument");
    if (smooth_passes > 8) return fail(LATS_ERR_INVALID_ARGUMENT, "at most 8 smoothing passes");
    m->smooth_passes = smooth_passes;
    m->half_step = half_step != 0;
    return LATS_OK;
}

int lats_mlp_set_stop_rule(lats_mlp* m, float rel_decrease) {
    if (!m) return fail(LATS_ERR_INVALID_ARGUMENT, "null argument");
    if (!(rel_decrease >= 0.f) || !std::isfinite(rel_decrease))
        return fail(LATS_ERR_INVALID_ARGUMENT, "the stopping tolerance must be finite and >= 0");
    m->stop_rel = rel_decrease;
    return LATS_OK;
}

int lats_mlp_reset_history(lats_mlp* m) {
    if (!m) return fail(LATS_ERR_INVALID_ARGUMENT, "null argument");
    if (m->corr1.p) HIPCHK(hipMemset(m->corr1.p, 0, m->corr1.n * 4));
    if (m->corr2.p) HIPCHK(hipMemset(m->corr2.p, 0, m->corr2.n * 4));
    return LATS_OK;
}

int lats_mlp_reserve_rows(lats_mlp* m, uint64_t rows) {
    if (!m) return fail(LATS_ERR_INVALID_ARGUMENT, "null argument");
    if (rows <= m->ds_cap) return LATS_OK;
    DBuf<uint16_t> nx;
    DBuf<float> nt;
    HIPCHK(nx.alloc(rows * kInP));
    HIPCHK(nt.alloc(rows * 4));
    HIPCHK(hipStreamSynchronize(m->stream));
    if (m->ds_rows) {
        // on the network's stream, finished before the old rows are freed
        HIPCHK(hipMemcpyAsync(nx.p, m->dsX.p, m->ds_rows * kInP * 2, hipMemcpyDeviceToDevice, m->stream));
        HIPCHK(hipMemcpyAsync(nt.p, m->dsT.p, m->ds_rows * 4 * 4, hipMemcpyDeviceToDevice, m->stream));
        HIPCHK(hipStreamSynchronize(m->stream));
    }
    std::swap(m->dsX.p, nx.p);
    std::swap(m->dsX.n, nx.n);
    std::swap(m->dsT.p, nt.p);
    std::swap(m->dsT.n, nt.n);
    m->ds_cap = rows;
    return LATS_OK;
}

int lats_mlp_dataset_clear(lats_mlp* m) {
    if (!m) return fail(LATS_ERR_INVALID_ARGUMENT, "null argument");
    m->ds_rows = 0;
    return LATS_OK;
}

int lats_mlp_record_frame(lats_mlp* m, lats_vbd* c) {
    int rc = check_bound(m, c);
    if (rc) return rc;
    const uint32_t V = c->V;
    if (m->ds_rows + V > m->ds_cap) {
        rc = lats_mlp_reserve_rows(m, std::max<uint64_t>(m->ds_rows + V, 2 * m->ds_cap));
        if (rc) return rc;
    }
    rc = lats_vbd_frame_begin(c);
    if (rc) return rc;
    HIPCHK(hipMemcpyAsync(m->xpred.p, c->x.p, (size_t)3 * V * 4, hipMemcpyDeviceToDevice, c->stream));
    enqueue_features(m, c);
    HIPCHK(hipMemcpyAsync(m->dsX.p + m->ds_rows * kInP, m->X.p, (size_t)V * kInP * 2, hipMemcpyDeviceToDevice,
                          c->stream));
    rc = lats_vbd_iterate(c, c->prm.constraint_iter_count);
    if (rc) return rc;
    hipLaunchKernelGGL(k_mlp_target, dim3(nblocks(V)), dim3(kThreads), 0, c->stream, c->x.p, m->xpred.p,
                       m->dsT.p + m->ds_rows * 4, V, 1.f / m->s0);
    hipLaunchKernelGGL(k_mlp_history, dim3(nblocks(3 * (uint64_t)V)), dim3(kThreads), 0, c->stream, c->x.p,
                       m->xpred.p, m->corr1.p, m->corr2.p, 3 * V);
    rc = lats_vbd_frame_end(c);
    if (rc) return rc;
    HIPCHK(hipGetLastError());
    HIPCHK(hipStreamSynchronize(c->stream));
    m->ds_rows += V;
    return LATS_OK;
}

int lats_mlp_dataset_append(lats_mlp* m, const float* X, const float* T, uint32_t M) {
    if (!m || (M && (!X || !T))) return fail(LATS_ERR_INVALID_ARGUMENT, "null argument");
    if (M == 0) return LATS_OK;
    if (m->ds_rows + M > m->ds_cap) {
        int rc = lats_mlp_reserve_rows(m, std::max<uint64_t>(m->ds_rows + M, 2 * m->ds_cap));
        if (rc) return rc;
    }
    std::vector<float> tb((size_t)M * 4, 0.f);
    for (uint32_t r = 0; r < M; r++)
        for (int c = 0; c < kOut; c++) tb[(size_t)r * 4 + c] = T[(size_t)r * kOut + c];
    const std::vector<uint16_t> xb = to_bf16_rows(X, M);
    HIPCHK(hipMemcpyAsync(m->dsX.p + m->ds_rows * kInP, xb.data(), xb.size() * 2, hipMemcpyHostToDevice, m->stream));
    HIPCHK(hipMemcpyAsync(m->dsT.p + m->ds_rows * 4, tb.data(), tb.size() * 4, hipMemcpyHostToDevice, m->stream));
    HIPCHK(hipStreamSynchronize(m->stream));
    m->ds_rows += M;
    return LATS_OK;
}

int lats_mlp_dataset_rows(lats_mlp* m, uint64_t* rows) {
    if (!m || !rows) return fail(LATS_ERR_INVALID_ARGUMENT, "null argument");
    *rows = m->ds_rows;
    return LATS_OK;
}

int lats_mlp_dataset_read(lats_mlp* m, uint64_t first, uint32_t count, float* X, float* T) {
    if (!m) return fail(LATS_ERR_INVALID_ARGUMENT, "null argument");
    if (first > m->ds_rows || count > m->ds_rows - first) return fail(LATS_ERR_INVALID_ARGUMENT, "rows out of range");
    if (count == 0) return LATS_OK;
    std::vector<uint16_t> xb((size_t)count * kInP);
    std::vector<float> tb((size_t)count * 4);
    HIPCHK(hipStreamSynchronize(m->stream));
    HIPCHK(hipMemcpy(xb.data(), m->dsX.p + first * kInP, xb.size() * 2, hipMemcpyDeviceToHost));
    HIPCHK(hipMemcpy(tb.data(), m->dsT.p + first * 4, tb.size() * 4, hipMemcpyDeviceToHost));
    for (uint32_t r = 0; r < count; r++) {
        if (X)
            for (int c = 0; c < kIn; c++) X[(size_t)r * kIn + c] = bf2f(xb[(size_t)r * kInP + c]);
        if (T)
            for (int c = 0; c < kOut; c++) T[(size_t)r * kOut + c] = tb[(size_t)r * 4 + c];
    }
    return LATS_OK;
}

int lats_mlp_fit_feature_scale(lats_mlp* m, float* scale48) {
    if (!m) return fail(LATS_ERR_INVALID_ARGUMENT, "null argument");
    if (m->ds_rows == 0) return fail(LATS_ERR_INVALID_ARGUMENT, "no recorded rows");
    DBuf<double> part;
    DBuf<float> mul;
    const uint32_t g = (uint32_t)std::min<uint64_t>(1024, (m->ds_rows + 3) / 4);
    HIPCHK(part.alloc((size_t)g * kIn));
    hipLaunchKernelGGL(k_mlp_feat_sumsq, dim3(g), dim3(kT), 0, m->stream, m->dsX.p, m->ds_rows, part.p);
    HIPCHK(hipGetLastError());
    std::vector<double> hp((size_t)g * kIn);
    HIPCHK(hipMemcpyAsync(hp.data(), part.p, hp.size() * 8, hipMemcpyDeviceToHost, m->stream));
    HIPCHK(hipStreamSynchronize(m->stream));
    std::vector<float> sc(kIn), mulv(kIn);
    HIPCHK(hipMemcpy(sc.data(), m->scale.p, kIn * 4, hipMemcpyDeviceToHost));
    for (int c = 0; c < kIn; c++) {
        double s = 0.0;
        for (uint32_t b = 0; b < g; b++) s += hp[(size_t)b * kIn + c];
        const double rms = std::sqrt(s / (double)m->ds_rows);
        // constant-zero features keep their scale
        const double r = (rms > 1e-20 && std::isfinite(rms)) ? rms : 1.0;
        mulv[c] = (float)(1.0 / r);
        sc[c] = (float)(sc[c] * r);
    }
    HIPCHK(mul.upload(mulv));
    const uint64_t n = m->ds_rows * kIn;
    hipLaunchKernelGGL(k_mlp_rescale, dim3((uint32_t)((n + 255) / 256)), dim3(256), 0, m->stream, m->dsX.p,
                       m->ds_rows, mul.p);
    HIPCHK(hipGetLastError());
    HIPCHK(hipMemcpyAsync(m->scale.p, sc.data(), kIn * 4, hipMemcpyHostToDevice, m->stream));
    HIPCHK(hipStreamSynchronize(m->stream));
    if (scale48) std::memcpy(scale48, sc.data(), kIn * 4);
    return LATS_OK;
}

int lats_mlp_time_kernels(lats_mlp* m, uint32_t rows, uint32_t iters, float* forward_ms, float* step_ms) {
    if (!m || !forward_ms || !step_ms) return fail(LATS_ERR_INVALID_ARGUMENT, "null argument");
    if (rows == 0 || iters == 0 || rows > m->ds_rows)
        return fail(LATS_ERR_INVALID_ARGUMENT, "rows must be in [1, recorded rows] and iters positive");
    int rc = ensure_train_buffers(m, rows);
    if (rc) return rc;
    hipEvent_t e0 = nullptr, e1 = nullptr, e2 = nullptr;
    HIPCHK(hipEventCreate(&e0));
    HIPCHK(hipEventCreate(&e1));
    HIPCHK(hipEventCreate(&e2));
    // forward passes over the first `rows` recorded rows (inference: no Z / H)
    enqueue_forward(m, m->dsX.p, RowMap{}, rows, m->tY.p, nullptr, nullptr, m->stream);   // warm-up
    hipError_t e = hipEventRecord(e0, m->stream);
    for (uint32_t k = 0; k < iters && e == hipSuccess; k++)
        enqueue_forward(m, m->dsX.p, RowMap{}, rows, m->tY.p, nullptr, nullptr, m->stream);
    if (e == hipSuccess) e = hipEventRecord(e1, m->stream);
    // whole training steps (forward with Z / H, backward, reduction, Adam, weight packing)
    const float b1 = 0.9f, b2 = 0.999f;
    for (uint32_t k = 0; k < iters && e == hipSuccess; k++) {
        enqueue_gradient(m, m->dsX.p, m->dsT.p, RowMap{}, rows, m->tY.p, m->tZ.p, m->tH.p, m->stream);
        m->step++;
        const float c1 = 1.f - std::pow(b1, (float)m->step), c2 = 1.f - std::pow(b2, (float)m->step);
        hipLaunchKernelGGL(k_mlp_adam, dim3((kNParam + 255) / 256), dim3(256), 0, m->stream, m->P.p, m->Mom.p,
                           m->Vel.p, m->G.p, 1e-4f, b1, b2, 1e-8f, c1, c2);
        enqueue_pack(m, m->stream);
    }
    if (e == hipSuccess) e = hipEventRecord(e2, m->stream);
    if (e == hipSuccess) e = hipStreamSynchronize(m->stream);
    float a = 0.f, b = 0.f;
    if (e == hipSuccess) e = hipEventElapsedTime(&a, e0, e1);
    if (e == hipSuccess) e = hipEventElapsedTime(&b, e1, e2);
    (void)hipEventDestroy(e0);
    (void)hipEventDestroy(e1);
    (void)hipEventDestroy(e2);
    if (e != hipSuccess) return fail(LATS_ERR_HIP, hipGetErrorString(e));
    HIPCHK(hipGetLastError());
    *forward_ms = a / (float)iters;
    *step_ms = b / (float)iters;
    return LATS_OK;
}

int lats_mlp_train(lats_mlp* m, uint32_t steps, uint32_t batch_rows, float lr, uint64_t seed, double* mean_loss) {
    if (!m) return fail(LATS_ERR_INVALID_ARGUMENT, "null argument");
    if (m->ds_rows == 0) return fail(LATS_ERR_INVALID_ARGUMENT, "no recorded rows");
    if (batch_rows == 0) return fail(LATS_ERR_INVALID_ARGUMENT, "batch_rows must be positive");
    int rc = ensure_train_buffers(m, batch_rows);
    if (rc) return rc;
    const uint32_t g = bwd_grid(batch_rows);
    std::vector<double> lp(g);
    double sum = 0.0;
    const float b1 = 0.9f, b2 = 0.999f;
    for (uint32_t it = 0; it < steps; it++) {
        const RowMap map{m->ds_rows, mix64(seed + it) | 1ull};
        enqueue_gradient(m, m->dsX.p, m->dsT.p, map, batch_rows, m->tY.p, m->tZ.p, m->tH.p, m->stream);
        m->step++;
        const float c1 = 1.f - std::pow(b1, (float)m->step), c2 = 1.f - std::pow(b2, (float)m->step);
        hipLaunchKernelGGL(k_mlp_adam, dim3((kNParam + 255) / 256), dim3(256), 0, m->stream, m->P.p, m->Mom.p,
                           m->Vel.p, m->G.p, lr, b1, b2, 1e-8f, c1, c2);
        enqueue_pack(m, m->stream);
        HIPCHK(hipGetLastError());
        HIPCHK(hipMemcpyAsync(lp.data(), m->loss_part.p, g * sizeof(double), hipMemcpyDeviceToHost, m->stream));
        HIPCHK(hipStreamSynchronize(m->stream));
        double s = 0.0;
        for (double v : lp) s += v;
        sum += s / (3.0 * batch_rows);
    }
    if (mean_loss) *mean_loss = steps ? sum / steps : 0.0;
    return LATS_OK;
}

int lats_mlp_warm_start_frame(lats_mlp* m, lats_vbd* c, float target, uint32_t max_sweeps, uint32_t first_check,
                              uint32_t check_every, int use_model, float* out) {
    int rc = check_bound(m, c);
    if (rc) return rc;
    return warm_frame(m, c, nullptr, target, max_sweeps, first_check, check_every, use_model, out);
}

int lats_mlp_record_guided_frame(lats_mlp* m, lats_vbd* c, lats_vbd* ref, uint32_t max_sweeps, uint32_t first_check,
                                 uint32_t check_every, int use_model, float* out) {
    int rc = check_bound(m, c);
    if (rc) return rc;
    if (!ref || ref == c) return fail(LATS_ERR_INVALID_ARGUMENT, "a second context of the same mesh is needed");
    if (ref->V != c->V || ref->C != c->C || ref->perm != c->perm || ref->ccount != c->ccount)
        return fail(LATS_ERR_INVALID_ARGUMENT, "reference context is not built from the same mesh");
    return warm_frame(m, c, ref, 0.f, max_sweeps, first_check, check_every, use_model, out);
}

}  // extern "C"
