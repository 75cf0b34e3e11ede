// mlp.hpp -- learned contention predictor: abort-log counters, transaction features, a
// 64-64-1 MLP with bf16 weights and fp32 accumulation on gfx950 MFMA, its training
// (forward, backward, Adam) and the scheduler's inference + stable routing partition.
//
// The predictor has no counterpart in the reference.  Its input is the STM's abort log
// (failed attempts per transaction; the reference keeps only per-class totals,
// bank/csmv_sh/API.cuh:112-119), and its output only chooses the lane a transfer runs
// on; commit semantics are those of the lane.
//
// Features of a transfer over items x_0..x_{nr-1} (x_0, x_1 written, nr <= 4):
//   f[h1(x) mod 32]      += 1 for a written item, 1/2 for a read-only one
//   f[32 + h2(x) mod 31] += floor(log2(1 + A[h0(x) mod 4096])) / 8
//   f[63]                 = max over the items of floor(log2(1 + A[h0(x) mod 4096])) / 8
// where A holds the per-bucket abort counters folded from the abort log (the last feature
// keeps the most-aborted item apart from the hashed sums).  Every feature is a multiple of
// 1/8 below 32, so the bf16 features are exact.
//
// Network: z = w2 . relu(W1^T x + b1) + b2, p = sigmoid(z), trained with the mean binary
// cross-entropy on "aborted at least once".  One wave owns a tile of 64 samples; the
// products X W1 (forward) and X^T dA (weight gradient) are 4 x 4 blocks of
// v_mfma_f32_16x16x16bf16_1k.  Operand layout of that instruction, lane l = 16 q + r:
//   A: rows r, k = 4q..4q+3;   B: k = 4q..4q+3, column r;   D: rows 4q..4q+3, column r.
// The D layout of a forward block is therefore the B layout of the backward product
// with the tile's samples as K, so dA never leaves registers.
#pragma once

#include "csmv_device.hpp"

namespace csmv {

constexpr int kFeat = 64, kHidden = 64, kFeatItems = 4;
constexpr uint32_t kAbortBuckets = 4096;
// parameter vector: W1[f][h] (f-major), b1[h], w2[h], b2
constexpr uint32_t kParamW1 = 0, kParamB1 = kFeat * kHidden, kParamW2 = kParamB1 + kHidden,
                   kParamB2 = kParamW2 + kHidden, kParamCount = kParamB2 + 1;
// per-wave partial gradient row: the parameters' gradients, then the loss
constexpr uint32_t kPartStride = kParamCount + 1;
constexpr uint32_t kMlpWaves = 4;          // waves (tiles in flight) per workgroup
constexpr uint32_t kMlpGradBlocks = 256;   // workgroups of one gradient pass at most
constexpr uint32_t kXStride = 68;          // bf16 row stride of the LDS copy of a tile
constexpr uint32_t kSaltRw = 0xA5A5A5A5u, kSaltAb = 0x5A5A5A5Au;
constexpr uint32_t kPartItems = 1024;      // items per workgroup of the routing partition

typedef short bf16x4 __attribute__((ext_vector_type(4)));
typedef float f32x4 __attribute__((ext_vector_type(4)));

// Float sums across lanes without shuffles.  Each step adds the partner lane's value to
// the lane's own; the partners are lane ^ 1 and lane ^ 2 (DPP quad_perm), then the mirror
// lane in the other half-row and in the other half of the row (DPP row_half_mirror,
// row_mirror), lane ^ 16 (ds_swizzle SWAP16) and lane ^ 32 (v_permlane32_swap).  After each
// step every lane of a group holds the same sum, so a mirror partner adds the same value
// as the xor partner would: the results are bitwise those of a butterfly over lane ^ m.
template <int Ctrl>
__device__ __forceinline__ float dpp_add(float v)
{
    return v + __int_as_float(__builtin_amdgcn_update_dpp(0, __float_as_int(v), Ctrl, 0xF, 0xF, false));
}
__device__ __forceinline__ float row_sum16(float v)
{
    v = dpp_add<dpp::kQuadSwap1>(v);
    v = dpp_add<dpp::kQuadSwap2>(v);
    v = dpp_add<dpp::kRowHalfMirror>(v);
    return dpp_add<dpp::kRowMirror>(v);
}
__device__ __forceinline__ float add_lane_xor16(float v)
{
    return v + __int_as_float(__builtin_amdgcn_ds_swizzle(__float_as_int(v), 0x401F));
}
__device__ __forceinline__ float add_lane_xor32(float v)
{
    // swapping v's upper half with a copy's lower half leaves [lo, lo] and [hi, hi]
    const auto p = __builtin_amdgcn_permlane32_swap(__float_as_uint(v), __float_as_uint(v), false, false);
    const uint32_t lane = lane_id();
    return v + __uint_as_float(lane < 32 ? p[1] : p[0]);
}
__device__ __forceinline__ float wave_sum_f(float v)
{
    return add_lane_xor32(add_lane_xor16(row_sum16(v)));
}

// 32-bit integer finaliser (xor-shift-multiply, constants of the "lowbias32" hash)
__host__ __device__ __forceinline__ uint32_t feat_mix(uint32_t x)
{
    x ^= x >> 16;
    x *= 0x7FEB352Du;
    x ^= x >> 15;
    x *= 0x846CA68Bu;
    x ^= x >> 16;
    return x;
}

// fp32 -> bf16 bits, round to nearest even (finite inputs)
__device__ __forceinline__ uint16_t to_bf16(float f)
{
    uint32_t u = __float_as_uint(f);
    u += 0x7FFFu + ((u >> 16) & 1u);
    return (uint16_t)(u >> 16);
}

__device__ __forceinline__ float sigmoid(float z) { return 1.f / (1.f + __expf(-z)); }

// floor(log2(1 + a)) for any 32-bit a
__device__ __forceinline__ uint32_t log2p1(uint32_t a)
{
    return 63u - (uint32_t)__builtin_clzll((uint64_t)a + 1ull);
}

// ---- abort log -> counters --------------------------------------------------------------
// A[b] >>= shift (shift >= 32 clears)
__global__ void __launch_bounds__(256) abort_decay_kernel(uint32_t *counters, uint32_t shift)
{
    const uint32_t i = blockIdx.x * blockDim.x + threadIdx.x;
    if (i < kAbortBuckets) counters[i] = shift >= 32u ? 0u : counters[i] >> shift;
}

// A[h0(x)] += failed attempts of every transfer touching x
__global__ void __launch_bounds__(256) abort_fold_kernel(const uint32_t *op, const uint32_t *addr,
                                                         const uint32_t *log_ab, uint64_t T,
                                                         uint32_t *counters)
{
    const uint64_t k = (uint64_t)blockIdx.x * blockDim.x + threadIdx.x;
    if (k >= T) return;
    const uint32_t o = op[k];
    const uint32_t ab = log_ab[k];
    if ((o & 0xFFu) != kKindTransfer || ab == 0u) return;
    const uint32_t nr = min((o >> 16) & 0xFFu, (uint32_t)kFeatItems);
    for (uint32_t j = 0; j < nr; ++j)
        atomicAdd(&counters[feat_mix(addr[(uint64_t)j * T + k]) & (kAbortBuckets - 1u)], ab);
}

// ---- features in MFMA A-operand layout ---------------------------------------------------
// Lane (q, r) holds, for the row it serves, features kk*16 + 4q + i in a[kk][i].
__device__ __forceinline__ void feat_add(float (&a)[4][4], uint32_t f, float v, uint32_t q)
{
    const bool mine = ((f >> 2) & 3u) == q;
#pragma unroll
    for (int kk = 0; kk < 4; ++kk)
#pragma unroll
        for (int i = 0; i < 4; ++i)
            a[kk][i] += (mine && (f >> 4) == (uint32_t)kk && (f & 3u) == (uint32_t)i) ? v : 0.f;
}

__device__ __forceinline__ void txn_features(const uint32_t *op, const uint32_t *addr, uint64_t T,
                                             uint64_t k, bool valid, const uint32_t *counters,
                                             uint32_t q, bf16x4 (&out)[4])
{
    float a[4][4];
#pragma unroll
    for (int kk = 0; kk < 4; ++kk)
#pragma unroll
        for (int i = 0; i < 4; ++i) a[kk][i] = 0.f;
    const uint32_t o = valid ? op[k] : 0xFFu;
    if ((o & 0xFFu) == kKindTransfer) {
        const uint32_t nr = min((o >> 16) & 0xFFu, (uint32_t)kFeatItems);
        uint32_t lgmax = 0;
        for (uint32_t j = 0; j < nr; ++j) {
            const uint32_t x = addr[(uint64_t)j * T + k];
            feat_add(a, feat_mix(x ^ kSaltRw) & 31u, j < 2 ? 1.f : 0.5f, q);
            const uint32_t lg = log2p1(counters[feat_mix(x) & (kAbortBuckets - 1u)]);
            feat_add(a, 32u + feat_mix(x ^ kSaltAb) % 31u, (float)lg * 0.125f, q);
            lgmax = lg > lgmax ? lg : lgmax;
        }
        // the most-aborted item apart from the hashed sums (a hot item among cold ones)
        feat_add(a, 63u, (float)lgmax * 0.125f, q);
    }
#pragma unroll
    for (int kk = 0; kk < 4; ++kk)
#pragma unroll
        for (int i = 0; i < 4; ++i) out[kk][i] = (short)to_bf16(a[kk][i]);
}

// ---- one 64-sample tile through the network -----------------------------------------------
// a[bm][kk]: A operands (rows bm*16 + r).  pre[bm][bn][j] = hidden pre-activation of sample
// bm*16 + 4q + j, unit bn*16 + r.  z[bm][j] = output logit of sample bm*16 + 4q + j
// (identical in the 16 lanes of quad q).
__device__ __forceinline__ void mlp_tile(const bf16x4 (&a)[4][4], const uint16_t *W1T,
                                         const float *params, uint32_t q, uint32_t r,
                                         f32x4 (&pre)[4][4], float (&z)[4][4])
{
#pragma unroll
    for (int bm = 0; bm < 4; ++bm)
#pragma unroll
        for (int bn = 0; bn < 4; ++bn) pre[bm][bn] = f32x4{0.f, 0.f, 0.f, 0.f};
#pragma unroll
    for (int kk = 0; kk < 4; ++kk) {
        bf16x4 b[4];
#pragma unroll
        for (int bn = 0; bn < 4; ++bn)
            b[bn] = *(const bf16x4 *)(W1T + (bn * 16 + r) * kFeat + kk * 16 + 4 * q);
#pragma unroll
        for (int bm = 0; bm < 4; ++bm)
#pragma unroll
            for (int bn = 0; bn < 4; ++bn)
                pre[bm][bn] = __builtin_amdgcn_mfma_f32_16x16x16bf16_1k(a[bm][kk], b[bn],
                                                                         pre[bm][bn], 0, 0, 0);
    }
    float zp[4][4];
#pragma unroll
    for (int bm = 0; bm < 4; ++bm)
#pragma unroll
        for (int j = 0; j < 4; ++j) zp[bm][j] = 0.f;
#pragma unroll
    for (int bn = 0; bn < 4; ++bn) {
        const float b1 = params[kParamB1 + bn * 16 + r];
        const float w2 = params[kParamW2 + bn * 16 + r];
#pragma unroll
        for (int bm = 0; bm < 4; ++bm)
#pragma unroll
            for (int j = 0; j < 4; ++j) {
                const float h = pre[bm][bn][j] + b1;
                pre[bm][bn][j] = h;
                zp[bm][j] += fmaxf(h, 0.f) * w2;
            }
    }
    const float b2 = params[kParamB2];
#pragma unroll
    for (int bm = 0; bm < 4; ++bm)
#pragma unroll
        for (int j = 0; j < 4; ++j) {
            z[bm][j] = row_sum16(zp[bm][j]) + b2;
        }
}

// The lane owning sample s = bm*16 + 4q + j is r = 4 bm + j: each sample has one owner.
__device__ __forceinline__ uint32_t own_sample(uint32_t q, uint32_t r)
{
    return (r >> 2) * 16u + 4u * q + (r & 3u);
}
__device__ __forceinline__ float own_value(const float (&z)[4][4], uint32_t r)
{
    float v = 0.f;
#pragma unroll
    for (int bm = 0; bm < 4; ++bm)
#pragma unroll
        for (int j = 0; j < 4; ++j) v = (r == (uint32_t)(bm * 4 + j)) ? z[bm][j] : v;
    return v;
}

__device__ __forceinline__ void load_x_tile(const uint16_t *X, uint64_t n, uint64_t base,
                                            uint32_t q, uint32_t r, bf16x4 (&a)[4][4])
{
#pragma unroll
    for (int bm = 0; bm < 4; ++bm) {
        const uint64_t row = base + bm * 16 + r;
#pragma unroll
        for (int kk = 0; kk < 4; ++kk)
            a[bm][kk] = row < n ? *(const bf16x4 *)(X + row * kFeat + kk * 16 + 4 * q)
                                : bf16x4{0, 0, 0, 0};
    }
}

// p[i] = sigmoid(z) of rows of X (n x 64 bf16)
__global__ void __launch_bounds__(64 * kMlpWaves) mlp_forward_kernel(const uint16_t *X, uint64_t n,
                                                                    const uint16_t *W1T,
                                                                    const float *params, float *p)
{
    const uint32_t lane = threadIdx.x & 63u, q = lane >> 4, r = lane & 15u;
    const uint64_t base = ((uint64_t)blockIdx.x * kMlpWaves + threadIdx.x / 64u) * 64u;
    if (base >= n) return;
    bf16x4 a[4][4];
    load_x_tile(X, n, base, q, r, a);
    f32x4 pre[4][4];
    float z[4][4];
    mlp_tile(a, W1T, params, q, r, pre, z);
    const uint64_t s = base + own_sample(q, r);
    if (s < n) p[s] = sigmoid(own_value(z, r));
}

// Gradient of the mean binary cross-entropy over the n rows of (X, y).  Each wave walks
// tiles blockIdx*4 + wave, stepping by the grid, and writes its partial gradient (and
// partial loss) to part[wave][kPartStride]; mlp_reduce_kernel sums the partials in a fixed
// order, so the result does not depend on scheduling.
__global__ void __launch_bounds__(64 * kMlpWaves) mlp_grad_kernel(const uint16_t *X, const float *y,
                                                                 uint64_t n, const uint16_t *W1T,
                                                                 const float *params, float *part)
{
    __shared__ uint16_t sX[kMlpWaves][64 * kXStride];
    const uint32_t w = threadIdx.x / 64u;
    const uint32_t lane = threadIdx.x & 63u, q = lane >> 4, r = lane & 15u;
    uint16_t *tx = sX[w];
    f32x4 dW[4][4];
#pragma unroll
    for (int fm = 0; fm < 4; ++fm)
#pragma unroll
        for (int bn = 0; bn < 4; ++bn) dW[fm][bn] = f32x4{0.f, 0.f, 0.f, 0.f};
    float gb1[4] = {0.f, 0.f, 0.f, 0.f}, gw2[4] = {0.f, 0.f, 0.f, 0.f};
    float gb2 = 0.f, loss = 0.f;
    const float inv_n = 1.f / (float)n;
    const uint64_t ntiles = (n + 63) / 64;
    for (uint64_t t = (uint64_t)blockIdx.x * kMlpWaves + w; t < ntiles;
         t += (uint64_t)gridDim.x * kMlpWaves) {
        const uint64_t base = t * 64;
        bf16x4 a[4][4];
        load_x_tile(X, n, base, q, r, a);
#pragma unroll
        for (int bm = 0; bm < 4; ++bm)
#pragma unroll
            for (int kk = 0; kk < 4; ++kk)
                *(bf16x4 *)(tx + (bm * 16 + r) * kXStride + kk * 16 + 4 * q) = a[bm][kk];
        f32x4 pre[4][4];
        float z[4][4];
        mlp_tile(a, W1T, params, q, r, pre, z);
        // dL/dz per sample; loss and db2 counted by the sample's owner lane only
        float dz[4][4];
#pragma unroll
        for (int bm = 0; bm < 4; ++bm)
#pragma unroll
            for (int j = 0; j < 4; ++j) {
                const uint64_t s = base + bm * 16 + 4 * q + j;
                const bool live = s < n;
                const float yy = live ? y[s] : 0.f;
                const float zz = z[bm][j];
                dz[bm][j] = live ? (sigmoid(zz) - yy) * inv_n : 0.f;
                if (live && r == (uint32_t)(bm * 4 + j)) {
                    // -[y log p + (1 - y) log(1 - p)] = max(z, 0) - z y + log(1 + e^-|z|)
                    loss += fmaxf(zz, 0.f) - zz * yy + log1pf(__expf(-fabsf(zz)));
                    gb2 += dz[bm][j];
                }
            }
        // dA = dz w2 relu'(pre), kept in the D layout = B layout of X^T dA
        bf16x4 dA[4][4];
#pragma unroll
        for (int bn = 0; bn < 4; ++bn) {
            const float w2 = params[kParamW2 + bn * 16 + r];
#pragma unroll
            for (int bm = 0; bm < 4; ++bm)
#pragma unroll
                for (int j = 0; j < 4; ++j) {
                    const float h = pre[bm][bn][j];
                    const bool act = h > 0.f;
                    const float da = act ? dz[bm][j] * w2 : 0.f;
                    gw2[bn] += act ? h * dz[bm][j] : 0.f;
                    gb1[bn] += da;
                    dA[bm][bn][j] = (short)to_bf16(da);
                }
        }
        wave_fence();
        // dW1[f][h] += sum_s X[s][f] dA[s][h]: A operand = X^T read back transposed
#pragma unroll
        for (int fm = 0; fm < 4; ++fm)
#pragma unroll
            for (int ks = 0; ks < 4; ++ks) {
                bf16x4 xt;
#pragma unroll
                for (int i = 0; i < 4; ++i)
                    xt[i] = (short)tx[(ks * 16 + 4 * q + i) * kXStride + fm * 16 + r];
#pragma unroll
                for (int bn = 0; bn < 4; ++bn)
                    dW[fm][bn] = __builtin_amdgcn_mfma_f32_16x16x16bf16_1k(xt, dA[ks][bn],
                                                                          dW[fm][bn], 0, 0, 0);
            }
        wave_fence();
    }
    float *out = part + ((size_t)blockIdx.x * kMlpWaves + w) * kPartStride;
#pragma unroll
    for (int fm = 0; fm < 4; ++fm)
#pragma unroll
        for (int bn = 0; bn < 4; ++bn)
#pragma unroll
            for (int j = 0; j < 4; ++j)
                out[kParamW1 + (fm * 16 + 4 * q + j) * kHidden + bn * 16 + r] = dW[fm][bn][j];
#pragma unroll
    for (int bn = 0; bn < 4; ++bn) {
        const float v1 = add_lane_xor32(add_lane_xor16(gb1[bn]));
        const float v2 = add_lane_xor32(add_lane_xor16(gw2[bn]));
        if (q == 0) {
            out[kParamB1 + bn * 16 + r] = v1;
            out[kParamW2 + bn * 16 + r] = v2;
        }
    }
    gb2 = wave_sum_f(gb2);
    loss = wave_sum_f(loss);
    if (lane == 0) {
        out[kParamB2] = gb2;
        out[kParamCount] = loss * inv_n;
    }
}

// grad[i] = sum over the partial rows (fixed order); grad[kParamCount] = loss
__global__ void __launch_bounds__(256) mlp_reduce_kernel(const float *part, uint32_t nparts,
                                                         float *grad)
{
    const uint32_t i = blockIdx.x * blockDim.x + threadIdx.x;
    if (i >= kPartStride) return;
    float s = 0.f;
    for (uint32_t b = 0; b < nparts; ++b) s += part[(size_t)b * kPartStride + i];
    grad[i] = s;
}

// W1T[h][f] = bf16(W1[f][h])
__global__ void __launch_bounds__(256) mlp_sync_kernel(const float *params, uint16_t *W1T)
{
    const uint32_t i = blockIdx.x * blockDim.x + threadIdx.x;
    if (i < (uint32_t)(kFeat * kHidden))
        W1T[(i % kHidden) * kFeat + i / kHidden] = to_bf16(params[kParamW1 + i]);
}

// Adam on the fp32 master parameters (bias corrections corr1 = 1 - beta1^t, corr2 =
// 1 - beta2^t); W1 is re-rounded to bf16 for the MFMA products
__global__ void __launch_bounds__(256) mlp_adam_kernel(float *params, float *m, float *v,
                                                       const float *grad, uint16_t *W1T, float lr,
                                                       float beta1, float beta2, float eps,
                                                       float corr1, float corr2)
{
    const uint32_t i = blockIdx.x * blockDim.x + threadIdx.x;
    if (i >= kParamCount) return;
    const float g = grad[i];
    const float mi = beta1 * m[i] + (1.f - beta1) * g;
    const float vi = beta2 * v[i] + (1.f - beta2) * g * g;
    m[i] = mi;
    v[i] = vi;
    const float wv = params[i] - lr * (mi / corr1) / (sqrtf(vi / corr2) + eps);
    params[i] = wv;
    if (i < (uint32_t)(kFeat * kHidden)) W1T[(i % kHidden) * kFeat + i / kHidden] = to_bf16(wv);
}

// ---- features of listed transactions (training set) -----------------------------------------
// X[i] = features of transaction list[i], y[i] = aborted at least once
__global__ void __launch_bounds__(64 * kMlpWaves) features_kernel(
    const uint32_t *op, const uint32_t *addr, uint64_t T, const uint32_t *log_ab,
    const uint32_t *list, uint64_t n, const uint32_t *counters, uint16_t *X, float *y)
{
    const uint32_t lane = threadIdx.x & 63u, q = lane >> 4, r = lane & 15u;
    const uint64_t base = ((uint64_t)blockIdx.x * kMlpWaves + threadIdx.x / 64u) * 64u;
    if (base >= n) return;
#pragma unroll
    for (int bm = 0; bm < 4; ++bm) {
        const uint64_t i = base + bm * 16 + r;
        const bool valid = i < n;
        const uint64_t k = valid ? list[i] : 0;
        bf16x4 a[4];
        txn_features(op, addr, T, k, valid, counters, q, a);
        if (valid) {
#pragma unroll
            for (int kk = 0; kk < 4; ++kk) *(bf16x4 *)(X + i * kFeat + kk * 16 + 4 * q) = a[kk];
            if (q == 0) y[i] = log_ab[k] ? 1.f : 0.f;
        }
    }
}

// ---- scheduler: inference over the resident stream + routing flags --------------------------
// p[k] = predicted abort probability of transfer k (0 for audits); flag[k] = p[k] > tau
__global__ void __launch_bounds__(64 * kMlpWaves) mlp_infer_kernel(
    const uint32_t *op, const uint32_t *addr, uint64_t T, const uint32_t *counters,
    const uint16_t *W1T, const float *params, float tau, float *p, uint8_t *flag)
{
    const uint32_t lane = threadIdx.x & 63u, q = lane >> 4, r = lane & 15u;
    const uint64_t base = ((uint64_t)blockIdx.x * kMlpWaves + threadIdx.x / 64u) * 64u;
    if (base >= T) return;
    bf16x4 a[4][4];
#pragma unroll
    for (int bm = 0; bm < 4; ++bm) {
        const uint64_t k = base + bm * 16 + r;
        txn_features(op, addr, T, k < T ? k : 0, k < T, counters, q, a[bm]);
    }
    f32x4 pre[4][4];
    float z[4][4];
    mlp_tile(a, W1T, params, q, r, pre, z);
    const uint64_t k = base + own_sample(q, r);
    if (k < T) {
        const bool transfer = (op[k] & 0xFFu) == kKindTransfer;
        const float pk = transfer ? sigmoid(own_value(z, r)) : 0.f;
        p[k] = pk;
        flag[k] = (transfer && pk > tau) ? 1 : 0;
    }
}

// ---- stable partition of the stream by flag ------------------------------------------------
// Workgroup b covers items [b*1024, b*1024 + 1024) as 4 passes of 256; item order is kept.
__global__ void __launch_bounds__(256) part_count_kernel(const uint8_t *flag, uint64_t T,
                                                         uint32_t *bcount)
{
    __shared__ uint32_t wsum[4];
    const uint64_t b0 = (uint64_t)blockIdx.x * kPartItems;
    uint32_t c = 0;
    for (uint32_t i = 0; i < 4; ++i) {
        const uint64_t k = b0 + i * 256 + threadIdx.x;
        c += (k < T && flag[k]) ? 1u : 0u;
    }
    c = wave_sum(c);
    if ((threadIdx.x & 63u) == 0) wsum[threadIdx.x / 64] = c;
    __syncthreads();
    if (threadIdx.x == 0) bcount[blockIdx.x] = wsum[0] + wsum[1] + wsum[2] + wsum[3];
}

// exclusive scan of bcount[0..nb) in place (one workgroup of 1024); total -> *total
__global__ void __launch_bounds__(1024) part_scan_kernel(uint32_t *bcount, uint32_t nb,
                                                         uint64_t *total)
{
    __shared__ uint32_t wsum[16];
    __shared__ uint32_t carry;
    if (threadIdx.x == 0) carry = 0;
    __syncthreads();
    const uint32_t lane = threadIdx.x & 63u, w = threadIdx.x / 64u;
    for (uint32_t c0 = 0; c0 < nb; c0 += 1024) {
        const uint32_t i = c0 + threadIdx.x;
        const uint32_t v = i < nb ? bcount[i] : 0u;
        const uint32_t incl = wave_incl_scan(v);
        if (lane == 63) wsum[w] = incl;
        __syncthreads();
        uint32_t before = carry;
        for (uint32_t j = 0; j < w; ++j) before += wsum[j];
        if (i < nb) bcount[i] = before + incl - v;
        __syncthreads();
        if (threadIdx.x == 1023) carry = before + incl;
        __syncthreads();
    }
    if (threadIdx.x == 0) *total = carry;
}

// hot items -> hot_list[rank among hot], others -> norm_list[rank among others]
__global__ void __launch_bounds__(256) part_scatter_kernel(const uint8_t *flag, uint64_t T,
                                                           const uint32_t *boff,
                                                           uint32_t *hot_list,
                                                           uint32_t *norm_list)
{
    __shared__ uint32_t wsum[4];
    const uint64_t b0 = (uint64_t)blockIdx.x * kPartItems;
    const uint32_t w = threadIdx.x / 64u;
    uint32_t carry = boff[blockIdx.x];   // hot items before this pass
    for (uint32_t i = 0; i < 4; ++i) {
        const uint64_t k = b0 + i * 256 + threadIdx.x;
        const bool hot = k < T && flag[k];
        const uint64_t m = ballot(hot);
        if ((threadIdx.x & 63u) == 0) wsum[w] = (uint32_t)__builtin_popcountll(m);
        __syncthreads();
        uint32_t before = carry;
        for (uint32_t j = 0; j < w; ++j) before += wsum[j];
        const uint32_t hb = before + prefix_count(m);   // hot items before k
        if (k < T) {
            if (hot) hot_list[hb] = (uint32_t)k;
            else norm_list[k - hb] = (uint32_t)k;
        }
        carry += wsum[0] + wsum[1] + wsum[2] + wsum[3];
        __syncthreads();
    }
}

}  // namespace csmv
