// FP32 MLP that predicts a bucket's distinct vertex count from its super k-mer
// statistics, replacing the reference's (disabled) HyperLogLog sizing hook
// (include/HyperLogLog.hpp:29-58, src/Subgraphs_Manager.cpp:98-115).  It only
// chooses table capacities and the bucket order; an undersized table is
// detected exactly and rebuilt, so the prediction never changes a result.
//
// Network: N_FEAT -> 64 -> 64 -> 1, ReLU, trained full-batch with Adam inside
// one 1024-thread workgroup (activations in LDS, per-thread register tiles,
// Adam moments in registers).  Inputs are standardised with the
// training set's mean / std, stored after the weights.
//
// Parameter layout (floats): W1[H][F] b1[H] W2[H][H] b2[H] w3[H] b3, then
// mean[F], inv_std[F].
#include "cf_kernels.hpp"

namespace cfgpu {
namespace {

constexpr int F = N_FEAT;
constexpr int H = MLP_H;
constexpr int OW1 = 0, OB1 = OW1 + H * F, OW2 = OB1 + H, OB2 = OW2 + H * H, OW3 = OB2 + H,
              OB3 = OW3 + H, OMEAN = OB3 + 1, OISTD = OMEAN + F;
constexpr int P_ALL = OISTD + F;
static_assert(OB3 + 1 == MLP_PARAMS, "layout");
constexpr int TRAIN_MAX = 128;
constexpr int TW = 1024;
static_assert(MLP_TRAIN_MAX == TRAIN_MAX && H == 64, "thread tiles below assume 128 samples x 64 units");

__device__ __forceinline__ float urand(uint32_t seed, uint32_t i) {
    uint64_t h = fmix64((uint64_t(seed) << 32) ^ i ^ 0x9E3779B97F4A7C15ull);
    return float(h >> 40) * (1.0f / 16777216.0f);  // [0,1)
}

// Sum over groups of `n` consecutive lanes (n a power of two <= 64).
__device__ __forceinline__ float group_sum(float v, int n) {
    for (int d = n >> 1; d > 0; d >>= 1) v += __shfl_xor(v, d);
    return v;
}

// Full-batch training, one workgroup, every activation in LDS.  Samples past
// n (n < 128) are zero rows with zero loss weight, so every thread tile below
// covers the fixed 128 x 64 shapes.  Thread t owns, for the 128x64 products,
// samples s0 = 2*(t/16) .. s0+1 and units j0 = 4*(t%16) .. j0+3.
__global__ __launch_bounds__(TW) void k_mlp_train(float* __restrict__ gparams,
                                                  const float* __restrict__ feat,
                                                  const uint32_t* __restrict__ idx,
                                                  const float* __restrict__ target, uint32_t n,
                                                  uint32_t epochs, float lr, uint32_t seed,
                                                  float* __restrict__ adam,  // 2*MLP_PARAMS
                                                  float* __restrict__ loss_out) {
    typedef __attribute__((ext_vector_type(4))) float f4;
    __shared__ __attribute__((aligned(16))) float P[P_ALL];
    __shared__ __attribute__((aligned(16))) float Gr[MLP_PARAMS];
    __shared__ __attribute__((aligned(16))) float X[TRAIN_MAX * F];
    __shared__ __attribute__((aligned(16))) float H1[TRAIN_MAX * H];  // h1, then dz1
    __shared__ __attribute__((aligned(16))) float B[TRAIN_MAX * H];   // h2, then dz2
    __shared__ float T[TRAIN_MAX];
    __shared__ float dy[TRAIN_MAX];
    __shared__ float red[TW / 64];
    const int tid = threadIdx.x;
    if (n > TRAIN_MAX) n = TRAIN_MAX;
    const float inv_n = 1.f / float(n > 0 ? n : 1);
    const int s0 = 2 * (tid >> 4), j0 = 4 * (tid & 15);

    // He-uniform initialisation, zero biases, zero Adam moments.
    for (int i = tid; i < MLP_PARAMS; i += TW) {
        float v = 0.f;
        if (i < OB1) v = (2.f * urand(seed, i) - 1.f) * sqrtf(6.f / F);
        else if (i >= OW2 && i < OB2) v = (2.f * urand(seed, i) - 1.f) * sqrtf(6.f / H);
        else if (i >= OW3 && i < OB3) v = (2.f * urand(seed, i) - 1.f) * sqrtf(3.f / H);
        P[i] = v;
    }
    float am[(MLP_PARAMS + TW - 1) / TW], av[(MLP_PARAMS + TW - 1) / TW];
#pragma unroll
    for (int c = 0; c < (MLP_PARAMS + TW - 1) / TW; ++c) am[c] = av[c] = 0.f;
    for (int i = tid; i < TRAIN_MAX * F; i += TW)
        X[i] = uint32_t(i / F) < n ? feat[size_t(idx[i / F]) * F + (i % F)] : 0.f;
    for (int i = tid; i < TRAIN_MAX; i += TW) T[i] = uint32_t(i) < n ? target[i] : 0.f;
    __syncthreads();
    // standardisation over the n real samples
    if (tid < F) {
        float m = 0.f, q = 0.f;
        for (uint32_t s = 0; s < n; ++s) m += X[s * F + tid];
        m *= inv_n;
        for (uint32_t s = 0; s < n; ++s) {
            const float d = X[s * F + tid] - m;
            q += d * d;
        }
        const float sd = sqrtf(q * inv_n);
        P[OMEAN + tid] = m;
        P[OISTD + tid] = sd > 1e-6f ? 1.f / sd : 0.f;
    }
    __syncthreads();
    for (int i = tid; i < TRAIN_MAX * F; i += TW)
        X[i] = uint32_t(i / F) < n ? (X[i] - P[OMEAN + i % F]) * P[OISTD + i % F] : 0.f;
    // Output bias starts at the mean target: the network then fits deviations
    // from the average distinct/instance ratio, and weight decay keeps those
    // small when the training buckets do not support them.
    if (tid < 64) {
        float m = 0.f;
        for (uint32_t s = tid; s < n; s += 64) m += T[s];
        m = group_sum(m, 64);
        if (tid == 0) P[OB3] = m * inv_n;
    }
    __syncthreads();
    // this thread's two input rows stay in registers for every epoch
    float xr[2][F];
#pragma unroll
    for (int a = 0; a < 2; ++a)
#pragma unroll
        for (int f = 0; f < F; ++f) xr[a][f] = X[(s0 + a) * F + f];

    const float b1 = 0.9f, b2 = 0.999f, eps = 1e-8f, wd = 1e-3f;
    float b1t = 1.f, b2t = 1.f;
    for (uint32_t ep = 0; ep < epochs; ++ep) {
        // h1 = relu(W1 x + b1)
        {
            float acc[2][4];
#pragma unroll
            for (int c = 0; c < 4; ++c) acc[0][c] = acc[1][c] = P[OB1 + j0 + c];
#pragma unroll
            for (int c = 0; c < 4; ++c)
#pragma unroll
                for (int f = 0; f < F; ++f) {
                    const float w = P[OW1 + (j0 + c) * F + f];
                    acc[0][c] += w * xr[0][f];
                    acc[1][c] += w * xr[1][f];
                }
#pragma unroll
            for (int a = 0; a < 2; ++a)
                *reinterpret_cast<f4*>(&H1[(s0 + a) * H + j0]) =
                    f4{fmaxf(acc[a][0], 0.f), fmaxf(acc[a][1], 0.f), fmaxf(acc[a][2], 0.f), fmaxf(acc[a][3], 0.f)};
        }
        __syncthreads();
        // h2 = relu(W2 h1 + b2)
        {
            float acc[2][4];
#pragma unroll
            for (int c = 0; c < 4; ++c) acc[0][c] = acc[1][c] = P[OB2 + j0 + c];
            for (int i = 0; i < H; i += 4) {
                const f4 h0 = *reinterpret_cast<const f4*>(&H1[s0 * H + i]);
                const f4 h1v = *reinterpret_cast<const f4*>(&H1[(s0 + 1) * H + i]);
#pragma unroll
                for (int c = 0; c < 4; ++c) {
                    const f4 w = *reinterpret_cast<const f4*>(&P[OW2 + (j0 + c) * H + i]);
                    acc[0][c] += w.x * h0.x + w.y * h0.y + w.z * h0.z + w.w * h0.w;
                    acc[1][c] += w.x * h1v.x + w.y * h1v.y + w.z * h1v.z + w.w * h1v.w;
                }
            }
#pragma unroll
            for (int a = 0; a < 2; ++a)
                *reinterpret_cast<f4*>(&B[(s0 + a) * H + j0]) =
                    f4{fmaxf(acc[a][0], 0.f), fmaxf(acc[a][1], 0.f), fmaxf(acc[a][2], 0.f), fmaxf(acc[a][3], 0.f)};
        }
        __syncthreads();
        // y = w3 . h2 + b3 (8 lanes per sample), dy = 2 (y - t) / n
        float my_loss = 0.f;
        {
            const int sm = tid >> 3, part = tid & 7;
            float y = 0.f;
#pragma unroll
            for (int c = 0; c < 8; ++c) y += P[OW3 + part * 8 + c] * B[sm * H + part * 8 + c];
            y = group_sum(y, 8) + P[OB3];
            const float d = uint32_t(sm) < n ? y - T[sm] : 0.f;
            if (part == 0) {
                dy[sm] = 2.f * d * inv_n;
                my_loss = d * d;
            }
        }
        __syncthreads();
        // dw3: 16 lanes per unit, 8 samples each; db3 by the first wave
        {
            const int j = tid >> 4, part = tid & 15;
            float g = 0.f;
#pragma unroll
            for (int c = 0; c < 8; ++c) {
                const int sm = part * 8 + c;
                g += dy[sm] * B[sm * H + j];
            }
            g = group_sum(g, 16);
            if (part == 0) Gr[OW3 + j] = g;
            if (tid < 64) {
                const float gb = group_sum(dy[tid] + dy[tid + 64], 64);
                if (tid == 0) Gr[OB3] = gb;
            }
        }
        __syncthreads();
        // dz2 in place of h2
#pragma unroll
        for (int a = 0; a < 2; ++a) {
            f4 v = *reinterpret_cast<f4*>(&B[(s0 + a) * H + j0]);
            const float d = dy[s0 + a];
            v.x = v.x > 0.f ? d * P[OW3 + j0 + 0] : 0.f;
            v.y = v.y > 0.f ? d * P[OW3 + j0 + 1] : 0.f;
            v.z = v.z > 0.f ? d * P[OW3 + j0 + 2] : 0.f;
            v.w = v.w > 0.f ? d * P[OW3 + j0 + 3] : 0.f;
            *reinterpret_cast<f4*>(&B[(s0 + a) * H + j0]) = v;
        }
        __syncthreads();
        // dW2[j][i] = sum_s dz2[s][j] h1[s][i]; thread: j = t/16, i = 4*(t%16)..+3
        {
            const int j = tid >> 4, i0 = 4 * (tid & 15);
            f4 g = {0.f, 0.f, 0.f, 0.f};
            float gb = 0.f;
            for (int sm = 0; sm < TRAIN_MAX; ++sm) {
                const float d = B[sm * H + j];
                const f4 h = *reinterpret_cast<const f4*>(&H1[sm * H + i0]);
                g += d * h;
                gb += d;
            }
            *reinterpret_cast<f4*>(&Gr[OW2 + j * H + i0]) = g;
            if ((tid & 15) == 0) Gr[OB2 + j] = gb;
        }
        __syncthreads();
        // dz1 = (W2^T dz2) * [h1 > 0] in place of h1 (each element read only by its owner)
        {
            f4 acc[2] = {{0.f, 0.f, 0.f, 0.f}, {0.f, 0.f, 0.f, 0.f}};
            for (int j = 0; j < H; ++j) {
                const f4 w = *reinterpret_cast<const f4*>(&P[OW2 + j * H + j0]);
                acc[0] += B[s0 * H + j] * w;
                acc[1] += B[(s0 + 1) * H + j] * w;
            }
#pragma unroll
            for (int a = 0; a < 2; ++a) {
                f4 h = *reinterpret_cast<f4*>(&H1[(s0 + a) * H + j0]);
                h.x = h.x > 0.f ? acc[a].x : 0.f;
                h.y = h.y > 0.f ? acc[a].y : 0.f;
                h.z = h.z > 0.f ? acc[a].z : 0.f;
                h.w = h.w > 0.f ? acc[a].w : 0.f;
                *reinterpret_cast<f4*>(&H1[(s0 + a) * H + j0]) = h;
            }
        }
        __syncthreads();
        // dW1[i][f] = sum_s dz1[s][i] x[s][f], db1[i] = sum_s dz1[s][i]:
        // 16 lanes per unit i, 8 samples each.
        {
            const int i = tid >> 4, part = tid & 15;
            float g[F + 1];
#pragma unroll
            for (int f = 0; f <= F; ++f) g[f] = 0.f;
#pragma unroll
            for (int c = 0; c < 8; ++c) {
                const int sm = part * 8 + c;
                const float d = H1[sm * H + i];
#pragma unroll
                for (int f = 0; f < F; ++f) g[f] += d * X[sm * F + f];
                g[F] += d;
            }
#pragma unroll
            for (int f = 0; f <= F; ++f) g[f] = group_sum(g[f], 16);
            if (part == 0) {
#pragma unroll
                for (int f = 0; f < F; ++f) Gr[OW1 + i * F + f] = g[f];
                Gr[OB1 + i] = g[F];
            }
        }
        __syncthreads();
        // Adam with decoupled weight decay on the weight matrices; moments in registers
        b1t *= b1;
        b2t *= b2;
        const float c1 = 1.f / (1.f - b1t), c2 = 1.f / (1.f - b2t);
#pragma unroll
        for (int c = 0; c < (MLP_PARAMS + TW - 1) / TW; ++c) {
            const int i = tid + c * TW;
            if (i < MLP_PARAMS) {
                const float g = Gr[i];
                am[c] = b1 * am[c] + (1.f - b1) * g;
                av[c] = b2 * av[c] + (1.f - b2) * g * g;
                const bool is_w = i < OB1 || (i >= OW2 && i < OB2) || (i >= OW3 && i < OB3);
                P[i] -= lr * ((am[c] * c1) / (sqrtf(av[c] * c2) + eps) + (is_w ? wd * P[i] : 0.f));
            }
        }
        if (ep + 1 == epochs) {
            float v = group_sum(my_loss, 64);
            if ((tid & 63) == 0) red[tid >> 6] = v;
        }
        __syncthreads();
    }
    if (tid == 0 && loss_out) {
        float t = 0.f;
        for (int j = 0; j < TW / 64; ++j) t += red[j];
        *loss_out = t * inv_n;
    }
    for (int i = tid; i < P_ALL; i += TW) gparams[i] = P[i];
    (void)adam;
}

__global__ __launch_bounds__(256) void k_mlp_infer(const float* __restrict__ gparams,
                                                   const float* __restrict__ feat, uint32_t n,
                                                   float* __restrict__ pred) {
    __shared__ float P[P_ALL];
    for (int i = threadIdx.x; i < P_ALL; i += 256) P[i] = gparams[i];
    __syncthreads();
    const uint32_t s = blockIdx.x * 256 + threadIdx.x;
    if (s >= n) return;
    float x[F];
    for (int f = 0; f < F; ++f) x[f] = (feat[size_t(s) * F + f] - P[OMEAN + f]) * P[OISTD + f];
    float h1[H];
    for (int j = 0; j < H; ++j) {
        float acc = P[OB1 + j];
        for (int f = 0; f < F; ++f) acc += P[OW1 + j * F + f] * x[f];
        h1[j] = acc > 0.f ? acc : 0.f;
    }
    float y = P[OB3];
    for (int j = 0; j < H; ++j) {
        float acc = P[OB2 + j];
        for (int i = 0; i < H; ++i) acc += P[OW2 + j * H + i] * h1[i];
        y += P[OW3 + j] * (acc > 0.f ? acc : 0.f);
    }
    pred[s] = y;
}

}  // namespace

int mlp_param_floats() { return P_ALL; }

void launch_mlp_train(float* params, const float* feat, const uint32_t* idx, const float* target,
                      uint32_t n, uint32_t epochs, float lr, uint32_t seed, float* adam,
                      float* loss_out, hipStream_t s) {
    hipLaunchKernelGGL(k_mlp_train, dim3(1), dim3(TW), 0, s, params, feat, idx, target, n, epochs,
                       lr, seed, adam, loss_out);
}

void launch_mlp_infer(const float* params, const float* feat, uint32_t n, float* pred,
                      hipStream_t s) {
    if (!n) return;
    hipLaunchKernelGGL(k_mlp_infer, dim3((n + 255) / 256), dim3(256), 0, s, params, feat, n, pred);
}

}  // namespace cfgpu
