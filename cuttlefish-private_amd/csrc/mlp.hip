// FP32 MLP that predicts a bucket's distinct vertex count from its super k-mer
// statistics, replacing the reference's (disabled) HyperLogLog sizing hook
// (include/HyperLogLog.hpp:29-58, src/Subgraphs_Manager.cpp:98-115).  It only
// chooses table capacities and the bucket order; an undersized table is
// detected exactly and rebuilt, so the prediction never changes a result.
//
// Network: N_FEAT -> 64 -> 64 -> 1, ReLU, trained full-batch with Adam inside
// one workgroup (all activations in LDS).  Inputs are standardised with the
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
constexpr int TW = 256;

__device__ __forceinline__ float urand(uint32_t seed, uint32_t i) {
    uint64_t h = fmix64((uint64_t(seed) << 32) ^ i ^ 0x9E3779B97F4A7C15ull);
    return float(h >> 40) * (1.0f / 16777216.0f);  // [0,1)
}

__global__ __launch_bounds__(TW) void k_mlp_train(float* __restrict__ gparams,
                                                  const float* __restrict__ feat,
                                                  const uint32_t* __restrict__ idx,
                                                  const float* __restrict__ target, uint32_t n,
                                                  uint32_t epochs, float lr, uint32_t seed,
                                                  float* __restrict__ adam,  // 2*MLP_PARAMS
                                                  float* __restrict__ loss_out) {
    __shared__ float P[P_ALL];
    __shared__ float Gr[MLP_PARAMS];
    __shared__ float X[TRAIN_MAX * F];
    __shared__ float H1[TRAIN_MAX * H];
    __shared__ float B[TRAIN_MAX * H];
    __shared__ float dy[TRAIN_MAX];
    __shared__ float red[TW / 64];
    const int tid = threadIdx.x;
    if (n > TRAIN_MAX) n = TRAIN_MAX;

    // He-uniform initialisation, zero biases, zero Adam moments.
    for (int i = tid; i < MLP_PARAMS; i += TW) {
        float v = 0.f;
        if (i < OB1) v = (2.f * urand(seed, i) - 1.f) * sqrtf(6.f / F);
        else if (i >= OW2 && i < OB2) v = (2.f * urand(seed, i) - 1.f) * sqrtf(6.f / H);
        else if (i >= OW3 && i < OB3) v = (2.f * urand(seed, i) - 1.f) * sqrtf(3.f / H);
        P[i] = v;
        adam[i] = 0.f;
        adam[MLP_PARAMS + i] = 0.f;
    }
    for (int i = tid; i < int(n) * F; i += TW) X[i] = feat[size_t(idx[i / F]) * F + (i % F)];
    __syncthreads();
    // standardisation
    if (tid < F) {
        float m = 0.f, q = 0.f;
        for (uint32_t s = 0; s < n; ++s) m += X[s * F + tid];
        m /= float(n > 0 ? n : 1);
        for (uint32_t s = 0; s < n; ++s) {
            const float d = X[s * F + tid] - m;
            q += d * d;
        }
        const float sd = sqrtf(q / float(n > 0 ? n : 1));
        P[OMEAN + tid] = m;
        P[OISTD + tid] = sd > 1e-6f ? 1.f / sd : 0.f;
    }
    __syncthreads();
    for (int i = tid; i < int(n) * F; i += TW) X[i] = (X[i] - P[OMEAN + i % F]) * P[OISTD + i % F];
    __syncthreads();

    // Output bias starts at the mean target: the network then fits deviations
    // from the average distinct/instance ratio, and weight decay keeps those
    // small when the training buckets do not support them.
    if (tid == 0) {
        float m = 0.f;
        for (uint32_t s = 0; s < n; ++s) m += target[s];
        P[OB3] = n ? m / float(n) : 0.f;
    }
    __syncthreads();
    const float b1 = 0.9f, b2 = 0.999f, eps = 1e-8f, wd = 1e-3f;
    float b1t = 1.f, b2t = 1.f;
    const int NH = int(n) * H;
    for (uint32_t ep = 0; ep < epochs; ++ep) {
        // forward, (sample, neuron)-parallel
        for (int q = tid; q < NH; q += TW) {
            const int sm = q / H, j = q % H;
            float acc = P[OB1 + j];
            for (int f = 0; f < F; ++f) acc += P[OW1 + j * F + f] * X[sm * F + f];
            H1[q] = acc > 0.f ? acc : 0.f;
        }
        __syncthreads();
        for (int q = tid; q < NH; q += TW) {
            const int sm = q / H, j = q % H;
            float acc = P[OB2 + j];
            for (int i = 0; i < H; ++i) acc += P[OW2 + j * H + i] * H1[sm * H + i];
            B[q] = acc > 0.f ? acc : 0.f;
        }
        __syncthreads();
        float my_loss = 0.f;
        if (uint32_t(tid) < n) {
            float y = P[OB3];
            for (int j = 0; j < H; ++j) y += P[OW3 + j] * B[tid * H + j];
            const float d = y - target[tid];
            my_loss = d * d;
            dy[tid] = 2.f * d / float(n);
        }
        __syncthreads();
        // dw3, db3 (weight-parallel)
        for (int j = tid; j < H + 1; j += TW) {
            float g = 0.f;
            for (uint32_t sm = 0; sm < n; ++sm) g += dy[sm] * (j < H ? B[sm * H + j] : 1.f);
            Gr[j < H ? OW3 + j : OB3] = g;
        }
        __syncthreads();
        // dh2 in place of h2
        for (int q = tid; q < NH; q += TW) B[q] = B[q] > 0.f ? dy[q / H] * P[OW3 + q % H] : 0.f;
        __syncthreads();
        // dW2, db2
        for (int w = tid; w < H * H + H; w += TW) {
            float g = 0.f;
            if (w < H * H) {
                const int j = w / H, i = w % H;
                for (uint32_t sm = 0; sm < n; ++sm) g += B[sm * H + j] * H1[sm * H + i];
                Gr[OW2 + w] = g;
            } else {
                const int j = w - H * H;
                for (uint32_t sm = 0; sm < n; ++sm) g += B[sm * H + j];
                Gr[OB2 + j] = g;
            }
        }
        __syncthreads();
        // dh1 in place of h1 (each (sample, i) reads only its own h1)
        for (int q = tid; q < NH; q += TW) {
            const int sm = q / H, i = q % H;
            float acc = 0.f;
            for (int j = 0; j < H; ++j) acc += B[sm * H + j] * P[OW2 + j * H + i];
            H1[q] = H1[q] > 0.f ? acc : 0.f;
        }
        __syncthreads();
        // dW1, db1
        for (int w = tid; w < H * F + H; w += TW) {
            float g = 0.f;
            if (w < H * F) {
                const int i = w / F, f = w % F;
                for (uint32_t sm = 0; sm < n; ++sm) g += H1[sm * H + i] * X[sm * F + f];
                Gr[OW1 + w] = g;
            } else {
                const int i = w - H * F;
                for (uint32_t sm = 0; sm < n; ++sm) g += H1[sm * H + i];
                Gr[OB1 + i] = g;
            }
        }
        __syncthreads();
        // Adam with decoupled weight decay on the weight matrices
        b1t *= b1;
        b2t *= b2;
        for (int i = tid; i < MLP_PARAMS; i += TW) {
            const float g = Gr[i];
            const float m = b1 * adam[i] + (1.f - b1) * g;
            const float v = b2 * adam[MLP_PARAMS + i] + (1.f - b2) * g * g;
            adam[i] = m;
            adam[MLP_PARAMS + i] = v;
            const bool is_w = i < OB1 || (i >= OW2 && i < OB2) || (i >= OW3 && i < OB3);
            P[i] -= lr * ((m / (1.f - b1t)) / (sqrtf(v / (1.f - b2t)) + eps) + (is_w ? wd * P[i] : 0.f));
        }
        if (ep + 1 == epochs) {
            float v = my_loss;
            for (int d = 32; d > 0; d >>= 1) v += __shfl_xor(v, d);
            if ((tid & 63) == 0) red[tid >> 6] = v;
        }
        __syncthreads();
    }
    if (tid == 0 && loss_out) {
        float t = 0.f;
        for (int j = 0; j < TW / 64; ++j) t += red[j];
        *loss_out = t / float(n > 0 ? n : 1);
    }
    for (int i = tid; i < P_ALL; i += TW) gparams[i] = P[i];
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
