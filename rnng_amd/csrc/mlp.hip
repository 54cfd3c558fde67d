// Cost model of the Voronoi cell query: a 4 -> 64 -> 64 -> 2 multilayer
// perceptron (ReLU hidden layers, linear output, FP32) with its own forward,
// backward, Adam and inference kernels.  No counterpart in the reference: it
// sizes the candidate and edge buffers of a cell graph build from per-cell
// features, so that one filter pass usually suffices where count-then-fill
// needs two.
//
// Parameters, packed: W1[64][4], b1[64], W2[64][64], b2[64], W3[2][64], b3[2]
// (row-major, out x in).  Loss: mean over batch and outputs of w (y - t)^2
// with w = the under-weight (rnng_mlp_set_under_weight, default MLP_UNDER =
// 4) where y < t: an estimate below the count costs a second filter pass, one
// above it only buffer memory.
#include <algorithm>
#include <atomic>
#include <cmath>
#include <mutex>
#include <vector>

#include "mfma.hpp"
#include "mlp.hpp"

namespace rnng {
namespace {

std::atomic<float> g_under{MLP_UNDER};   // loss weight of an estimate below its target
float under_weight() { return g_under.load(); }

constexpr int OW1 = 0, OB1 = OW1 + MLP_H * MLP_I, OW2 = OB1 + MLP_H, OB2 = OW2 + MLP_H * MLP_H,
              OW3 = OB2 + MLP_H, OB3 = OW3 + MLP_O * MLP_H;
static_assert(OB3 + MLP_O == MLP_NPARAMS, "parameter packing");

// Forward pass, one thread per sample; parameters staged in LDS.  A1/A2 (the
// post-ReLU activations, [B][64]) are stored when non-null for the backward
// pass; without them this is the inference kernel.
__global__ __launch_bounds__(256) void k_mlp_forward(const float* __restrict__ P, const float* __restrict__ X,
                                                     int B, float* __restrict__ Y, float* __restrict__ A1,
                                                     float* __restrict__ A2) {
    __shared__ float w[MLP_NPARAMS];
    for (int i = threadIdx.x; i < MLP_NPARAMS; i += blockDim.x) w[i] = P[i];
    __syncthreads();
    const int s = blockIdx.x * blockDim.x + threadIdx.x;
    if (s >= B) return;
    float x[MLP_I];
#pragma unroll
    for (int i = 0; i < MLP_I; ++i) x[i] = X[(size_t)s * MLP_I + i];
    float h1[MLP_H];
#pragma unroll
    for (int j = 0; j < MLP_H; ++j) {
        float a = w[OB1 + j];
#pragma unroll
        for (int i = 0; i < MLP_I; ++i) a += w[OW1 + j * MLP_I + i] * x[i];
        h1[j] = fmaxf(a, 0.0f);
        if (A1) A1[(size_t)s * MLP_H + j] = h1[j];
    }
    float y[MLP_O];
#pragma unroll
    for (int o = 0; o < MLP_O; ++o) y[o] = w[OB3 + o];
    for (int k = 0; k < MLP_H; ++k) {
        float a = w[OB2 + k];
#pragma unroll
        for (int j = 0; j < MLP_H; ++j) a += w[OW2 + k * MLP_H + j] * h1[j];
        const float h2 = fmaxf(a, 0.0f);
        if (A2) A2[(size_t)s * MLP_H + k] = h2;
#pragma unroll
        for (int o = 0; o < MLP_O; ++o) y[o] += w[OW3 + o * MLP_H + k] * h2;
    }
#pragma unroll
    for (int o = 0; o < MLP_O; ++o) Y[(size_t)s * MLP_O + o] = y[o];
}

// Backward pass of the mean squared error for BS samples per workgroup: the
// per-sample vectors go to LDS, then each thread sums its parameters'
// gradient over the block's samples into part[block][0..MLP_NPARAMS), and
// the block's sum of squared errors into part[block][MLP_NPARAMS].  The
// blocks are summed in a fixed order (k_mlp_sum), so gradients are
// deterministic.
constexpr int BS = 32;
constexpr int NPL = MLP_NPARAMS + 1;

__global__ __launch_bounds__(256) void k_mlp_backward(const float* __restrict__ P, const float* __restrict__ X,
                                                      const float* __restrict__ T, const float* __restrict__ Y,
                                                      const float* __restrict__ A1, const float* __restrict__ A2,
                                                      int B, float under, float* __restrict__ part) {
    __shared__ float xs[BS][MLP_I];
    __shared__ float h1[BS][MLP_H + 1];
    __shared__ float h2[BS][MLP_H + 1];
    __shared__ float dz1[BS][MLP_H + 1];
    __shared__ float dz2[BS][MLP_H + 1];
    __shared__ float dy[BS][MLP_O];
    __shared__ float lsum[BS];
    const int s0 = blockIdx.x * BS;
    const int tid = threadIdx.x;
    const float scale = 2.0f / ((float)B * (float)MLP_O);
    for (int e = tid; e < BS * MLP_H; e += 256) {
        const int s = e / MLP_H, j = e % MLP_H;
        const bool ok = s0 + s < B;
        h1[s][j] = ok ? A1[(size_t)(s0 + s) * MLP_H + j] : 0.0f;
        h2[s][j] = ok ? A2[(size_t)(s0 + s) * MLP_H + j] : 0.0f;
    }
    if (tid < BS) {
        const int s = tid;
        const bool ok = s0 + s < B;
        float l = 0.0f;
#pragma unroll
        for (int i = 0; i < MLP_I; ++i) xs[s][i] = ok ? X[(size_t)(s0 + s) * MLP_I + i] : 0.0f;
#pragma unroll
        for (int o = 0; o < MLP_O; ++o) {
            const float r = ok ? Y[(size_t)(s0 + s) * MLP_O + o] - T[(size_t)(s0 + s) * MLP_O + o] : 0.0f;
            const float wr = r < 0.0f ? under : 1.0f;
            dy[s][o] = scale * wr * r;
            l += wr * r * r;
        }
        lsum[s] = l;
    }
    __syncthreads();
    // dz2[s][k] = (sum_o W3[o][k] dy[s][o]) * [h2 > 0]
    for (int e = tid; e < BS * MLP_H; e += 256) {
        const int s = e / MLP_H, k = e % MLP_H;
        float g = 0.0f;
#pragma unroll
        for (int o = 0; o < MLP_O; ++o) g += P[OW3 + o * MLP_H + k] * dy[s][o];
        dz2[s][k] = h2[s][k] > 0.0f ? g : 0.0f;
    }
    __syncthreads();
    // dz1[s][j] = (sum_k W2[k][j] dz2[s][k]) * [h1 > 0]
    for (int e = tid; e < BS * MLP_H; e += 256) {
        const int s = e / MLP_H, j = e % MLP_H;
        float g = 0.0f;
        for (int k = 0; k < MLP_H; ++k) g += P[OW2 + k * MLP_H + j] * dz2[s][k];
        dz1[s][j] = h1[s][j] > 0.0f ? g : 0.0f;
    }
    __syncthreads();
    float* out = part + (size_t)blockIdx.x * NPL;
    for (int p = tid; p < MLP_NPARAMS; p += 256) {
        float g = 0.0f;
        if (p < OB1) {                       // W1[j][i]
            const int j = (p - OW1) / MLP_I, i = (p - OW1) % MLP_I;
            for (int s = 0; s < BS; ++s) g += dz1[s][j] * xs[s][i];
        } else if (p < OW2) {                // b1[j]
            const int j = p - OB1;
            for (int s = 0; s < BS; ++s) g += dz1[s][j];
        } else if (p < OB2) {                // W2[k][j]
            const int k = (p - OW2) / MLP_H, j = (p - OW2) % MLP_H;
            for (int s = 0; s < BS; ++s) g += dz2[s][k] * h1[s][j];
        } else if (p < OW3) {                // b2[k]
            const int k = p - OB2;
            for (int s = 0; s < BS; ++s) g += dz2[s][k];
        } else if (p < OB3) {                // W3[o][k]
            const int o = (p - OW3) / MLP_H, k = (p - OW3) % MLP_H;
            for (int s = 0; s < BS; ++s) g += dy[s][o] * h2[s][k];
        } else {                             // b3[o]
            const int o = p - OB3;
            for (int s = 0; s < BS; ++s) g += dy[s][o];
        }
        out[p] = g;
    }
    if (tid == 0) {
        float l = 0.0f;
        for (int s = 0; s < BS; ++s) l += lsum[s];
        out[MLP_NPARAMS] = l;
    }
}

// Deterministic two-level sum over blocks: out[g][p] = sum of in[r][p] for
// rows r of group g (rows in order); run once with groups of GR rows, then
// once over the groups.  Column MLP_NPARAMS of each row is the block's loss.
constexpr int GR = 64;

__global__ void k_mlp_sum(const float* __restrict__ in, int nrows, int rows_per_group, float* __restrict__ out) {
    const int p = blockIdx.x * blockDim.x + threadIdx.x;
    const int g = blockIdx.y;
    if (p >= NPL) return;
    const int r0 = g * rows_per_group, r1 = min(nrows, r0 + rows_per_group);
    float acc = 0.0f;
    for (int r = r0; r < r1; ++r) acc += in[(size_t)r * NPL + p];
    out[(size_t)g * NPL + p] = acc;
}

// Adam step t (1-based), bias-corrected.
__global__ void k_adam(float* __restrict__ P, const float* __restrict__ G, float* __restrict__ M,
                       float* __restrict__ V, int t, float lr) {
    const int p = blockIdx.x * blockDim.x + threadIdx.x;
    if (p >= MLP_NPARAMS) return;
    const float b1 = 0.9f, b2 = 0.999f, eps = 1e-8f;
    const float g = G[p];
    const float m = b1 * M[p] + (1.0f - b1) * g;
    const float v = b2 * V[p] + (1.0f - b2) * g * g;
    M[p] = m;
    V[p] = v;
    const float mh = m / (1.0f - powf(b1, (float)t));
    const float vh = v / (1.0f - powf(b2, (float)t));
    P[p] -= lr * mh / (sqrtf(vh) + eps);
}

struct Batch {
    LocalBuf P, X, T, Y, A1, A2, part, part2, G;
    int nblocks = 0, ngroups = 0;
    int setup(const float* params, const float* x, const float* t, int B) {
        nblocks = (B + BS - 1) / BS;
        ngroups = (nblocks + GR - 1) / GR;
        RNNG_TRY(P.reserve(MLP_NPARAMS * sizeof(float)));
        RNNG_TRY(X.reserve((size_t)B * MLP_I * sizeof(float)));
        RNNG_TRY(Y.reserve((size_t)B * MLP_O * sizeof(float)));
        RNNG_TRY(A1.reserve((size_t)B * MLP_H * sizeof(float)));
        RNNG_TRY(A2.reserve((size_t)B * MLP_H * sizeof(float)));
        RNNG_TRY(part.reserve((size_t)nblocks * NPL * sizeof(float)));
        RNNG_TRY(part2.reserve((size_t)ngroups * NPL * sizeof(float)));
        RNNG_TRY(G.reserve(NPL * sizeof(float)));
        RNNG_HIP_TRY(hipMemcpyAsync(P.ptr, params, MLP_NPARAMS * sizeof(float), hipMemcpyHostToDevice, stream()));
        RNNG_HIP_TRY(hipMemcpyAsync(X.ptr, x, (size_t)B * MLP_I * sizeof(float), hipMemcpyHostToDevice, stream()));
        if (t) {
            RNNG_TRY(T.reserve((size_t)B * MLP_O * sizeof(float)));
            RNNG_HIP_TRY(
                hipMemcpyAsync(T.ptr, t, (size_t)B * MLP_O * sizeof(float), hipMemcpyHostToDevice, stream()));
        }
        RNNG_HIP_TRY(hipStreamSynchronize(stream()));
        return RNNG_OK;
    }
    // G[0..MLP_NPARAMS) = gradient, G[MLP_NPARAMS] = sum of squared errors
    int loss_grad(int B) {
        k_mlp_forward<<<blocks_for(B, 256), 256, 0, stream()>>>(P.as<float>(), X.as<float>(), B, Y.as<float>(),
                                                                 A1.as<float>(), A2.as<float>());
        k_mlp_backward<<<nblocks, 256, 0, stream()>>>(P.as<float>(), X.as<float>(), T.as<float>(), Y.as<float>(),
                                                      A1.as<float>(), A2.as<float>(), B, under_weight(),
                                                      part.as<float>());
        k_mlp_sum<<<dim3(blocks_for(NPL, 256), ngroups), 256, 0, stream()>>>(part.as<float>(), nblocks, GR,
                                                                              part2.as<float>());
        k_mlp_sum<<<dim3(blocks_for(NPL, 256), 1), 256, 0, stream()>>>(part2.as<float>(), ngroups, ngroups,
                                                                        G.as<float>());
        RNNG_HIP_TRY(hipGetLastError());
        return RNNG_OK;
    }
    int read(float* grad, float* loss, int B) {
        std::vector<float> h(NPL);
        RNNG_HIP_TRY(hipMemcpyAsync(h.data(), G.ptr, NPL * sizeof(float), hipMemcpyDeviceToHost, stream()));
        RNNG_HIP_TRY(hipStreamSynchronize(stream()));
        if (grad) std::copy(h.begin(), h.begin() + MLP_NPARAMS, grad);
        if (loss) *loss = (float)((double)h[MLP_NPARAMS] / ((double)B * MLP_O));
        return RNNG_OK;
    }
};

int check_batch(const float* params, const float* x, int64_t B) {
    if (!params || !x || B <= 0 || B > (1 << 26)) {
        set_error("mlp: need parameters, inputs and 0 < batch <= 2^26");
        return RNNG_ERR_ARG;
    }
    return ensure_init();
}

std::mutex g_model_mu;
DevBuf g_model;            // device parameters of the cost model in use
bool g_model_set = false;

}  // namespace

int mlp_infer_device(const float* P, const float* X, int B, float* Y) {
    if (B <= 0) return RNNG_OK;
    k_mlp_forward<<<blocks_for(B, 256), 256, 0, stream()>>>(P, X, B, Y, nullptr, nullptr);
    RNNG_HIP_TRY(hipGetLastError());
    return RNNG_OK;
}

const float* cost_model_device() {
    std::lock_guard<std::mutex> lk(g_model_mu);
    return g_model_set ? g_model.as<float>() : nullptr;
}

}  // namespace rnng

using namespace rnng;

extern "C" {

int rnng_mlp_infer(const float* params, const float* x, int64_t batch, float* y) {
    RNNG_TRY(check_batch(params, x, batch));
    if (!y) {
        set_error("mlp_infer: null output");
        return RNNG_ERR_ARG;
    }
    Batch b;
    RNNG_TRY(b.setup(params, x, nullptr, (int)batch));
    RNNG_TRY(mlp_infer_device(b.P.as<float>(), b.X.as<float>(), (int)batch, b.Y.as<float>()));
    RNNG_HIP_TRY(hipMemcpyAsync(y, b.Y.ptr, (size_t)batch * MLP_O * sizeof(float), hipMemcpyDeviceToHost, stream()));
    RNNG_HIP_TRY(hipStreamSynchronize(stream()));
    return RNNG_OK;
}

int rnng_mlp_loss_grad(const float* params, const float* x, const float* target, int64_t batch, float* loss,
                       float* grad) {
    RNNG_TRY(check_batch(params, x, batch));
    if (!target || !loss || !grad) {
        set_error("mlp_loss_grad: null target or output");
        return RNNG_ERR_ARG;
    }
    Batch b;
    RNNG_TRY(b.setup(params, x, target, (int)batch));
    RNNG_TRY(b.loss_grad((int)batch));
    return b.read(grad, loss, (int)batch);
}

int rnng_mlp_train(float* params, const float* x, const float* target, int64_t batch, int32_t iters, float lr,
                   float* final_loss) {
    RNNG_TRY(check_batch(params, x, batch));
    if (!target || iters < 0 || !(lr > 0.0f)) {
        set_error("mlp_train: bad arguments");
        return RNNG_ERR_ARG;
    }
    Batch b;
    RNNG_TRY(b.setup(params, x, target, (int)batch));
    LocalBuf M, V;
    RNNG_TRY(M.reserve(MLP_NPARAMS * sizeof(float)));
    RNNG_TRY(V.reserve(MLP_NPARAMS * sizeof(float)));
    RNNG_TRY(launch_fill_u32(M.as<unsigned>(), MLP_NPARAMS, 0u));
    RNNG_TRY(launch_fill_u32(V.as<unsigned>(), MLP_NPARAMS, 0u));
    for (int t = 1; t <= iters; ++t) {
        RNNG_TRY(b.loss_grad((int)batch));
        k_adam<<<blocks_for(MLP_NPARAMS, 256), 256, 0, stream()>>>(b.P.as<float>(), b.G.as<float>(), M.as<float>(),
                                                                   V.as<float>(), t, lr);
    }
    RNNG_HIP_TRY(hipGetLastError());
    RNNG_TRY(b.loss_grad((int)batch));   // loss of the final parameters
    RNNG_HIP_TRY(hipMemcpyAsync(params, b.P.ptr, MLP_NPARAMS * sizeof(float), hipMemcpyDeviceToHost, stream()));
    return b.read(nullptr, final_loss, (int)batch);
}

int rnng_mlp_set_under_weight(float w) {
    if (!(w > 0.0f) || !(w < 1e6f)) {
        set_error("mlp_set_under_weight: need 0 < w < 1e6");
        return RNNG_ERR_ARG;
    }
    g_under.store(w);
    return RNNG_OK;
}

float rnng_mlp_under_weight(void) { return g_under.load(); }

int rnng_set_cost_model(const float* params) {
    RNNG_TRY(ensure_init());
    std::lock_guard<std::mutex> lk(g_model_mu);
    if (!params) {
        g_model_set = false;
        return RNNG_OK;
    }
    RNNG_TRY(g_model.reserve(MLP_NPARAMS * sizeof(float)));
    RNNG_HIP_TRY(hipMemcpyAsync(g_model.ptr, params, MLP_NPARAMS * sizeof(float), hipMemcpyHostToDevice, stream()));
    RNNG_HIP_TRY(hipStreamSynchronize(stream()));
    g_model_set = true;
    return RNNG_OK;
}

}  // extern "C"
