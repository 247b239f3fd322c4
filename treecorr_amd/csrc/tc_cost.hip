// treecorr_amd: the learned cost model of the traversal's jobs.
//
// A persistent traversal hands its jobs (node pairs, each the root of a subtree of
// visits) to workgroups.  How long a job takes varies by orders of magnitude, so a
// static partition with equal job counts leaves most workgroups idle while a few finish
// the dense parts.  The cost model predicts each job's work, y = log2(1 + visits), from
// its two nodes: a perceptron of three FP32 layers (CM_NF inputs, two hidden layers of
// CM_H ReLU units, one linear output) trained on the traversal's own logs.  The
// traversal then gives every workgroup an equal share of predicted work, or takes the
// jobs in order of decreasing predicted work; its results do not depend on either.
//
// Kernels: inference/forward (one sample per thread, parameters in LDS), backward (one
// sample per thread: the two layers' error terms), the gradient reduction (one
// parameter per thread, a fixed-order sum over the batch, so training is reproducible)
// and Adam.  The loss is the mean squared error on y.
#include "tc_engine.h"

#include <cmath>
#include <random>
#include <vector>

namespace tc {

namespace {

constexpr int NF = CM_NF, H = CM_H;
constexpr int O_W1 = 0, O_B1 = O_W1 + H * NF, O_W2 = O_B1 + H, O_B2 = O_W2 + H * H,
              O_W3 = O_B2 + H, O_B3 = O_W3 + H, NP = O_B3 + 1;
static_assert(NP == CM_NP, "parameter layout");

// Forward pass of samples idx[s] (or s when idx is null) of X (n x NF, row major):
// yhat[s], and with act the hidden activations h1, h2 (H each) of sample s.
__global__ __launch_bounds__(256) void k_mlp_forward(const float* __restrict__ params,
                                                     const float* __restrict__ X,
                                                     const long long* __restrict__ idx,
                                                     long long n, float* __restrict__ yhat,
                                                     float* __restrict__ act) {
    __shared__ float P[NP];
    for (int t = threadIdx.x; t < NP; t += blockDim.x) P[t] = params[t];
    __syncthreads();
    const long long stride = (long long)gridDim.x * blockDim.x;
    for (long long s = (long long)blockIdx.x * blockDim.x + threadIdx.x; s < n; s += stride) {
        const long long r = idx ? idx[s] : s;
        float x[NF];
#pragma unroll
        for (int f = 0; f < NF; ++f) x[f] = X[r * NF + f];
        float h1[H];
#pragma unroll
        for (int o = 0; o < H; ++o) {
            float a = P[O_B1 + o];
#pragma unroll
            for (int f = 0; f < NF; ++f) a = fmaf(P[O_W1 + o * NF + f], x[f], a);
            h1[o] = a > 0.f ? a : 0.f;
        }
        float y = P[O_B3];
        for (int o = 0; o < H; ++o) {
            float a = P[O_B2 + o];
#pragma unroll
            for (int k = 0; k < H; ++k) a = fmaf(P[O_W2 + o * H + k], h1[k], a);
            a = a > 0.f ? a : 0.f;
            if (act) act[s * 2 * H + H + o] = a;
            y = fmaf(P[O_W3 + o], a, y);
        }
        if (act)
#pragma unroll
            for (int k = 0; k < H; ++k) act[s * 2 * H + k] = h1[k];
        yhat[s] = y;
    }
}

// Backward pass, one sample per thread: with dy = 2 (yhat - y) / n, the error terms
// d2 = dy W3 [h2 > 0] and d1 = (W2^T d2) [h1 > 0], stored in del (2H per sample).
__global__ __launch_bounds__(256) void k_mlp_backward(const float* __restrict__ params,
                                                      const float* __restrict__ act,
                                                      const float* __restrict__ yhat,
                                                      const float* __restrict__ Y,
                                                      const long long* __restrict__ idx,
                                                      long long n, float* __restrict__ dyv,
                                                      float* __restrict__ del) {
    __shared__ float P[NP];
    for (int t = threadIdx.x; t < NP; t += blockDim.x) P[t] = params[t];
    __syncthreads();
    const long long s = (long long)blockIdx.x * blockDim.x + threadIdx.x;
    if (s >= n) return;
    const long long r = idx ? idx[s] : s;
    const float dy = 2.f * (yhat[s] - Y[r]) / (float)n;
    dyv[s] = dy;
    float d2[H];
#pragma unroll
    for (int k = 0; k < H; ++k) {
        d2[k] = act[s * 2 * H + H + k] > 0.f ? dy * P[O_W3 + k] : 0.f;
        del[s * 2 * H + H + k] = d2[k];
    }
    for (int k = 0; k < H; ++k) {
        float a = 0.f;
#pragma unroll
        for (int o = 0; o < H; ++o) a = fmaf(P[O_W2 + o * H + k], d2[o], a);
        del[s * 2 * H + k] = act[s * 2 * H + k] > 0.f ? a : 0.f;
    }
}

// Gradient of every parameter, one per thread, summed over the batch in sample order.
__global__ __launch_bounds__(256) void k_mlp_grad(const float* __restrict__ X,
                                                  const long long* __restrict__ idx,
                                                  const float* __restrict__ act,
                                                  const float* __restrict__ dyv,
                                                  const float* __restrict__ del, long long n,
                                                  float* __restrict__ grad) {
    const int p = blockIdx.x * blockDim.x + threadIdx.x;
    if (p >= NP) return;
    float g = 0.f;
    if (p < O_B1) {                       // W1[o][f] = sum d1[o] x[f]
        const int o = p / NF, f = p % NF;
        for (long long s = 0; s < n; ++s)
            g = fmaf(del[s * 2 * H + o], X[(idx ? idx[s] : s) * NF + f], g);
    } else if (p < O_W2) {                // b1[o]
        const int o = p - O_B1;
        for (long long s = 0; s < n; ++s) g += del[s * 2 * H + o];
    } else if (p < O_B2) {                // W2[o][k] = sum d2[o] h1[k]
        const int o = (p - O_W2) / H, k = (p - O_W2) % H;
        for (long long s = 0; s < n; ++s)
            g = fmaf(del[s * 2 * H + H + o], act[s * 2 * H + k], g);
    } else if (p < O_W3) {                // b2[o]
        const int o = p - O_B2;
        for (long long s = 0; s < n; ++s) g += del[s * 2 * H + H + o];
    } else if (p < O_B3) {                // W3[k] = sum dy h2[k]
        const int k = p - O_W3;
        for (long long s = 0; s < n; ++s) g = fmaf(dyv[s], act[s * 2 * H + H + k], g);
    } else {                              // b3
        for (long long s = 0; s < n; ++s) g += dyv[s];
    }
    grad[p] = g;
}

__global__ void k_adam(float* __restrict__ params, const float* __restrict__ grad,
                       float* __restrict__ m, float* __restrict__ v, float lr, float c1,
                       float c2) {
    const int p = blockIdx.x * blockDim.x + threadIdx.x;
    if (p >= NP) return;
    const float b1 = 0.9f, b2 = 0.999f, eps = 1e-8f;
    const float g = grad[p];
    const float mm = b1 * m[p] + (1.f - b1) * g;
    const float vv = b2 * v[p] + (1.f - b2) * g * g;
    m[p] = mm;
    v[p] = vv;
    params[p] -= lr * (mm / c1) / (sqrtf(vv / c2) + eps);
}

// Mini-batch indices: a counter-based hash of (seed, step, s) modulo n.
__global__ void k_batch_idx(unsigned long long seed, long long step, long long n, int batch,
                            long long* __restrict__ idx) {
    const int s = blockIdx.x * blockDim.x + threadIdx.x;
    if (s >= batch) return;
    unsigned long long z = seed + 0x9e3779b97f4a7c15ull * (unsigned long long)(step * batch + s + 1);
    z = (z ^ (z >> 30)) * 0xbf58476d1ce4e5b9ull;
    z = (z ^ (z >> 27)) * 0x94d049bb133111ebull;
    z ^= z >> 31;
    idx[s] = (long long)(z % (unsigned long long)n);
}

// Squared errors of a forward pass, summed per workgroup (then on the host, in order).
__global__ __launch_bounds__(256) void k_sq_err(const float* __restrict__ yhat,
                                                const float* __restrict__ Y, long long n,
                                                double* __restrict__ part) {
    __shared__ double s_sum[256];
    double a = 0.;
    for (long long s = (long long)blockIdx.x * blockDim.x + threadIdx.x; s < n;
         s += (long long)gridDim.x * blockDim.x) {
        const double d = (double)yhat[s] - (double)Y[s];
        a += d * d;
    }
    s_sum[threadIdx.x] = a;
    __syncthreads();
    for (int w = 128; w > 0; w >>= 1) {
        if (threadIdx.x < w) s_sum[threadIdx.x] += s_sum[threadIdx.x + w];
        __syncthreads();
    }
    if (threadIdx.x == 0) part[blockIdx.x] = s_sum[0];
}

struct DevBuf {
    void* p = nullptr;
    size_t bytes = 0;
    int get(size_t b) {
        if (b <= bytes) return 0;
        dev_free(p, bytes);
        bytes = 0;
        p = dev_alloc(b);
        if (!p) return fail("out of device memory for the cost model");
        bytes = b;
        return 0;
    }
    ~DevBuf() { dev_free(p, bytes); }
    template <class T> T* as() const { return static_cast<T*>(p); }
};

}  // namespace

struct CostModel {
    DevBuf params, m, v, grad;
    DevBuf X, Y, idx, yhat, act, dyv, del, part;
    long long t = 0;           // Adam steps taken
    unsigned long long seed = 0;
};

CostModel* cost_model_new(unsigned long long seed) {
    CostModel* M = new CostModel;
    M->seed = seed;
    if (M->params.get(sizeof(float) * NP) || M->m.get(sizeof(float) * NP) ||
        M->v.get(sizeof(float) * NP) || M->grad.get(sizeof(float) * NP)) {
        delete M;
        return nullptr;
    }
    // He-uniform initial weights, zero biases
    std::mt19937_64 rng(seed);
    std::vector<float> p(NP, 0.f);
    auto fill = [&](int off, int count, int fan_in) {
        std::uniform_real_distribution<float> u(-std::sqrt(6.f / fan_in), std::sqrt(6.f / fan_in));
        for (int k = 0; k < count; ++k) p[off + k] = u(rng);
    };
    fill(O_W1, H * NF, NF);
    fill(O_W2, H * H, H);
    fill(O_W3, H, H);
    if (hipMemcpy(M->params.p, p.data(), sizeof(float) * NP, hipMemcpyHostToDevice) != hipSuccess ||
        hipMemset(M->m.p, 0, sizeof(float) * NP) != hipSuccess ||
        hipMemset(M->v.p, 0, sizeof(float) * NP) != hipSuccess) {
        set_error("cost model: device copy failed");
        delete M;
        return nullptr;
    }
    return M;
}

void cost_model_delete(CostModel* M) { delete M; }

int cost_model_params(CostModel* M, float* host, bool set) {
    if (set) TC_HIP(hipMemcpy(M->params.p, host, sizeof(float) * NP, hipMemcpyHostToDevice));
    else TC_HIP(hipMemcpy(host, M->params.p, sizeof(float) * NP, hipMemcpyDeviceToHost));
    return 0;
}

// Forward over device features dX (n x NF) into device dyhat.
int cost_model_predict_device(const CostModel* M, const float* dX, long long n, float* dyhat,
                              hipStream_t st) {
    if (n <= 0) return 0;
    const long long blocks = std::min<long long>((n + 255) / 256, 8192);
    hipLaunchKernelGGL(k_mlp_forward, dim3((unsigned)blocks), dim3(256), 0, st,
                       M->params.as<float>(), dX, (const long long*)nullptr, n, dyhat,
                       (float*)nullptr);
    TC_HIP(hipGetLastError());
    return 0;
}

int cost_model_predict(CostModel* M, const float* X, long long n, float* yhat) {
    hipStream_t st = stream();
    if (n <= 0) return 0;
    if (M->X.get(sizeof(float) * NF * n) || M->yhat.get(sizeof(float) * n)) return 2;
    TC_HIP(hipMemcpyAsync(M->X.p, X, sizeof(float) * NF * n, hipMemcpyHostToDevice, st));
    if (int rc = cost_model_predict_device(M, M->X.as<float>(), n, M->yhat.as<float>(), st))
        return rc;
    TC_HIP(hipMemcpyAsync(yhat, M->yhat.p, sizeof(float) * n, hipMemcpyDeviceToHost, st));
    TC_HIP(hipStreamSynchronize(st));
    return 0;
}

namespace {
// Forward + backward + gradient over nb samples (idx, or the first nb rows) of the
// device data in M->X, M->Y; the gradient lands in M->grad.
int forward_backward(CostModel* M, const long long* idx, long long nb, hipStream_t st) {
    if (M->act.get(sizeof(float) * 2 * H * nb) || M->del.get(sizeof(float) * 2 * H * nb) ||
        M->dyv.get(sizeof(float) * nb) || M->yhat.get(sizeof(float) * nb))
        return 2;
    const unsigned blocks = (unsigned)((nb + 255) / 256);
    hipLaunchKernelGGL(k_mlp_forward, dim3(blocks), dim3(256), 0, st, M->params.as<float>(),
                       M->X.as<float>(), idx, nb, M->yhat.as<float>(), M->act.as<float>());
    hipLaunchKernelGGL(k_mlp_backward, dim3(blocks), dim3(256), 0, st, M->params.as<float>(),
                       M->act.as<float>(), M->yhat.as<float>(), M->Y.as<float>(), idx, nb,
                       M->dyv.as<float>(), M->del.as<float>());
    hipLaunchKernelGGL(k_mlp_grad, dim3((NP + 255) / 256), dim3(256), 0, st, M->X.as<float>(),
                       idx, M->act.as<float>(), M->dyv.as<float>(), M->del.as<float>(), nb,
                       M->grad.as<float>());
    TC_HIP(hipGetLastError());
    return 0;
}

int upload_data(CostModel* M, const float* X, const float* Y, long long n, hipStream_t st) {
    if (M->X.get(sizeof(float) * NF * n) || M->Y.get(sizeof(float) * n)) return 2;
    TC_HIP(hipMemcpyAsync(M->X.p, X, sizeof(float) * NF * n, hipMemcpyHostToDevice, st));
    TC_HIP(hipMemcpyAsync(M->Y.p, Y, sizeof(float) * n, hipMemcpyHostToDevice, st));
    return 0;
}

int mean_sq_err(CostModel* M, long long n, double* loss, hipStream_t st) {
    if (M->yhat.get(sizeof(float) * n) || M->part.get(sizeof(double) * 1024)) return 2;
    if (int rc = cost_model_predict_device(M, M->X.as<float>(), n, M->yhat.as<float>(), st))
        return rc;
    hipLaunchKernelGGL(k_sq_err, dim3(1024), dim3(256), 0, st, M->yhat.as<float>(),
                       M->Y.as<float>(), n, M->part.as<double>());
    TC_HIP(hipGetLastError());
    std::vector<double> part(1024);
    TC_HIP(hipMemcpyAsync(part.data(), M->part.p, sizeof(double) * 1024, hipMemcpyDeviceToHost, st));
    TC_HIP(hipStreamSynchronize(st));
    double s = 0.;
    for (double v : part) s += v;
    *loss = s / (double)n;
    return 0;
}
}  // namespace

int cost_model_grad(CostModel* M, const float* X, const float* Y, long long n, float* grad,
                    double* loss) {
    hipStream_t st = stream();
    if (n <= 0) return fail("no samples");
    if (int rc = upload_data(M, X, Y, n, st)) return rc;
    if (int rc = forward_backward(M, nullptr, n, st)) return rc;
    TC_HIP(hipMemcpyAsync(grad, M->grad.p, sizeof(float) * NP, hipMemcpyDeviceToHost, st));
    TC_HIP(hipStreamSynchronize(st));
    return mean_sq_err(M, n, loss, st);
}

int cost_model_train(CostModel* M, const float* X, const float* Y, long long n, int steps,
                     int batch, float lr, double* loss) {
    hipStream_t st = stream();
    if (n <= 0 || steps < 0 || batch <= 0) return fail("invalid training arguments");
    if (int rc = upload_data(M, X, Y, n, st)) return rc;
    if (M->idx.get(sizeof(long long) * batch)) return 2;
    for (int k = 0; k < steps; ++k) {
        hipLaunchKernelGGL(k_batch_idx, dim3((batch + 255) / 256), dim3(256), 0, st, M->seed,
                           (long long)M->t, n, batch, M->idx.as<long long>());
        if (int rc = forward_backward(M, M->idx.as<long long>(), batch, st)) return rc;
        ++M->t;
        const float c1 = 1.f - std::pow(0.9f, (float)M->t), c2 = 1.f - std::pow(0.999f, (float)M->t);
        hipLaunchKernelGGL(k_adam, dim3((NP + 255) / 256), dim3(256), 0, st, M->params.as<float>(),
                           M->grad.as<float>(), M->m.as<float>(), M->v.as<float>(), lr, c1, c2);
        TC_HIP(hipGetLastError());
    }
    return mean_sq_err(M, n, loss, st);
}

}  // namespace tc
