// Tile-skip network on the device: forward, backward, weight gradients and
// Adam, and the m1_mlp_* entry points (include/m1march.h).  No reference
// counterpart.
//
// Training runs entirely on the GPU: per minibatch one forward launch (one
// thread per sample, weights staged in LDS as bf16), one backward launch
// (per-sample deltas of the mean binary cross-entropy), one weight-gradient
// launch (one thread per parameter summing over the minibatch in sample
// order, so gradients are deterministic) and one Adam launch that also
// refreshes the bf16 copy of the weights.
#include <hip/hip_runtime.h>

#include <algorithm>
#include <cmath>
#include <cstring>
#include <numeric>
#include <random>
#include <string>
#include <vector>

#include "m1_host.h"
#include "m1_mlp.h"

using namespace m1;
using namespace m1::mlp;

namespace m1 {
namespace mlp {
namespace {

constexpr int NACT = 3 * NH;      // h1, h2, h3 per sample
constexpr int NDELTA = 3 * NH + 1;  // d1, d2, d3, dz per sample
constexpr int FWD_THREADS = WG;

__device__ inline void stage_weights(const uint16_t* w, uint16_t* wl)
{
    for (int k = threadIdx.x; k < NPARAM; k += blockDim.x)
        wl[k] = w[k];
    __syncthreads();
}

__global__ void __launch_bounds__(FWD_THREADS)
    forward_kernel(DeviceParams prm, const float* x, int n, float* logit, float* acts)
{
    __shared__ uint16_t wl[NPARAM];
    __shared__ float A[NH][WG], B[NH][WG];
    stage_weights(prm.w, wl);
    const int t = threadIdx.x;
    const int s = blockIdx.x * blockDim.x + t;
    if (s >= n)
        return;  // no barrier follows
    float xi[NIN];
#pragma unroll
    for (int i = 0; i < NIN; ++i)
        xi[i] = x[long(s) * NIN + i];
    float* a = acts ? acts + long(s) * NACT : nullptr;
    logit[s] = forward_one(wl, prm.p, xi, A, B, t, a, a ? a + NH : nullptr,
                           a ? a + 2 * NH : nullptr);
}

// Per-sample deltas of the mean loss L = (1/n) sum BCE(sigmoid(z), y) and
// the per-sample loss (stable form max(z,0) - z y + log(1 + exp(-|z|))).
__global__ void __launch_bounds__(FWD_THREADS)
    backward_kernel(DeviceParams prm, const float* acts, const float* logit, const float* y,
                    int n, float* deltas, float* loss)
{
    __shared__ uint16_t wl[NPARAM];
    __shared__ float D3[NH][WG], D2[NH][WG];
    stage_weights(prm.w, wl);
    const int t = threadIdx.x;
    const int s = blockIdx.x * blockDim.x + t;
    if (s >= n)
        return;  // no barrier follows
    const float z = logit[s];
    const float lab = y[s];
    loss[s] = fmaxf(z, 0.f) - z * lab + log1pf(expf(-fabsf(z)));
    const float dz = (sigmoid(z) - lab) / float(n);
    const float* h1 = acts + long(s) * NACT;
    const float* h2 = h1 + NH;
    const float* h3 = h2 + NH;
    float* d = deltas + long(s) * NDELTA;
    // deltas of the current sample in LDS, [unit][thread]
    for (int j = 0; j < NH; ++j)
        D3[j][t] = h3[j] > 0.f ? dz * bf16_to_float(wl[OFF_W4 + j]) : 0.f;
    for (int i = 0; i < NH; ++i) {
        float g = 0.f;
#pragma unroll 8
        for (int j = 0; j < NH; ++j)
            g = fmaf(D3[j][t], bf16_to_float(wl[OFF_W3 + i * NH + j]), g);
        D2[i][t] = h2[i] > 0.f ? g : 0.f;
    }
    for (int i = 0; i < NH; ++i) {
        float g = 0.f;
#pragma unroll 8
        for (int j = 0; j < NH; ++j)
            g = fmaf(D2[j][t], bf16_to_float(wl[OFF_W2 + i * NH + j]), g);
        d[i] = h1[i] > 0.f ? g : 0.f;
    }
    for (int j = 0; j < NH; ++j) {
        d[NH + j] = D2[j][t];
        d[2 * NH + j] = D3[j][t];
    }
    d[3 * NH] = dz;
}

// One thread per parameter: dL/dW[i][j] = sum_s a_s[i] delta_s[j] and
// dL/db[j] = sum_s delta_s[j], summed in sample order.
__global__ void weight_grad_kernel(const float* x, const float* acts, const float* deltas, int n,
                                   float* grad)
{
    const int k = blockIdx.x * blockDim.x + threadIdx.x;
    if (k >= NPARAM)
        return;
    // layer l: input stride/offset (x or acts) and delta offset
    const float* in;
    long in_stride;
    int in_off, i, d_off, j;
    bool bias;
    if (k < OFF_B1) {
        in = x; in_stride = NIN; in_off = 0; i = k / NH; j = k % NH; d_off = 0; bias = false;
    } else if (k < OFF_W2) {
        in = nullptr; in_stride = 0; in_off = 0; i = 0; j = k - OFF_B1; d_off = 0; bias = true;
    } else if (k < OFF_B2) {
        const int q = k - OFF_W2;
        in = acts; in_stride = NACT; in_off = 0; i = q / NH; j = q % NH; d_off = NH; bias = false;
    } else if (k < OFF_W3) {
        in = nullptr; in_stride = 0; in_off = 0; i = 0; j = k - OFF_B2; d_off = NH; bias = true;
    } else if (k < OFF_B3) {
        const int q = k - OFF_W3;
        in = acts; in_stride = NACT; in_off = NH; i = q / NH; j = q % NH; d_off = 2 * NH;
        bias = false;
    } else if (k < OFF_W4) {
        in = nullptr; in_stride = 0; in_off = 0; i = 0; j = k - OFF_B3; d_off = 2 * NH;
        bias = true;
    } else if (k < OFF_B4) {
        in = acts; in_stride = NACT; in_off = 2 * NH; i = k - OFF_W4; j = 0; d_off = 3 * NH;
        bias = false;
    } else {
        in = nullptr; in_stride = 0; in_off = 0; i = 0; j = 0; d_off = 3 * NH; bias = true;
    }
    float g = 0.f;
    for (int s = 0; s < n; ++s) {
        const float dl = deltas[long(s) * NDELTA + d_off + j];
        g = bias ? g + dl : fmaf(in[long(s) * in_stride + in_off + i], dl, g);
    }
    grad[k] = g;
}

__global__ void adam_kernel(float* p, uint16_t* w16, const float* grad, float* m, float* v,
                            float lr, float beta1, float beta2, float eps, float c1, float c2)
{
    const int k = blockIdx.x * blockDim.x + threadIdx.x;
    if (k >= NPARAM)
        return;
    const float g = grad[k];
    const float mk = beta1 * m[k] + (1.f - beta1) * g;
    const float vk = beta2 * v[k] + (1.f - beta2) * g * g;
    m[k] = mk;
    v[k] = vk;
    const float pk = p[k] - lr * (mk / c1) / (sqrtf(vk / c2) + eps);
    p[k] = pk;
    w16[k] = float_to_bf16(pk);
}

__global__ void round_kernel(const float* p, uint16_t* w16)
{
    const int k = blockIdx.x * blockDim.x + threadIdx.x;
    if (k < NPARAM)
        w16[k] = float_to_bf16(p[k]);
}

__global__ void gather_kernel(const float* x, const float* y, const int* idx, int n, float* xb,
                              float* yb)
{
    const int s = blockIdx.x * blockDim.x + threadIdx.x;
    if (s >= n)
        return;
    const long r = idx[s];
    for (int i = 0; i < NIN; ++i)
        xb[long(s) * NIN + i] = x[r * NIN + i];
    yb[s] = y[r];
}

unsigned blocks(long n, int t) { return unsigned((n + t - 1) / t); }

}  // namespace

hipError_t launch_forward(DeviceParams prm, const float* x, int n, float* logit, float* acts,
                          hipStream_t s)
{
    if (n <= 0)
        return hipSuccess;
    hipLaunchKernelGGL(forward_kernel, dim3(blocks(n, FWD_THREADS)), dim3(FWD_THREADS), 0, s,
                       prm, x, n, logit, acts);
    return hipGetLastError();
}

hipError_t launch_backward(DeviceParams prm, const float* acts, const float* logit,
                           const float* y, int n, float* deltas, float* loss, hipStream_t s)
{
    if (n <= 0)
        return hipSuccess;
    hipLaunchKernelGGL(backward_kernel, dim3(blocks(n, FWD_THREADS)), dim3(FWD_THREADS), 0, s,
                       prm, acts, logit, y, n, deltas, loss);
    return hipGetLastError();
}

hipError_t launch_weight_grad(const float* x, const float* acts, const float* deltas, int n,
                              float* grad, hipStream_t s)
{
    hipLaunchKernelGGL(weight_grad_kernel, dim3(blocks(NPARAM, 256)), dim3(256), 0, s, x, acts,
                       deltas, n, grad);
    return hipGetLastError();
}

hipError_t launch_adam(float* p, uint16_t* w16, const float* grad, float* m, float* v,
                       float lr, float beta1, float beta2, float eps, int step, hipStream_t s)
{
    const float c1 = 1.f - std::pow(beta1, float(step));
    const float c2 = 1.f - std::pow(beta2, float(step));
    hipLaunchKernelGGL(adam_kernel, dim3(blocks(NPARAM, 256)), dim3(256), 0, s, p, w16, grad, m,
                       v, lr, beta1, beta2, eps, c1, c2);
    return hipGetLastError();
}

hipError_t launch_round_bf16(const float* p, uint16_t* w16, hipStream_t s)
{
    hipLaunchKernelGGL(round_kernel, dim3(blocks(NPARAM, 256)), dim3(256), 0, s, p, w16);
    return hipGetLastError();
}

}  // namespace mlp
}  // namespace m1

// ------------------------------------------------------------------ C ABI
struct m1_mlp {
    int device = 0;
    hipStream_t stream = nullptr;
    std::string err;
    float *p = nullptr, *m = nullptr, *v = nullptr, *grad = nullptr;
    uint16_t* w16 = nullptr;
    int adam_step = 0;
    // scratch for n samples
    long cap = 0;
    float *x = nullptr, *y = nullptr, *acts = nullptr, *deltas = nullptr, *logit = nullptr,
          *loss = nullptr;
    // training data (full set) and minibatch index
    long data_cap = 0;
    float *dx = nullptr, *dy = nullptr;
    int* idx = nullptr;

    void free_scratch();
    ~m1_mlp();
};

void m1_mlp::free_scratch()
{
    for (void* q : {(void*)x, (void*)y, (void*)acts, (void*)deltas, (void*)logit, (void*)loss})
        if (q)
            (void)hipFree(q);
    x = y = acts = deltas = logit = loss = nullptr;
    cap = 0;
}

m1_mlp::~m1_mlp()
{
    (void)hipSetDevice(device);
    if (stream)
        (void)hipStreamSynchronize(stream);
    free_scratch();
    for (void* q : {(void*)p, (void*)m, (void*)v, (void*)grad, (void*)w16, (void*)dx, (void*)dy,
                    (void*)idx})
        if (q)
            (void)hipFree(q);
    if (stream)
        (void)hipStreamDestroy(stream);
}

namespace m1 {
namespace mlp {
DeviceParams device_params_of(const m1_mlp* model) { return {model->w16, model->p}; }
int device_of(const m1_mlp* model) { return model->device; }
hipError_t finish(const m1_mlp* model) { return hipStreamSynchronize(model->stream); }
}  // namespace mlp
}  // namespace m1

namespace {

std::string g_mlp_create_error;

template <class F>
int guarded(m1_mlp* c, F&& f)
{
    if (!c)
        return M1_ERR_STATE;
    return run_guarded(c->device, c->err, f);
}

void ensure_scratch(m1_mlp* c, long n)
{
    if (n <= c->cap)
        return;
    c->free_scratch();
    dev_alloc(&c->x, size_t(n) * NIN, "hipMalloc mlp inputs");
    dev_alloc(&c->y, size_t(n), "hipMalloc mlp labels");
    dev_alloc(&c->acts, size_t(n) * 3 * NH, "hipMalloc mlp activations");
    dev_alloc(&c->deltas, size_t(n) * (3 * NH + 1), "hipMalloc mlp deltas");
    dev_alloc(&c->logit, size_t(n), "hipMalloc mlp logits");
    dev_alloc(&c->loss, size_t(n), "hipMalloc mlp loss");
    c->cap = n;
}

void check_n(long n)
{
    if (n <= 0)
        throw ValueError{"at least one sample is needed"};
    if (n > (1L << 26))
        throw ValueError{"too many samples for one call"};
}

}  // namespace

extern "C" {

int32_t m1_mlp_param_count(void) { return NPARAM; }

int m1_mlp_create(m1_mlp** out, int device)
{
    if (!out) {
        g_mlp_create_error = "out must not be NULL";
        return M1_ERR_VALUE;
    }
    *out = nullptr;
    int count = 0;
    if (hipGetDeviceCount(&count) != hipSuccess || count == 0) {
        g_mlp_create_error = "no HIP device is available";
        return M1_ERR_HIP;
    }
    if (device < 0 || device >= count) {
        g_mlp_create_error = "device ordinal out of range";
        return M1_ERR_VALUE;
    }
    m1_mlp* c = new m1_mlp;
    c->device = device;
    try {
        hip_check(hipSetDevice(device), "hipSetDevice");
        hip_check(hipStreamCreateWithFlags(&c->stream, hipStreamNonBlocking), "hipStreamCreate");
        dev_alloc(&c->p, NPARAM, "hipMalloc mlp parameters");
        dev_alloc(&c->m, NPARAM, "hipMalloc mlp moments");
        dev_alloc(&c->v, NPARAM, "hipMalloc mlp moments");
        dev_alloc(&c->grad, NPARAM, "hipMalloc mlp gradient");
        dev_alloc(&c->w16, NPARAM, "hipMalloc mlp bf16 weights");
        hip_check(hipMemsetAsync(c->p, 0, NPARAM * sizeof(float), c->stream), "mlp reset");
        hip_check(hipMemsetAsync(c->w16, 0, NPARAM * sizeof(uint16_t), c->stream), "mlp reset");
        hip_check(hipMemsetAsync(c->m, 0, NPARAM * sizeof(float), c->stream), "mlp reset");
        hip_check(hipMemsetAsync(c->v, 0, NPARAM * sizeof(float), c->stream), "mlp reset");
        hip_check(hipStreamSynchronize(c->stream), "hipStreamSynchronize");
    } catch (const HipError& e) {
        g_mlp_create_error = e.msg;
        delete c;
        return M1_ERR_HIP;
    }
    *out = c;
    return M1_OK;
}

void m1_mlp_destroy(m1_mlp* model) { delete model; }

const char* m1_mlp_last_error(const m1_mlp* model)
{
    return model ? model->err.c_str() : g_mlp_create_error.c_str();
}

int m1_mlp_set_params(m1_mlp* c, const float* params)
{
    return guarded(c, [&] {
        if (!params)
            throw ValueError{"params must not be NULL"};
        for (int k = 0; k < NPARAM; ++k)
            if (!std::isfinite(params[k]))
                throw ValueError{"parameters must be finite"};
        hip_check(hipMemcpyAsync(c->p, params, NPARAM * sizeof(float), hipMemcpyHostToDevice,
                                 c->stream),
                  "parameter upload");
        hip_check(launch_round_bf16(c->p, c->w16, c->stream), "bf16 rounding kernel");
        hip_check(hipMemsetAsync(c->m, 0, NPARAM * sizeof(float), c->stream), "mlp reset");
        hip_check(hipMemsetAsync(c->v, 0, NPARAM * sizeof(float), c->stream), "mlp reset");
        c->adam_step = 0;
        hip_check(hipStreamSynchronize(c->stream), "hipStreamSynchronize");
    });
}

int m1_mlp_get_params(m1_mlp* c, float* params)
{
    return guarded(c, [&] {
        if (!params)
            throw ValueError{"params must not be NULL"};
        hip_check(hipMemcpyAsync(params, c->p, NPARAM * sizeof(float), hipMemcpyDeviceToHost,
                                 c->stream),
                  "parameter download");
        hip_check(hipStreamSynchronize(c->stream), "hipStreamSynchronize");
    });
}

int m1_mlp_forward(m1_mlp* c, const float* x, int32_t n, float* prob)
{
    return guarded(c, [&] {
        check_n(n);
        if (!x || !prob)
            throw ValueError{"x and prob must not be NULL"};
        ensure_scratch(c, n);
        hip_check(hipMemcpyAsync(c->x, x, size_t(n) * NIN * sizeof(float),
                                 hipMemcpyHostToDevice, c->stream),
                  "input upload");
        hip_check(launch_forward({c->w16, c->p}, c->x, n, c->logit, nullptr, c->stream),
                  "mlp forward kernel");
        std::vector<float> z(static_cast<size_t>(n));
        hip_check(hipMemcpyAsync(z.data(), c->logit, size_t(n) * sizeof(float),
                                 hipMemcpyDeviceToHost, c->stream),
                  "logit download");
        hip_check(hipStreamSynchronize(c->stream), "hipStreamSynchronize");
        for (int s = 0; s < n; ++s)
            prob[s] = 1.f / (1.f + std::exp(-z[size_t(s)]));
    });
}

int m1_mlp_loss_grad(m1_mlp* c, const float* x, const float* y, int32_t n, double* loss,
                     float* grad)
{
    return guarded(c, [&] {
        check_n(n);
        if (!x || !y)
            throw ValueError{"x and y must not be NULL"};
        ensure_scratch(c, n);
        hip_check(hipMemcpyAsync(c->x, x, size_t(n) * NIN * sizeof(float),
                                 hipMemcpyHostToDevice, c->stream),
                  "input upload");
        hip_check(hipMemcpyAsync(c->y, y, size_t(n) * sizeof(float), hipMemcpyHostToDevice,
                                 c->stream),
                  "label upload");
        const DeviceParams prm{c->w16, c->p};
        hip_check(launch_forward(prm, c->x, n, c->logit, c->acts, c->stream),
                  "mlp forward kernel");
        hip_check(launch_backward(prm, c->acts, c->logit, c->y, n, c->deltas, c->loss, c->stream),
                  "mlp backward kernel");
        hip_check(launch_weight_grad(c->x, c->acts, c->deltas, n, c->grad, c->stream),
                  "mlp gradient kernel");
        std::vector<float> l(static_cast<size_t>(n));
        hip_check(hipMemcpyAsync(l.data(), c->loss, size_t(n) * sizeof(float),
                                 hipMemcpyDeviceToHost, c->stream),
                  "loss download");
        if (grad)
            hip_check(hipMemcpyAsync(grad, c->grad, NPARAM * sizeof(float),
                                     hipMemcpyDeviceToHost, c->stream),
                      "gradient download");
        hip_check(hipStreamSynchronize(c->stream), "hipStreamSynchronize");
        if (loss) {
            double sum = 0.0;
            for (float v : l)
                sum += v;
            *loss = sum / n;
        }
    });
}

int m1_mlp_train(m1_mlp* c, const float* x, const float* y, int32_t n, int32_t epochs,
                 int32_t batch, float lr, uint64_t seed, double* final_loss)
{
    return guarded(c, [&] {
        check_n(n);
        if (!x || !y)
            throw ValueError{"x and y must not be NULL"};
        if (epochs < 1 || batch < 1)
            throw ValueError{"epochs and batch must be positive"};
        if (!(lr > 0.f && std::isfinite(lr)))
            throw ValueError{"learning rate must be positive"};
        const int bs = std::min(batch, n);
        ensure_scratch(c, bs);
        if (n > c->data_cap) {
            for (void* q : {(void*)c->dx, (void*)c->dy, (void*)c->idx})
                if (q)
                    (void)hipFree(q);
            c->dx = c->dy = nullptr;
            c->idx = nullptr;
            dev_alloc(&c->dx, size_t(n) * NIN, "hipMalloc training inputs");
            dev_alloc(&c->dy, size_t(n), "hipMalloc training labels");
            dev_alloc(&c->idx, size_t(n), "hipMalloc training order");
            c->data_cap = n;
        }
        hip_check(hipMemcpyAsync(c->dx, x, size_t(n) * NIN * sizeof(float),
                                 hipMemcpyHostToDevice, c->stream),
                  "training input upload");
        hip_check(hipMemcpyAsync(c->dy, y, size_t(n) * sizeof(float), hipMemcpyHostToDevice,
                                 c->stream),
                  "training label upload");
        std::vector<int> order(static_cast<size_t>(n));
        std::iota(order.begin(), order.end(), 0);
        std::mt19937_64 rng(seed);
        const DeviceParams prm{c->w16, c->p};
        double last = 0.0;
        std::vector<float> l(static_cast<size_t>(bs));
        for (int e = 0; e < epochs; ++e) {
            std::shuffle(order.begin(), order.end(), rng);
            hip_check(hipStreamSynchronize(c->stream), "hipStreamSynchronize");
            hip_check(hipMemcpyAsync(c->idx, order.data(), size_t(n) * sizeof(int),
                                     hipMemcpyHostToDevice, c->stream),
                      "order upload");
            double sum = 0.0;
            long cnt = 0;
            for (int b0 = 0; b0 + bs <= n; b0 += bs) {
                hipLaunchKernelGGL(gather_kernel, dim3(unsigned((bs + 255) / 256)), dim3(256), 0,
                                   c->stream, c->dx, c->dy, c->idx + b0, bs, c->x, c->y);
                hip_check(hipGetLastError(), "gather kernel");
                hip_check(launch_forward(prm, c->x, bs, c->logit, c->acts, c->stream),
                          "mlp forward kernel");
                hip_check(launch_backward(prm, c->acts, c->logit, c->y, bs, c->deltas, c->loss,
                                          c->stream),
                          "mlp backward kernel");
                hip_check(launch_weight_grad(c->x, c->acts, c->deltas, bs, c->grad, c->stream),
                          "mlp gradient kernel");
                c->adam_step += 1;
                hip_check(launch_adam(c->p, c->w16, c->grad, c->m, c->v, lr, 0.9f, 0.999f, 1e-8f,
                                      c->adam_step, c->stream),
                          "adam kernel");
                if (e == epochs - 1) {
                    hip_check(hipMemcpyAsync(l.data(), c->loss, size_t(bs) * sizeof(float),
                                             hipMemcpyDeviceToHost, c->stream),
                              "loss download");
                    hip_check(hipStreamSynchronize(c->stream), "hipStreamSynchronize");
                    for (float v : l)
                        sum += v;
                    cnt += bs;
                }
            }
            if (e == epochs - 1 && cnt > 0)
                last = sum / double(cnt);
        }
        hip_check(hipStreamSynchronize(c->stream), "hipStreamSynchronize");
        if (final_loss)
            *final_loss = last;
    });
}

}  // extern "C"
