// sbc_mlp.hip -- the learned cost model (include/sbc_costmodel.h): window
// feature extraction, FP32 training kernels and BF16-MFMA inference for a
// 12-64-64-1 ReLU MLP.  Everything here only feeds scheduling decisions.
#include <hip/hip_runtime.h>

#include <cmath>
#include <cstdint>
#include <cstdio>
#include <cstring>
#include <vector>

#include "sbc_costmodel.h"
#include "sbc_kernels.h"

namespace {

constexpr int NF = SBC_COST_FEATURES;  // 12 real inputs
constexpr int NI = 16;                 // padded input width (one MFMA K block)
constexpr int NH = 64;
constexpr int OFF_W1 = 0;
constexpr int OFF_B1 = OFF_W1 + NI * NH;
constexpr int OFF_W2 = OFF_B1 + NH;
constexpr int OFF_B2 = OFF_W2 + NH * NH;
constexpr int OFF_W3 = OFF_B2 + NH;
constexpr int OFF_B3 = OFF_W3 + NH;
constexpr int NP = OFF_B3 + 1;

typedef short bf16x4 __attribute__((ext_vector_type(4)));
typedef float f32x4 __attribute__((ext_vector_type(4)));

// FP32 -> BF16, round to nearest even (finite inputs).
__host__ __device__ inline unsigned short f2bf(float x)
{
    uint32_t b;
    memcpy(&b, &x, 4);
    b += 0x7FFFu + ((b >> 16) & 1u);
    return (unsigned short)(b >> 16);
}

__device__ inline float relu(float x) { return x > 0.f ? x : 0.f; }

// ---------------------------------------------------------------- training

// Forward pass, one thread per sample; keeps pre-activations for backward.
__global__ void mlp_fwd_kernel(const float* __restrict__ X, int n, const float* __restrict__ p, float* z1,
                               float* z2, float* out)
{
    const int s = blockIdx.x * blockDim.x + threadIdx.x;
    if (s >= n) return;
    float x[NF];
    for (int i = 0; i < NF; ++i) x[i] = X[(size_t)s * NF + i];
    float h1[NH];
    for (int j = 0; j < NH; ++j) {
        float a = p[OFF_B1 + j];
        for (int i = 0; i < NF; ++i) a += x[i] * p[OFF_W1 + i * NH + j];
        z1[(size_t)s * NH + j] = a;
        h1[j] = relu(a);
    }
    float o = p[OFF_B3];
    for (int k = 0; k < NH; ++k) {
        float a = p[OFF_B2 + k];
        for (int j = 0; j < NH; ++j) a += h1[j] * p[OFF_W2 + j * NH + k];
        z2[(size_t)s * NH + k] = a;
        o += relu(a) * p[OFF_W3 + k];
    }
    out[s] = o;
}

// Backward through the layers for each sample of the mean squared error
// (1/n) * sum (out - y)^2.
__global__ void mlp_bwd_kernel(const float* __restrict__ y, int n, float inv_n, const float* __restrict__ p,
                               const float* __restrict__ z1, const float* __restrict__ z2,
                               const float* __restrict__ out, float* dz1, float* dz2, float* dout)
{
    const int s = blockIdx.x * blockDim.x + threadIdx.x;
    if (s >= n) return;
    const float d = 2.f * (out[s] - y[s]) * inv_n;
    dout[s] = d;
    float g2[NH];
    for (int k = 0; k < NH; ++k) {
        g2[k] = z2[(size_t)s * NH + k] > 0.f ? d * p[OFF_W3 + k] : 0.f;
        dz2[(size_t)s * NH + k] = g2[k];
    }
    for (int j = 0; j < NH; ++j) {
        float a = 0.f;
        for (int k = 0; k < NH; ++k) a += p[OFF_W2 + j * NH + k] * g2[k];
        dz1[(size_t)s * NH + j] = z1[(size_t)s * NH + j] > 0.f ? a : 0.f;
    }
}

// Gradient of each parameter: one thread per parameter summing its
// per-sample contributions in sample order (deterministic).
__global__ void mlp_wgrad_kernel(const float* __restrict__ X, int n, const float* __restrict__ z1,
                                 const float* __restrict__ z2, const float* __restrict__ dz1,
                                 const float* __restrict__ dz2, const float* __restrict__ dout, float* g)
{
    const int q = blockIdx.x * blockDim.x + threadIdx.x;
    if (q >= NP) return;
    float a = 0.f;
    if (q < OFF_B1) {
        const int i = q / NH, j = q % NH;
        if (i < NF)
            for (int s = 0; s < n; ++s) a += X[(size_t)s * NF + i] * dz1[(size_t)s * NH + j];
    } else if (q < OFF_W2) {
        const int j = q - OFF_B1;
        for (int s = 0; s < n; ++s) a += dz1[(size_t)s * NH + j];
    } else if (q < OFF_B2) {
        const int j = (q - OFF_W2) / NH, k = (q - OFF_W2) % NH;
        for (int s = 0; s < n; ++s) a += relu(z1[(size_t)s * NH + j]) * dz2[(size_t)s * NH + k];
    } else if (q < OFF_W3) {
        const int k = q - OFF_B2;
        for (int s = 0; s < n; ++s) a += dz2[(size_t)s * NH + k];
    } else if (q < OFF_B3) {
        const int k = q - OFF_W3;
        for (int s = 0; s < n; ++s) a += relu(z2[(size_t)s * NH + k]) * dout[s];
    } else {
        for (int s = 0; s < n; ++s) a += dout[s];
    }
    g[q] = a;
}

__global__ void mlp_adam_kernel(float* p, const float* __restrict__ g, float* m, float* v, float lr, float b1,
                                float b2, float eps, float c1, float c2)
{
    const int q = blockIdx.x * blockDim.x + threadIdx.x;
    if (q >= NP) return;
    if (q < OFF_B1 && q / NH >= NF) return;  // padding rows of W1 stay 0
    const float gq = g[q];
    const float mq = b1 * m[q] + (1.f - b1) * gq;
    const float vq = b2 * v[q] + (1.f - b2) * gq * gq;
    m[q] = mq;
    v[q] = vq;
    p[q] -= lr * (mq / c1) / (sqrtf(vq / c2) + eps);
}

__global__ void mlp_gather_kernel(const float* __restrict__ X, const float* __restrict__ y,
                                  const int32_t* __restrict__ idx, int n, float* Xb, float* yb)
{
    const int s = blockIdx.x * blockDim.x + threadIdx.x;
    if (s >= n) return;
    const int src = idx[s];
    for (int i = 0; i < NF; ++i) Xb[(size_t)s * NF + i] = X[(size_t)src * NF + i];
    yb[s] = y[src];
}

// ---------------------------------------------------------------- inference

__global__ void mlp_fp32_kernel(const float* __restrict__ X, int n, const float* __restrict__ p, float* out)
{
    const int s = blockIdx.x * blockDim.x + threadIdx.x;
    if (s >= n) return;
    float x[NF];
    for (int i = 0; i < NF; ++i) x[i] = X[(size_t)s * NF + i];
    float h1[NH];
    for (int j = 0; j < NH; ++j) {
        float a = p[OFF_B1 + j];
        for (int i = 0; i < NF; ++i) a += x[i] * p[OFF_W1 + i * NH + j];
        h1[j] = relu(a);
    }
    float o = p[OFF_B3];
    for (int k = 0; k < NH; ++k) {
        float a = p[OFF_B2 + k];
        for (int j = 0; j < NH; ++j) a += h1[j] * p[OFF_W2 + j * NH + k];
        o += relu(a) * p[OFF_W3 + k];
    }
    out[s] = o;
}

// BF16 weights and activations on v_mfma_f32_16x16x16_bf16 with FP32
// accumulation.  One wavefront evaluates 16 samples.  Operand layouts of the
// 16x16x16 MFMA (lane l, j = 0..3):
//   A[row = l % 16][k = 4 * (l / 16) + j],  B[k = 4 * (l / 16) + j][col = l % 16],
//   D[row = 4 * (l / 16) + j][col = l % 16].
// Layer-1 outputs go through LDS to be re-read in the A layout of layer 2.
__global__ __launch_bounds__(256) void mlp_mfma_kernel(const float* __restrict__ X, int n,
                                                       const unsigned short* __restrict__ w1b,
                                                       const unsigned short* __restrict__ w2b,
                                                       const float* __restrict__ p, float* out)
{
    __shared__ unsigned short h1s[4][16][NH + 4];
    const int wave = threadIdx.x >> 6, lane = threadIdx.x & 63;
    const int s0 = (blockIdx.x * 4 + wave) * 16;
    const int r16 = lane & 15, q4 = lane >> 4;
    bf16x4 a;
    for (int j = 0; j < 4; ++j) {
        const int k = 4 * q4 + j, row = s0 + r16;
        const float xv = (row < n && k < NF) ? X[(size_t)row * NF + k] : 0.f;
        a[j] = (short)f2bf(xv);
    }
    for (int nb = 0; nb < 4; ++nb) {
        bf16x4 b;
        for (int j = 0; j < 4; ++j) b[j] = (short)w1b[(4 * q4 + j) * NH + nb * 16 + r16];
        f32x4 acc = {0.f, 0.f, 0.f, 0.f};
        acc = __builtin_amdgcn_mfma_f32_16x16x16bf16_1k(a, b, acc, 0, 0, 0);
        const float bias = p[OFF_B1 + nb * 16 + r16];
        for (int j = 0; j < 4; ++j) h1s[wave][4 * q4 + j][nb * 16 + r16] = f2bf(relu(acc[j] + bias));
    }
    __syncthreads();
    float part[4] = {0.f, 0.f, 0.f, 0.f};
    for (int nb = 0; nb < 4; ++nb) {
        f32x4 acc = {0.f, 0.f, 0.f, 0.f};
        for (int kb = 0; kb < 4; ++kb) {
            bf16x4 aa, bb;
            for (int j = 0; j < 4; ++j) {
                aa[j] = (short)h1s[wave][r16][kb * 16 + 4 * q4 + j];
                bb[j] = (short)w2b[(kb * 16 + 4 * q4 + j) * NH + nb * 16 + r16];
            }
            acc = __builtin_amdgcn_mfma_f32_16x16x16bf16_1k(aa, bb, acc, 0, 0, 0);
        }
        const int col = nb * 16 + r16;
        const float bias = p[OFF_B2 + col], w3 = p[OFF_W3 + col];
        for (int j = 0; j < 4; ++j) part[j] += relu(acc[j] + bias) * w3;
    }
    for (int j = 0; j < 4; ++j) {
        float v = part[j];
        for (int d = 1; d < 16; d <<= 1) v += __shfl_xor(v, d);
        const int row = s0 + 4 * q4 + j;
        if (r16 == 0 && row < n) out[row] = v + p[OFF_B3];
    }
}

__global__ void mlp_tobf16_kernel(const float* __restrict__ p, unsigned short* w1b, unsigned short* w2b)
{
    const int q = blockIdx.x * blockDim.x + threadIdx.x;
    if (q < NI * NH) w1b[q] = f2bf(p[OFF_W1 + q]);
    if (q < NH * NH) w2b[q] = f2bf(p[OFF_W2 + q]);
}

// ---------------------------------------------------------------- features

__global__ void sbc_feature_kernel(SbcFeatArgs A)
{
    const int64_t t = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
    if (t >= (int64_t)A.count * A.K) return;
    const int i = (int)(t / A.K), r = (int)(t % A.K);
    const size_t off = (size_t)r * (size_t)A.nv;
    const int32_t u = A.u[i], v = A.v[i];
    const int32_t du = A.dist[off + u], dv = A.dist[off + v];
    float f[NF];
    for (int q = 0; q < NF; ++q) f[q] = 0.f;
    const int32_t INF = 0x7fffffff;
    if (!(du == INF && dv == INF) && du != dv) {
        const int64_t gap = (int64_t)du - (int64_t)dv;
        const int32_t s = gap > 0 ? u : v, pv = gap > 0 ? v : u;
        const int64_t ag = gap > 0 ? gap : -gap;
        const int32_t ds = A.dist[off + s], dp = A.dist[off + pv];
        f[0] = 1.f;
        f[1] = ag >= 2 ? 1.f : 0.f;
        f[2] = (float)(ag < 16 ? ag : 16) / 16.f;
        f[3] = (float)(ds < 16 ? ds : 16) / 16.f;
        f[4] = (float)(dp < 16 ? dp : 16) / 16.f;
        f[5] = log1pf((float)A.vinfo[s].deg) / 10.f;
        f[6] = log1pf((float)A.vinfo[pv].deg) / 10.f;
        f[7] = dp == INF ? 0.f : (float)log1p((double)A.vr[off + pv].sigma) / 20.f;
        f[8] = ds == INF ? 0.f : (float)log1p((double)A.vr[off + s].sigma) / 20.f;
        f[9] = (float)log1p(A.vr[off + s].delta) / 10.f;
        f[10] = (float)log1p(A.vr[off + pv].delta) / 10.f;
    }
    f[11] = A.hist ? A.hist[r] / 10.f : 0.f;
    for (int q = 0; q < NF; ++q) A.X[t * NF + q] = f[q];
}

}  // namespace

extern "C" hipError_t sbc_launch_features(const SbcFeatArgs* A, hipStream_t st)
{
    const int64_t n = (int64_t)A->count * A->K;
    if (n <= 0) return hipSuccess;
    hipLaunchKernelGGL(sbc_feature_kernel, dim3((unsigned)((n + 255) / 256)), dim3(256), 0, st, *A);
    return hipGetLastError();
}

// Per-root history feature, kept on the device: hist[r] = decay * hist[r] +
// (1 - decay) * mean over the window's items of log1p(affected), from the
// item records of the last batched call (rec[i * K + r]).
__global__ void sbc_hist_kernel(const SbcItemRec* __restrict__ rec, int count, int K, float decay, float* hist)
{
    const int r = blockIdx.x * blockDim.x + threadIdx.x;
    if (r >= K) return;
    float acc = 0.0f;
    for (int i = 0; i < count; ++i) acc += log1pf((float)rec[(size_t)i * K + r].affected);
    hist[r] = decay * hist[r] + (1.0f - decay) * (acc / (float)count);
}

extern "C" hipError_t sbc_launch_hist_update(const SbcItemRec* rec, int count, int K, float decay, float* hist,
                                             hipStream_t st)
{
    if (count <= 0 || K <= 0) return hipSuccess;
    hipLaunchKernelGGL(sbc_hist_kernel, dim3((unsigned)((K + 255) / 256)), dim3(256), 0, st, rec, count, K, decay, hist);
    return hipGetLastError();
}

// ---------------------------------------------------------------- model object

struct sbcCostModel {
    float* d_p = nullptr;
    float* d_m = nullptr;
    float* d_v = nullptr;
    float* d_g = nullptr;
    unsigned short* d_w1b = nullptr;
    unsigned short* d_w2b = nullptr;
    int64_t t = 0;  // Adam step
    bool bf16_stale = true;
};

static thread_local char g_mlp_err[256] = "";
static sbcStatus mlp_err(sbcStatus st, const char* what, hipError_t e)
{
    snprintf(g_mlp_err, sizeof(g_mlp_err), "%s: %s", what, hipGetErrorString(e));
    fprintf(stderr, "streamingbc_hip cost model: %s\n", g_mlp_err);
    return st;
}
#define MCHK(call)                                              \
    do {                                                        \
        hipError_t e_ = (call);                                 \
        if (e_ != hipSuccess) return mlp_err(SBC_ERR_HIP, #call, e_); \
    } while (0)

extern "C" int32_t sbcCostModelNumParams(void) { return NP; }

extern "C" sbcStatus sbcCostModelCreate(uint64_t seed, sbcCostModel** out)
{
    if (!out) return SBC_ERR_ARG;
    sbcCostModel* m = new sbcCostModel();
    MCHK(hipMalloc(&m->d_p, sizeof(float) * NP));
    MCHK(hipMalloc(&m->d_m, sizeof(float) * NP));
    MCHK(hipMalloc(&m->d_v, sizeof(float) * NP));
    MCHK(hipMalloc(&m->d_g, sizeof(float) * NP));
    MCHK(hipMalloc(&m->d_w1b, sizeof(unsigned short) * NI * NH));
    MCHK(hipMalloc(&m->d_w2b, sizeof(unsigned short) * NH * NH));
    // He-uniform initial weights from a splitmix64 stream, zero biases
    std::vector<float> p(NP, 0.f);
    uint64_t x = seed;
    auto unif = [&x]() {
        uint64_t z = (x += 0x9E3779B97F4A7C15ull);
        z = (z ^ (z >> 30)) * 0xBF58476D1CE4E5B9ull;
        z = (z ^ (z >> 27)) * 0x94D049BB133111EBull;
        z ^= z >> 31;
        return (double)(z >> 11) * 0x1.0p-53;
    };
    const float l1 = sqrtf(6.f / NF), l2 = sqrtf(6.f / NH);
    for (int i = 0; i < NF; ++i)
        for (int j = 0; j < NH; ++j) p[OFF_W1 + i * NH + j] = (float)((2.0 * unif() - 1.0) * l1);
    for (int q = 0; q < NH * NH; ++q) p[OFF_W2 + q] = (float)((2.0 * unif() - 1.0) * l2);
    for (int k = 0; k < NH; ++k) p[OFF_W3 + k] = (float)((2.0 * unif() - 1.0) * l2);
    MCHK(hipMemcpy(m->d_p, p.data(), sizeof(float) * NP, hipMemcpyHostToDevice));
    MCHK(hipMemset(m->d_m, 0, sizeof(float) * NP));
    MCHK(hipMemset(m->d_v, 0, sizeof(float) * NP));
    *out = m;
    return SBC_OK;
}

extern "C" void sbcCostModelFree(sbcCostModel* m)
{
    if (!m) return;
    (void)hipFree(m->d_p);
    (void)hipFree(m->d_m);
    (void)hipFree(m->d_v);
    (void)hipFree(m->d_g);
    (void)hipFree(m->d_w1b);
    (void)hipFree(m->d_w2b);
    delete m;
}

extern "C" sbcStatus sbcCostModelGetParams(sbcCostModel* m, float* params)
{
    if (!m || !params) return SBC_ERR_ARG;
    MCHK(hipMemcpy(params, m->d_p, sizeof(float) * NP, hipMemcpyDeviceToHost));
    return SBC_OK;
}

extern "C" sbcStatus sbcCostModelSetParams(sbcCostModel* m, const float* params)
{
    if (!m || !params) return SBC_ERR_ARG;
    MCHK(hipMemcpy(m->d_p, params, sizeof(float) * NP, hipMemcpyHostToDevice));
    m->bf16_stale = true;
    return SBC_OK;
}

// activations and gradients of one batch
struct Acts {
    float *X = nullptr, *y = nullptr, *z1 = nullptr, *z2 = nullptr, *out = nullptr, *dz1 = nullptr, *dz2 = nullptr,
          *dout = nullptr;
    int cap = 0;
    ~Acts()
    {
        for (float* q : {X, y, z1, z2, out, dz1, dz2, dout})
            if (q) (void)hipFree(q);
    }
    hipError_t alloc(int n)
    {
        hipError_t e = hipSuccess;
        cap = n;
        if ((e = hipMalloc(&X, sizeof(float) * n * NF))) return e;
        if ((e = hipMalloc(&y, sizeof(float) * n))) return e;
        if ((e = hipMalloc(&z1, sizeof(float) * n * NH))) return e;
        if ((e = hipMalloc(&z2, sizeof(float) * n * NH))) return e;
        if ((e = hipMalloc(&out, sizeof(float) * n))) return e;
        if ((e = hipMalloc(&dz1, sizeof(float) * n * NH))) return e;
        if ((e = hipMalloc(&dz2, sizeof(float) * n * NH))) return e;
        return hipMalloc(&dout, sizeof(float) * n);
    }
};

static hipError_t grad_batch(sbcCostModel* m, Acts& A, int n)
{
    const int bs = 128;
    const int gb = (n + bs - 1) / bs;
    hipLaunchKernelGGL(mlp_fwd_kernel, dim3(gb), dim3(bs), 0, 0, A.X, n, m->d_p, A.z1, A.z2, A.out);
    hipLaunchKernelGGL(mlp_bwd_kernel, dim3(gb), dim3(bs), 0, 0, A.y, n, 1.f / (float)n, m->d_p, A.z1, A.z2, A.out,
                       A.dz1, A.dz2, A.dout);
    hipLaunchKernelGGL(mlp_wgrad_kernel, dim3((NP + 255) / 256), dim3(256), 0, 0, A.X, n, A.z1, A.z2, A.dz1, A.dz2,
                       A.dout, m->d_g);
    return hipGetLastError();
}

static sbcStatus mse(sbcCostModel* m, const float* dX, const float* dy, int64_t n, float* loss)
{
    float* d_out = nullptr;
    MCHK(hipMalloc(&d_out, sizeof(float) * n));
    hipLaunchKernelGGL(mlp_fp32_kernel, dim3((unsigned)((n + 255) / 256)), dim3(256), 0, 0, dX, (int)n, m->d_p, d_out);
    std::vector<float> o(n), yy(n);
    MCHK(hipMemcpy(o.data(), d_out, sizeof(float) * n, hipMemcpyDeviceToHost));
    MCHK(hipMemcpy(yy.data(), dy, sizeof(float) * n, hipMemcpyDeviceToHost));
    MCHK(hipFree(d_out));
    double acc = 0;
    for (int64_t i = 0; i < n; ++i) acc += (double)(o[i] - yy[i]) * (o[i] - yy[i]);
    *loss = (float)(acc / (double)(n > 0 ? n : 1));
    return SBC_OK;
}

extern "C" sbcStatus sbcCostModelGradients(sbcCostModel* m, const float* X, const float* y, int64_t n, float* grad,
                                           float* loss)
{
    if (!m || !X || !y || n <= 0 || n > (1 << 20)) return SBC_ERR_ARG;
    Acts A;
    MCHK(A.alloc((int)n));
    MCHK(hipMemcpy(A.X, X, sizeof(float) * n * NF, hipMemcpyHostToDevice));
    MCHK(hipMemcpy(A.y, y, sizeof(float) * n, hipMemcpyHostToDevice));
    MCHK(grad_batch(m, A, (int)n));
    MCHK(hipDeviceSynchronize());
    if (grad) MCHK(hipMemcpy(grad, m->d_g, sizeof(float) * NP, hipMemcpyDeviceToHost));
    if (loss) {
        sbcStatus st = mse(m, A.X, A.y, n, loss);
        if (st) return st;
    }
    return SBC_OK;
}

extern "C" sbcStatus sbcCostModelTrain(sbcCostModel* m, const float* X, const float* y, int64_t n, int32_t epochs,
                                       int32_t batch, float lr, uint64_t seed, float* final_loss)
{
    if (!m || !X || !y || n <= 0 || epochs < 0 || batch <= 0 || batch > (1 << 16)) return SBC_ERR_ARG;
    float *dX = nullptr, *dy = nullptr;
    int32_t* didx = nullptr;
    MCHK(hipMalloc(&dX, sizeof(float) * n * NF));
    MCHK(hipMalloc(&dy, sizeof(float) * n));
    MCHK(hipMalloc(&didx, sizeof(int32_t) * n));
    MCHK(hipMemcpy(dX, X, sizeof(float) * n * NF, hipMemcpyHostToDevice));
    MCHK(hipMemcpy(dy, y, sizeof(float) * n, hipMemcpyHostToDevice));
    Acts A;
    MCHK(A.alloc(batch));
    std::vector<int32_t> perm(n);
    uint64_t x = seed;
    auto rnd = [&x]() {
        uint64_t z = (x += 0x9E3779B97F4A7C15ull);
        z = (z ^ (z >> 30)) * 0xBF58476D1CE4E5B9ull;
        z = (z ^ (z >> 27)) * 0x94D049BB133111EBull;
        return z ^ (z >> 31);
    };
    const float b1 = 0.9f, b2 = 0.999f, eps = 1e-8f;
    for (int ep = 0; ep < epochs; ++ep) {
        for (int64_t i = 0; i < n; ++i) perm[i] = (int32_t)i;
        for (int64_t i = n - 1; i > 0; --i) {
            const int64_t j = (int64_t)(rnd() % (uint64_t)(i + 1));
            std::swap(perm[i], perm[j]);
        }
        MCHK(hipMemcpy(didx, perm.data(), sizeof(int32_t) * n, hipMemcpyHostToDevice));
        for (int64_t b0 = 0; b0 < n; b0 += batch) {
            const int nb = (int)((n - b0) < batch ? (n - b0) : batch);
            hipLaunchKernelGGL(mlp_gather_kernel, dim3((nb + 255) / 256), dim3(256), 0, 0, dX, dy, didx + b0, nb, A.X,
                               A.y);
            MCHK(grad_batch(m, A, nb));
            m->t++;
            const float c1 = 1.f - powf(b1, (float)m->t), c2 = 1.f - powf(b2, (float)m->t);
            hipLaunchKernelGGL(mlp_adam_kernel, dim3((NP + 255) / 256), dim3(256), 0, 0, m->d_p, m->d_g, m->d_m,
                               m->d_v, lr, b1, b2, eps, c1, c2);
        }
    }
    MCHK(hipDeviceSynchronize());
    m->bf16_stale = true;
    sbcStatus st = SBC_OK;
    if (final_loss) st = mse(m, dX, dy, n, final_loss);
    (void)hipFree(dX);
    (void)hipFree(dy);
    (void)hipFree(didx);
    return st;
}

static sbcStatus predict(sbcCostModel* m, const float* X, int64_t n, float* out, bool mfma)
{
    if (!m || !X || !out || n < 0) return SBC_ERR_ARG;
    if (n == 0) return SBC_OK;
    float *dX = nullptr, *dout = nullptr;
    MCHK(hipMalloc(&dX, sizeof(float) * n * NF));
    MCHK(hipMalloc(&dout, sizeof(float) * n));
    MCHK(hipMemcpy(dX, X, sizeof(float) * n * NF, hipMemcpyHostToDevice));
    if (mfma) {
        if (m->bf16_stale) {
            hipLaunchKernelGGL(mlp_tobf16_kernel, dim3((NH * NH + 255) / 256), dim3(256), 0, 0, m->d_p, m->d_w1b,
                               m->d_w2b);
            m->bf16_stale = false;
        }
        hipLaunchKernelGGL(mlp_mfma_kernel, dim3((unsigned)((n + 63) / 64)), dim3(256), 0, 0, dX, (int)n, m->d_w1b,
                           m->d_w2b, m->d_p, dout);
    } else {
        hipLaunchKernelGGL(mlp_fp32_kernel, dim3((unsigned)((n + 255) / 256)), dim3(256), 0, 0, dX, (int)n, m->d_p,
                           dout);
    }
    MCHK(hipGetLastError());
    MCHK(hipMemcpy(out, dout, sizeof(float) * n, hipMemcpyDeviceToHost));
    (void)hipFree(dX);
    (void)hipFree(dout);
    return SBC_OK;
}

// ---------------------------------------------------------------- schedule

namespace {

// cost[r] = sum over the window's items of expm1(max(pred, 0)) + overhead
__global__ void root_cost_kernel(const float* __restrict__ pred, int n, int K, float overhead, float* cost)
{
    const int r = blockIdx.x * blockDim.x + threadIdx.x;
    if (r >= K) return;
    float c = 0.f;
    for (int i = 0; i < n; ++i) c += expm1f(fmaxf(pred[(size_t)i * K + r], 0.f)) + overhead;
    cost[r] = c;
}

__device__ inline uint32_t orderable(float x)
{
    uint32_t u = __float_as_uint(x);
    return (u & 0x80000000u) ? ~u : (u | 0x80000000u);
}

// One workgroup: roots sorted by descending cost (ties: lower index first)
// into order[K]; heavy = roots with cost >= light_max.
constexpr int SORT_CAP = 4096;
__global__ __launch_bounds__(1024) void root_order_kernel(const float* __restrict__ cost, int K, float light_max,
                                                          int32_t* order, int32_t* heavy)
{
    __shared__ unsigned long long key[SORT_CAP];
    __shared__ int nheavy;
    int n2 = 1;
    while (n2 < K) n2 <<= 1;
    if (threadIdx.x == 0) nheavy = 0;
    for (int i = threadIdx.x; i < n2; i += blockDim.x) {
        if (i < K) {
            // ascending key = descending cost
            key[i] = ((unsigned long long)orderable(-cost[i]) << 32) | (uint32_t)i;
        } else {
            key[i] = ~0ull;
        }
    }
    __syncthreads();
    for (int k2 = 2; k2 <= n2; k2 <<= 1) {
        for (int j = k2 >> 1; j > 0; j >>= 1) {
            for (int i = threadIdx.x; i < n2; i += blockDim.x) {
                const int p = i ^ j;
                if (p > i) {
                    const bool up = (i & k2) == 0;
                    const unsigned long long a = key[i], b = key[p];
                    if ((a > b) == up) {
                        key[i] = b;
                        key[p] = a;
                    }
                }
            }
            __syncthreads();
        }
    }
    for (int i = threadIdx.x; i < K; i += blockDim.x) {
        const int r = (int)(uint32_t)key[i];
        order[i] = r;
        if (cost[r] >= light_max) atomicAdd(&nheavy, 1);
    }
    __syncthreads();
    if (threadIdx.x == 0) *heavy = nheavy;
}

}  // namespace

extern "C" hipError_t sbc_cost_schedule(sbcCostModel* m, const float* dX, int n, int K, float light_max,
                                        float overhead, float* d_pred, float* d_cost, int32_t* d_order,
                                        int32_t* d_heavy, hipStream_t st)
{
    if (!m || K <= 0 || K > SORT_CAP || n <= 0) return hipErrorInvalidValue;
    if (m->bf16_stale) {
        hipLaunchKernelGGL(mlp_tobf16_kernel, dim3((NH * NH + 255) / 256), dim3(256), 0, st, m->d_p, m->d_w1b,
                           m->d_w2b);
        m->bf16_stale = false;
    }
    const int64_t items = (int64_t)n * K;
    hipLaunchKernelGGL(mlp_mfma_kernel, dim3((unsigned)((items + 63) / 64)), dim3(256), 0, st, dX, (int)items,
                       m->d_w1b, m->d_w2b, m->d_p, d_pred);
    hipLaunchKernelGGL(root_cost_kernel, dim3((K + 255) / 256), dim3(256), 0, st, d_pred, n, K, overhead, d_cost);
    hipLaunchKernelGGL(root_order_kernel, dim3(1), dim3(1024), 0, st, d_cost, K, light_max, d_order, d_heavy);
    return hipGetLastError();
}

extern "C" sbcStatus sbcCostModelPredict(sbcCostModel* m, const float* X, int64_t n, float* out)
{
    return predict(m, X, n, out, true);
}

extern "C" sbcStatus sbcCostModelPredictFP32(sbcCostModel* m, const float* X, int64_t n, float* out)
{
    return predict(m, X, n, out, false);
}
