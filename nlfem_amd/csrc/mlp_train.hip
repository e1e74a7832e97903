// Training and stand-alone inference of the cut-pair surrogate network (mlp.hpp) with
// hand-written BF16 MFMA kernels: fused forward + backward per 16-sample tile, per-wave
// gradient partials, and an Adam step that also refreshes the packed BF16 fragments.
#include <hip/hip_runtime.h>

#include <algorithm>
#include <mutex>
#include <cmath>
#include <numeric>
#include <random>
#include <stdexcept>
#include <string>
#include <vector>

#include "mlp.hpp"

namespace nlfem {
namespace mlp {

#define HIPCHECK_MLP(x)                                                                                      \
    do {                                                                                                     \
        hipError_t e_ = (x);                                                                                 \
        if (e_ != hipSuccess)                                                                                \
            throw std::runtime_error(std::string("HIP error ") + hipGetErrorString(e_) + " (" #x ")");        \
    } while (0)

struct GradArgs {
    const float* X;     // [n][12] raw inputs
    const float* T;     // [n][9] raw targets
    const int* idx;     // batch slot -> sample
    long long n;        // slots in this launch (the rest of the grid is masked)
    const float* norm;  // NNORM
    const bf16x4* wf;
    const bf16x4* wb;
    const float* bias;
    float* gpart;       // [4 * gridDim.x][NPARAM]
    float* loss;        // += sum of squared normalised errors * scale
    float scale;        // 1 / (batch * 9): gradient of the mean squared error
};

__device__ inline float wave_sum16(float v) {  // sum over the 16 lanes sharing lane / 16
    v += __shfl_xor(v, 1, 64);
    v += __shfl_xor(v, 2, 64);
    v += __shfl_xor(v, 4, 64);
    v += __shfl_xor(v, 8, 64);
    return v;
}

// One wavefront per tile of 16 samples.  LDS keeps the activations and the current layer's
// output gradient as [feature][sample] BF16 so that the weight-gradient products (which
// reduce over samples) read their operands with one 8-byte load per lane.
__global__ __launch_bounds__(256) void k_mlp_grad(GradArgs a) {
    __shared__ uint16_t hT[4][3][NH][16];
    __shared__ uint16_t h0T[4][16][16];
    __shared__ uint16_t dT[4][NH][16];
    const int lane = threadIdx.x & 63, w = threadIdx.x >> 6;
    const int j = lane & 15, q = 4 * (lane >> 4);
    const long long slot = (long long)blockIdx.x * 64 + w * 16 + j;
    const bool valid = slot < a.n;
    const long long s = valid ? a.idx[slot] : 0;
    float* gp = a.gpart + ((size_t)blockIdx.x * 4 + w) * NPARAM;

    f32x4 x = {0, 0, 0, 0};
    if (valid && q < NIN) {
#pragma unroll
        for (int i = 0; i < 4; i++) x[i] = (a.X[s * NIN + q + i] - a.norm[NORM_MU + q + i]) * a.norm[NORM_ISD + q + i];
    }
    bf16x4 h0[1] = {pack4(x)};
    f32x4 z1[4], z2[4], z3[4], y[1];
    bf16x4 h1[4], h2[4], h3[4];
    layer<4, 1>(a.wf + FWD_L1 * 64, a.bias, h0, z1, lane);
    activate<4>(z1, h1);
    layer<4, 4>(a.wf + FWD_L2 * 64, a.bias + NH, h1, z2, lane);
    activate<4>(z2, h2);
    layer<4, 4>(a.wf + FWD_L3 * 64, a.bias + 2 * NH, h2, z3, lane);
    activate<4>(z3, h3);
    layer<1, 4>(a.wf + FWD_L4 * 64, a.bias + 3 * NH, h3, y, lane);
#pragma unroll
    for (int i = 0; i < 4; i++) {
        h0T[w][q + i][j] = (uint16_t)h0[0][i];
#pragma unroll
        for (int t = 0; t < 4; t++) {
            hT[w][0][16 * t + q + i][j] = (uint16_t)h1[t][i];
            hT[w][1][16 * t + q + i][j] = (uint16_t)h2[t][i];
            hT[w][2][16 * t + q + i][j] = (uint16_t)h3[t][i];
        }
    }
    // output gradient of the mean squared error in normalised units
    f32x4 dy = {0, 0, 0, 0};
    float lsum = 0;
#pragma unroll
    for (int i = 0; i < 4; i++) {
        const int o = q + i;
        if (valid && o < NOUT) {
            const float t = (a.T[s * NOUT + o] - a.norm[NORM_OMU + o]) / a.norm[NORM_OSD + o];
            const float d = y[0][i] - t;
            lsum += d * d;
            dy[i] = 2.0f * d * a.scale;
        }
    }
    for (int m = 32; m > 0; m >>= 1) lsum += __shfl_xor(lsum, m, 64);
    if (lane == 0) atomicAdd(a.loss, lsum * a.scale);
    const bf16x4 dyb = pack4(dy);
#pragma unroll
    for (int i = 0; i < 4; i++) dT[w][q + i][j] = (uint16_t)dyb[i];
    __syncthreads();

    const f32x4 zero = {0, 0, 0, 0};
    auto lds4 = [&](const uint16_t* row) {  // four consecutive samples q..q+3 of one feature row
        bf16x4 r;
        r[0] = (short)row[q]; r[1] = (short)row[q + 1]; r[2] = (short)row[q + 2]; r[3] = (short)row[q + 3];
        return r;
    };
    // weight tile (A = dZ rows, B = H rows), one k-step of 16 samples; lane holds
    // dW[16 t + q + i][16 c + j]
    auto store_w = [&](f32x4 g, int off, int nin, int t, int c, int nout) {
#pragma unroll
        for (int i = 0; i < 4; i++) {
            const int o = 16 * t + q + i, k = 16 * c + j;
            if (o < nout && k < nin) gp[off + o * nin + k] = g[i];
        }
    };
    auto store_b = [&](const f32x4* dz, int nt, int off, int nout) {
        for (int t = 0; t < nt; t++)
#pragma unroll
            for (int i = 0; i < 4; i++) {
                const float v = wave_sum16(dz[t][i]);
                const int o = 16 * t + q + i;
                if (j == 0 && o < nout) gp[off + o] = v;
            }
    };

    // layer 4
    {
        const bf16x4 A = lds4(&dT[w][j][0]);
#pragma unroll
        for (int c = 0; c < 4; c++) store_w(mfma(A, lds4(&hT[w][2][16 * c + j][0]), zero), OFF_W4, NH, 0, c, NOUT);
        store_b(&dy, 1, OFF_B4, NOUT);
    }
    f32x4 dz3[4], dz2[4], dz1[4];
    bf16x4 g3[4], g2[4];
    {
        bf16x4 B[1] = {dyb};
#pragma unroll
        for (int t = 0; t < 4; t++) {
            f32x4 acc = mfma(a.wb[(BWD_L4 + t) * 64 + lane], B[0], zero);
#pragma unroll
            for (int i = 0; i < 4; i++) acc[i] *= dsilu(z3[t][i]);
            dz3[t] = acc;
            g3[t] = pack4(acc);
        }
    }
    __syncthreads();
#pragma unroll
    for (int t = 0; t < 4; t++)
#pragma unroll
        for (int i = 0; i < 4; i++) dT[w][16 * t + q + i][j] = (uint16_t)g3[t][i];
    __syncthreads();
    // layer 3
#pragma unroll
    for (int t = 0; t < 4; t++) {
        const bf16x4 A = lds4(&dT[w][16 * t + j][0]);
#pragma unroll
        for (int c = 0; c < 4; c++) store_w(mfma(A, lds4(&hT[w][1][16 * c + j][0]), zero), OFF_W3, NH, t, c, NH);
    }
    store_b(dz3, 4, OFF_B3, NH);
#pragma unroll
    for (int t = 0; t < 4; t++) {
        f32x4 acc = zero;
#pragma unroll
        for (int c = 0; c < 4; c++) acc = mfma(a.wb[(BWD_L3 + t * 4 + c) * 64 + lane], g3[c], acc);
#pragma unroll
        for (int i = 0; i < 4; i++) acc[i] *= dsilu(z2[t][i]);
        dz2[t] = acc;
        g2[t] = pack4(acc);
    }
    __syncthreads();
#pragma unroll
    for (int t = 0; t < 4; t++)
#pragma unroll
        for (int i = 0; i < 4; i++) dT[w][16 * t + q + i][j] = (uint16_t)g2[t][i];
    __syncthreads();
    // layer 2
#pragma unroll
    for (int t = 0; t < 4; t++) {
        const bf16x4 A = lds4(&dT[w][16 * t + j][0]);
#pragma unroll
        for (int c = 0; c < 4; c++) store_w(mfma(A, lds4(&hT[w][0][16 * c + j][0]), zero), OFF_W2, NH, t, c, NH);
    }
    store_b(dz2, 4, OFF_B2, NH);
    bf16x4 g1[4];
#pragma unroll
    for (int t = 0; t < 4; t++) {
        f32x4 acc = zero;
#pragma unroll
        for (int c = 0; c < 4; c++) acc = mfma(a.wb[(BWD_L2 + t * 4 + c) * 64 + lane], g2[c], acc);
#pragma unroll
        for (int i = 0; i < 4; i++) acc[i] *= dsilu(z1[t][i]);
        dz1[t] = acc;
        g1[t] = pack4(acc);
    }
    __syncthreads();
#pragma unroll
    for (int t = 0; t < 4; t++)
#pragma unroll
        for (int i = 0; i < 4; i++) dT[w][16 * t + q + i][j] = (uint16_t)g1[t][i];
    __syncthreads();
    // layer 1
#pragma unroll
    for (int t = 0; t < 4; t++) {
        const bf16x4 A = lds4(&dT[w][16 * t + j][0]);
        store_w(mfma(A, lds4(&h0T[w][j][0]), zero), OFF_W1, NIN, t, 0, NH);
    }
    store_b(dz1, 4, OFF_B1, NH);
}

// Sum of the per-wave partials (and, with lr > 0, an Adam step).  One thread per parameter.
__global__ void k_mlp_update(int nw, const float* __restrict__ gpart, float* w, float* m, float* v, float* gout,
                             float lr, float b1, float b2, float eps, float bc1, float bc2, uint16_t* wf, uint16_t* wb,
                             float* bias) {
    const int p = blockIdx.x * blockDim.x + threadIdx.x;
    if (p >= NPARAM) return;
    float g = 0;
    for (int i = 0; i < nw; i++) g += gpart[(size_t)i * NPARAM + p];
    if (gout) gout[p] = g;
    if (lr > 0) {
        const float mm = b1 * m[p] + (1 - b1) * g;
        const float vv = b2 * v[p] + (1 - b2) * g * g;
        m[p] = mm;
        v[p] = vv;
        w[p] -= lr * (mm / bc1) / (sqrtf(vv / bc2) + eps);
    }
    int fo, bo;
    const int bi = param_place(p, fo, bo);
    if (bi >= 0) bias[bi] = w[p];
    else {
        const uint16_t hb = to_bf16(w[p]);
        wf[fo] = hb;
        if (bo >= 0) wb[bo] = hb;
    }
}

__global__ __launch_bounds__(256) void k_mlp_forward(const float* X, long long n, const float* norm,
                                                     const bf16x4* wf, const float* bias, float* Y) {
    const int lane = threadIdx.x & 63, w = threadIdx.x >> 6;
    const int j = lane & 15, q = 4 * (lane >> 4);
    const long long s = ((long long)blockIdx.x * 4 + w) * 16 + j;
    const bool valid = s < n;
    f32x4 x = {0, 0, 0, 0};
    if (valid && q < NIN) {
#pragma unroll
        for (int i = 0; i < 4; i++) x[i] = (X[s * NIN + q + i] - norm[NORM_MU + q + i]) * norm[NORM_ISD + q + i];
    }
    const f32x4 y = forward_tile(wf, bias, x, lane);
#pragma unroll
    for (int i = 0; i < 4; i++) {
        const int o = q + i;
        if (valid && o < NOUT) Y[s * NOUT + o] = y[i] * norm[NORM_OSD + o] + norm[NORM_OMU + o];
    }
}

// ---------------------------------------------------------------------------- host side
template <typename Tp>
static Tp* dmalloc(size_t n) {
    Tp* p = nullptr;
    HIPCHECK_MLP(hipMalloc((void**)&p, std::max<size_t>(n, 1) * sizeof(Tp)));
    return p;
}

struct DevNet {
    float *w = nullptr, *m = nullptr, *v = nullptr, *bias = nullptr, *norm = nullptr;
    uint16_t *wf = nullptr, *wb = nullptr;
    void alloc(const float* params, const float* nrm) {
        w = dmalloc<float>(NPARAM);
        m = dmalloc<float>(NPARAM);
        v = dmalloc<float>(NPARAM);
        bias = dmalloc<float>(BIAS_PAD);
        norm = dmalloc<float>(NNORM);
        wf = dmalloc<uint16_t>(NFRAG_FWD * 256);
        wb = dmalloc<uint16_t>(NFRAG_BWD * 256);
        HIPCHECK_MLP(hipMemcpy(w, params, NPARAM * sizeof(float), hipMemcpyHostToDevice));
        HIPCHECK_MLP(hipMemcpy(norm, nrm, NNORM * sizeof(float), hipMemcpyHostToDevice));
        HIPCHECK_MLP(hipMemset(m, 0, NPARAM * sizeof(float)));
        HIPCHECK_MLP(hipMemset(v, 0, NPARAM * sizeof(float)));
        HIPCHECK_MLP(hipMemset(bias, 0, BIAS_PAD * sizeof(float)));
        HIPCHECK_MLP(hipMemset(wf, 0, NFRAG_FWD * 256 * sizeof(uint16_t)));
        HIPCHECK_MLP(hipMemset(wb, 0, NFRAG_BWD * 256 * sizeof(uint16_t)));
    }
    void release() {
        void* ps[] = {w, m, v, bias, norm, wf, wb};
        for (void* p : ps)
            if (p) (void)hipFree(p);
    }
};

static void check_norm(const float* norm) {
    for (int i = 0; i < NNORM; i++)
        if (!std::isfinite(norm[i])) throw std::invalid_argument("surrogate normalisation contains a non-finite value");
    for (int i = 0; i < NOUT; i++)
        if (!(norm[NORM_OSD + i] > 0)) throw std::invalid_argument("surrogate output scale must be positive");
}

// Pack the FP32 parameters into the BF16 fragments used by every kernel.
static void pack(DevNet& d, hipStream_t st) {
    k_mlp_update<<<(NPARAM + 255) / 256, 256, 0, st>>>(0, nullptr, d.w, d.m, d.v, nullptr, 0.f, 0.f, 0.f, 0.f, 1.f,
                                                       1.f, d.wf, d.wb, d.bias);
    HIPCHECK_MLP(hipGetLastError());
}

// Pack one network into caller-owned device buffers (fragments NFRAG_FWD*256 bf16, biases
// BIAS_PAD, normalisation NNORM) on stream st.
void upload(const float* params, const float* norm, uint16_t* d_wf, float* d_bias, float* d_norm, hipStream_t st) {
    check_norm(norm);
    float* w = dmalloc<float>(NPARAM);
    uint16_t* wb = nullptr;
    try {
        wb = dmalloc<uint16_t>(NFRAG_BWD * 256);
        HIPCHECK_MLP(hipMemcpyAsync(w, params, NPARAM * sizeof(float), hipMemcpyHostToDevice, st));
        HIPCHECK_MLP(hipMemcpyAsync(d_norm, norm, NNORM * sizeof(float), hipMemcpyHostToDevice, st));
        HIPCHECK_MLP(hipMemsetAsync(d_wf, 0, NFRAG_FWD * 256 * sizeof(uint16_t), st));
        HIPCHECK_MLP(hipMemsetAsync(d_bias, 0, BIAS_PAD * sizeof(float), st));
        k_mlp_update<<<(NPARAM + 255) / 256, 256, 0, st>>>(0, nullptr, w, nullptr, nullptr, nullptr, 0.f, 0.f, 0.f, 0.f,
                                                           1.f, 1.f, d_wf, wb, d_bias);
        HIPCHECK_MLP(hipGetLastError());
        HIPCHECK_MLP(hipStreamSynchronize(st));
    } catch (...) {
        (void)hipFree(w); (void)hipFree(wb);
        throw;
    }
    (void)hipFree(w);
    (void)hipFree(wb);
}

void mlp_forward(const float* params, const float* norm, const float* X, long long n, float* Y) {
    check_norm(norm);
    if (n <= 0) return;
    DevNet d;
    float *dX = nullptr, *dY = nullptr;
    try {
        d.alloc(params, norm);
        dX = dmalloc<float>((size_t)n * NIN);
        dY = dmalloc<float>((size_t)n * NOUT);
        HIPCHECK_MLP(hipMemcpy(dX, X, (size_t)n * NIN * sizeof(float), hipMemcpyHostToDevice));
        pack(d, 0);
        const long long tiles = (n + 15) / 16;
        k_mlp_forward<<<(unsigned)((tiles + 3) / 4), 256>>>(dX, n, d.norm, (const bf16x4*)d.wf, d.bias, dY);
        HIPCHECK_MLP(hipGetLastError());
        HIPCHECK_MLP(hipMemcpy(Y, dY, (size_t)n * NOUT * sizeof(float), hipMemcpyDeviceToHost));
    } catch (...) {
        d.release(); (void)hipFree(dX); (void)hipFree(dY);
        throw;
    }
    d.release(); (void)hipFree(dX); (void)hipFree(dY);
}

// ---------------------------------------------------------------------------- allocation-free forward
// nlfem_mlp_forward_noalloc keeps the packed network in this file-scope device array instead of
// hipMalloc'd buffers: forward fragments, the transposed fragments that the packing kernel also
// writes, padded biases.  Every device holds its own copy until the process ends.  In host code
// the array's name is only a host-side stand-in: kernels get the device address that
// hipGetSymbolAddress returns for the current device (scratch_ptr()).
constexpr size_t SCR_WF = 0, SCR_WB = SCR_WF + (size_t)NFRAG_FWD * 256 * sizeof(uint16_t),
                 SCR_BIAS = SCR_WB + (size_t)NFRAG_BWD * 256 * sizeof(uint16_t),
                 SCRATCH_BYTES = SCR_BIAS + (size_t)BIAS_PAD * sizeof(float);
static_assert(SCR_WB % 256 == 0 && SCR_BIAS % 256 == 0 && SCRATCH_BYTES % 4 == 0, "scratch layout");
static __device__ __align__(256) unsigned char scratch[SCRATCH_BYTES];
constexpr long long NOALLOC_MAX_SAMPLES = 2147483647LL;

static std::mutex g_scratch_mu;
static void* g_scratch_dev[64] = {};

// Device address of `scratch` on the current device, looked up once per device.
static unsigned char* scratch_ptr() {
    int dev = -1;
    HIPCHECK_MLP(hipGetDevice(&dev));
    if (dev < 0 || dev >= 64) throw std::runtime_error("device index out of range for the scratch table");
    std::lock_guard<std::mutex> lock(g_scratch_mu);
    if (!g_scratch_dev[dev]) {
        void* p = nullptr;
        HIPCHECK_MLP(hipGetSymbolAddress(&p, HIP_SYMBOL(scratch)));
        if (!p) throw std::runtime_error("hipGetSymbolAddress returned no address for the scratch array");
        g_scratch_dev[dev] = p;
    }
    return (unsigned char*)g_scratch_dev[dev];
}

__global__ void k_zero_words(uint32_t* p, int n) {
    const int i = blockIdx.x * blockDim.x + threadIdx.x;
    if (i < n) p[i] = 0u;
}

void mlp_forward_noalloc(const float* d_params, const float* d_norm, const float* d_X, long long n, float* d_Y,
                         hipStream_t st) {
    if (n < 0 || n > NOALLOC_MAX_SAMPLES)
        throw std::invalid_argument("nlfem_mlp_forward_noalloc: n must lie in [0, 2^31 - 1]");
    if (n == 0) return;
    unsigned char* s = scratch_ptr();
    uint16_t* wf = (uint16_t*)(s + SCR_WF);
    uint16_t* wb = (uint16_t*)(s + SCR_WB);
    float* bias = (float*)(s + SCR_BIAS);
    // the packing kernel writes the weights and biases only: padding entries must read zero
    // (a fill kernel, not a memset node, so that the sequence can be captured in a graph)
    const int words = (int)(SCRATCH_BYTES / 4);
    k_zero_words<<<(words + 255) / 256, 256, 0, st>>>((uint32_t*)s, words);
    k_mlp_update<<<(NPARAM + 255) / 256, 256, 0, st>>>(0, nullptr, const_cast<float*>(d_params), nullptr, nullptr,
                                                       nullptr, 0.f, 0.f, 0.f, 0.f, 1.f, 1.f, wf, wb, bias);
    const long long tiles = (n + 15) / 16;
    k_mlp_forward<<<(unsigned)((tiles + 3) / 4), 256, 0, st>>>(d_X, n, d_norm, (const bf16x4*)wf, bias, d_Y);
    HIPCHECK_MLP(hipGetLastError());
}

// Gradient of the mean squared (normalised) error over all n samples, no update.
void mlp_grad(const float* params, const float* norm, const float* X, const float* T, long long n, float* grad,
              float* loss) {
    check_norm(norm);
    if (n <= 0) throw std::invalid_argument("empty batch");
    DevNet d;
    float *dX = nullptr, *dT = nullptr, *gpart = nullptr, *dg = nullptr, *dl = nullptr;
    int* idx = nullptr;
    auto cleanup = [&]() {
        d.release();
        void* ps[] = {dX, dT, gpart, dg, dl, idx};
        for (void* p : ps)
            if (p) (void)hipFree(p);
    };
    try {
        d.alloc(params, norm);
        const unsigned nblk = (unsigned)((n + 63) / 64);
        dX = dmalloc<float>((size_t)n * NIN);
        dT = dmalloc<float>((size_t)n * NOUT);
        gpart = dmalloc<float>((size_t)nblk * 4 * NPARAM);
        dg = dmalloc<float>(NPARAM);
        dl = dmalloc<float>(1);
        idx = dmalloc<int>((size_t)n);
        std::vector<int> id((size_t)n);
        std::iota(id.begin(), id.end(), 0);
        HIPCHECK_MLP(hipMemcpy(dX, X, (size_t)n * NIN * sizeof(float), hipMemcpyHostToDevice));
        HIPCHECK_MLP(hipMemcpy(dT, T, (size_t)n * NOUT * sizeof(float), hipMemcpyHostToDevice));
        HIPCHECK_MLP(hipMemcpy(idx, id.data(), (size_t)n * sizeof(int), hipMemcpyHostToDevice));
        HIPCHECK_MLP(hipMemset(gpart, 0, (size_t)nblk * 4 * NPARAM * sizeof(float)));
        HIPCHECK_MLP(hipMemset(dl, 0, sizeof(float)));
        pack(d, 0);
        GradArgs a{dX, dT, idx, n, d.norm, (const bf16x4*)d.wf, (const bf16x4*)d.wb, d.bias, gpart, dl,
                   1.0f / (float)(n * NOUT)};
        k_mlp_grad<<<nblk, 256>>>(a);
        HIPCHECK_MLP(hipGetLastError());
        k_mlp_update<<<(NPARAM + 255) / 256, 256>>>((int)nblk * 4, gpart, d.w, d.m, d.v, dg, 0.f, 0.f, 0.f, 0.f, 1.f,
                                                    1.f, d.wf, d.wb, d.bias);
        HIPCHECK_MLP(hipGetLastError());
        HIPCHECK_MLP(hipMemcpy(grad, dg, NPARAM * sizeof(float), hipMemcpyDeviceToHost));
        HIPCHECK_MLP(hipMemcpy(loss, dl, sizeof(float), hipMemcpyDeviceToHost));
    } catch (...) {
        cleanup();
        throw;
    }
    cleanup();
}

// Adam with linear warm-up and cosine decay; minibatches of `batch` samples (a multiple of
// 64) drawn without replacement from a fresh permutation each epoch; the last partial batch
// of an epoch is dropped.  loss_hist[e] is the mean minibatch loss of epoch e.
void mlp_train(float* params, const float* norm, const float* X, const float* T, long long n, int epochs, int batch,
               float lr_max, unsigned seed, float* loss_hist) {
    check_norm(norm);
    if (batch <= 0 || batch % 64) throw std::invalid_argument("batch must be a positive multiple of 64");
    if (n < batch) throw std::invalid_argument("fewer samples than one batch");
    if (epochs <= 0) throw std::invalid_argument("epochs must be positive");
    DevNet d;
    float *dX = nullptr, *dT = nullptr, *gpart = nullptr, *dl = nullptr;
    int* idx = nullptr;
    hipStream_t st = nullptr;
    auto cleanup = [&]() {
        d.release();
        void* ps[] = {dX, dT, gpart, dl, idx};
        for (void* p : ps)
            if (p) (void)hipFree(p);
        if (st) (void)hipStreamDestroy(st);
    };
    try {
        HIPCHECK_MLP(hipStreamCreateWithFlags(&st, hipStreamNonBlocking));
        d.alloc(params, norm);
        const long long nb = n / batch;
        const unsigned nblk = (unsigned)(batch / 64);
        dX = dmalloc<float>((size_t)n * NIN);
        dT = dmalloc<float>((size_t)n * NOUT);
        gpart = dmalloc<float>((size_t)nblk * 4 * NPARAM);
        dl = dmalloc<float>((size_t)epochs);
        idx = dmalloc<int>((size_t)n);
        HIPCHECK_MLP(hipMemcpyAsync(dX, X, (size_t)n * NIN * sizeof(float), hipMemcpyHostToDevice, st));
        HIPCHECK_MLP(hipMemcpyAsync(dT, T, (size_t)n * NOUT * sizeof(float), hipMemcpyHostToDevice, st));
        HIPCHECK_MLP(hipMemsetAsync(dl, 0, (size_t)epochs * sizeof(float), st));
        pack(d, st);
        std::mt19937_64 rng(seed);
        std::vector<int> perm((size_t)n);
        std::iota(perm.begin(), perm.end(), 0);
        const double b1 = 0.9, b2 = 0.999;
        const long long total = nb * epochs, warm = std::max<long long>(1, total / 20);
        long long step = 0;
        for (int e = 0; e < epochs; e++) {
            std::shuffle(perm.begin(), perm.end(), rng);
            HIPCHECK_MLP(hipMemcpyAsync(idx, perm.data(), (size_t)n * sizeof(int), hipMemcpyHostToDevice, st));
            // the permutation buffer is reused next epoch: wait for the copy before rewriting it
            HIPCHECK_MLP(hipStreamSynchronize(st));
            for (long long b = 0; b < nb; b++) {
                step++;
                const double prog = (double)step / (double)total;
                double lr = lr_max * 0.5 * (1.0 + std::cos(M_PI * prog));
                if (step < warm) lr = lr_max * (double)step / (double)warm;
                GradArgs a{dX, dT, idx + b * batch, batch, d.norm, (const bf16x4*)d.wf, (const bf16x4*)d.wb,
                           d.bias, gpart, dl + e, 1.0f / (float)((long long)batch * NOUT)};
                k_mlp_grad<<<nblk, 256, 0, st>>>(a);
                const float bc1 = (float)(1.0 - std::pow(b1, (double)step));
                const float bc2 = (float)(1.0 - std::pow(b2, (double)step));
                k_mlp_update<<<(NPARAM + 255) / 256, 256, 0, st>>>((int)nblk * 4, gpart, d.w, d.m, d.v, nullptr,
                                                                   (float)lr, (float)b1, (float)b2, 1e-8f, bc1, bc2,
                                                                   d.wf, d.wb, d.bias);
            }
            HIPCHECK_MLP(hipGetLastError());
        }
        std::vector<float> lh((size_t)epochs);
        HIPCHECK_MLP(hipMemcpyAsync(lh.data(), dl, (size_t)epochs * sizeof(float), hipMemcpyDeviceToHost, st));
        HIPCHECK_MLP(hipMemcpyAsync(params, d.w, NPARAM * sizeof(float), hipMemcpyDeviceToHost, st));
        HIPCHECK_MLP(hipStreamSynchronize(st));
        for (int e = 0; e < epochs; e++)
            if (loss_hist) loss_hist[e] = lh[e] / (float)nb;
    } catch (...) {
        cleanup();
        throw;
    }
    cleanup();
}

}  // namespace mlp
}  // namespace nlfem
