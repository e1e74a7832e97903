// Surrogate network for cut pairs: a 12 -> 64 -> 64 -> 64 -> 9 perceptron with SiLU
// activations, evaluated with BF16 MFMA (v_mfma_f32_16x16x16_bf16, FP32 accumulation).
//
// Orientation.  Every layer computes Z = W H + b with the features along the MFMA row
// dimension and 16 samples along the column dimension: the weight fragment is the A
// operand, the activations are the B operand.  The 16x16 FP32 result of an MFMA leaves
// lane l holding sample (l % 16) and features 4 (l / 16) + {0..3} of its 16-feature tile,
// which is exactly the B-operand layout of the next layer's 16-feature k-chunk, so a
// forward pass never moves data between lanes.
//
// Parameter vector (FP32 master copy, row-major W[out][in]):
//   W1[64][12] b1[64] W2[64][64] b2[64] W3[64][64] b3[64] W4[9][64] b4[9]   (9737 floats)
// Packed BF16 fragments, 64 lanes x 4 values each (512 B per fragment):
//   forward  (A = W):   L1 4 (4 tiles x 1 chunk, inputs padded to 16), L2 16, L3 16,
//                       L4 4 (1 tile, outputs padded to 16, x 4 chunks)          = 40
//   backward (A = W^T): L2 16, L3 16, L4 4 (4 tiles over inputs x 1 chunk)       = 36
// Fragment (tile t, chunk c) of a matrix M with NC chunks sits at index t*NC + c; lane l
// holds M[16 t + l % 16][16 c + 4 (l / 16) + i], i = 0..3.  Padding entries are zero.
#pragma once
#include <hip/hip_runtime.h>
#include <stdint.h>

namespace nlfem {
namespace mlp {

constexpr int NIN = 12, NH = 64, NOUT = 9;
constexpr int OFF_W1 = 0, OFF_B1 = OFF_W1 + NH * NIN, OFF_W2 = OFF_B1 + NH, OFF_B2 = OFF_W2 + NH * NH,
              OFF_W3 = OFF_B2 + NH, OFF_B3 = OFF_W3 + NH * NH, OFF_W4 = OFF_B3 + NH, OFF_B4 = OFF_W4 + NOUT * NH,
              NPARAM = OFF_B4 + NOUT;
constexpr int FWD_L1 = 0, FWD_L2 = 4, FWD_L3 = 20, FWD_L4 = 36, NFRAG_FWD = 40;
constexpr int BWD_L2 = 0, BWD_L3 = 16, BWD_L4 = 32, NFRAG_BWD = 36;
// normalisation block: input mean[12], 1/std[12], output mean[9], std[9]
constexpr int NORM_MU = 0, NORM_ISD = 12, NORM_OMU = 24, NORM_OSD = 33, NNORM = 42;
// biases padded per layer: b1[64] b2[64] b3[64] b4[16]
constexpr int BIAS_PAD = 3 * NH + 16;

typedef short bf16x4 __attribute__((ext_vector_type(4)));
typedef float f32x4 __attribute__((ext_vector_type(4)));

// FP32 -> BF16, round to nearest even (finite inputs).
__host__ __device__ inline uint16_t to_bf16(float f) {
    union { float f; uint32_t u; } v;
    v.f = f;
    const uint32_t r = v.u + 0x7fffu + ((v.u >> 16) & 1u);
    return (uint16_t)(r >> 16);
}
__host__ __device__ inline float from_bf16(uint16_t h) {
    union { float f; uint32_t u; } v;
    v.u = (uint32_t)h << 16;
    return v.f;
}

// Position of weight (o, k) of a layer (nin inputs, nc = input chunks) in the forward pack,
// relative to the layer's first fragment, in units of bf16 values.
__host__ __device__ inline int fwd_pos(int o, int k, int nc) {
    const int t = o >> 4, r = o & 15, c = k >> 4, kk = k & 15;
    return (((t * nc + c) * 64) + r + 16 * (kk >> 2)) * 4 + (kk & 3);
}
// Position of weight (o, k) in the transposed pack (rows = inputs, k = outputs; nco = output chunks).
__host__ __device__ inline int bwd_pos(int o, int k, int nco) { return fwd_pos(k, o, nco); }

// Where parameter p lives in the packs: returns the bias index (>= 0) or -1 for a weight,
// and the forward / backward pack offsets (bwd -1 when the layer has no transposed pack).
__host__ __device__ inline int param_place(int p, int& fwd, int& bwd) {
    fwd = -1; bwd = -1;
    if (p < OFF_B1) { int o = p / NIN, k = p % NIN; fwd = FWD_L1 * 256 + fwd_pos(o, k, 1); return -1; }
    if (p < OFF_W2) return p - OFF_B1;
    if (p < OFF_B2) { int q = p - OFF_W2, o = q / NH, k = q % NH;
        fwd = FWD_L2 * 256 + fwd_pos(o, k, 4); bwd = BWD_L2 * 256 + bwd_pos(o, k, 4); return -1; }
    if (p < OFF_W3) return NH + (p - OFF_B2);
    if (p < OFF_B3) { int q = p - OFF_W3, o = q / NH, k = q % NH;
        fwd = FWD_L3 * 256 + fwd_pos(o, k, 4); bwd = BWD_L3 * 256 + bwd_pos(o, k, 4); return -1; }
    if (p < OFF_W4) return 2 * NH + (p - OFF_B3);
    if (p < OFF_B4) { int q = p - OFF_W4, o = q / NH, k = q % NH;
        fwd = FWD_L4 * 256 + fwd_pos(o, k, 4); bwd = BWD_L4 * 256 + bwd_pos(o, k, 1); return -1; }
    return 3 * NH + (p - OFF_B4);
}

#if defined(__HIP_DEVICE_COMPILE__)
__device__ inline f32x4 mfma(bf16x4 a, bf16x4 b, f32x4 c) {
    return __builtin_amdgcn_mfma_f32_16x16x16bf16_1k(a, b, c, 0, 0, 0);
}
#else
__device__ inline f32x4 mfma(bf16x4, bf16x4, f32x4 c) { return c; }
#endif

// SiLU with the hardware exponential and reciprocal (v_exp_f32, v_rcp_f32).
__device__ inline float sigmoid(float z) { return __builtin_amdgcn_rcpf(1.0f + __expf(-z)); }
__device__ inline float silu(float z) { return z * sigmoid(z); }
__device__ inline float dsilu(float z) {
    const float s = sigmoid(z);
    return s * (1.0f + z * (1.0f - s));
}
// FP32 -> BF16 with round to nearest even, two values per v_cvt_pk_bf16_f32.
__device__ inline short bf16_bits(float f) {
    const __bf16 h = (__bf16)f;
    return __builtin_bit_cast(short, h);
}
__device__ inline bf16x4 pack4(f32x4 v) {
    bf16x4 r;
    r[0] = bf16_bits(v[0]); r[1] = bf16_bits(v[1]);
    r[2] = bf16_bits(v[2]); r[3] = bf16_bits(v[3]);
    return r;
}

// Z[t] = b + sum_c W[t][c] H[c] for NT output tiles and NC input chunks.
template <int NT, int NC>
__device__ inline void layer(const bf16x4* __restrict__ wf, const float* __restrict__ b, const bf16x4 (&h)[NC],
                             f32x4 (&z)[NT], int lane) {
    const int q = 4 * (lane >> 4);
#pragma unroll
    for (int t = 0; t < NT; t++) {
        f32x4 acc = {b[16 * t + q], b[16 * t + q + 1], b[16 * t + q + 2], b[16 * t + q + 3]};
#pragma unroll
        for (int c = 0; c < NC; c++) acc = mfma(wf[(t * NC + c) * 64 + lane], h[c], acc);
        z[t] = acc;
    }
}

template <int NT>
__device__ inline void activate(const f32x4 (&z)[NT], bf16x4 (&h)[NT]) {
#pragma unroll
    for (int t = 0; t < NT; t++) {
        f32x4 a = {silu(z[t][0]), silu(z[t][1]), silu(z[t][2]), silu(z[t][3])};
        h[t] = pack4(a);
    }
}

// Forward pass of one 16-sample tile held by one wavefront.  x: this lane's four normalised
// inputs (sample lane % 16, inputs 4 (lane / 16) + i; lanes 48..63 carry the zero padding).
// Returns outputs 4 (lane / 16) + i of sample lane % 16 in normalised units.
__device__ inline f32x4 forward_tile(const bf16x4* __restrict__ wf, const float* __restrict__ bias, f32x4 x,
                                     int lane) {
    bf16x4 h0[1] = {pack4(x)};
    f32x4 z1[4], z2[4], z3[4], y[1];
    bf16x4 h1[4], h2[4], h3[4];
    layer<4, 1>(wf + FWD_L1 * 64, bias, h0, z1, lane);
    activate<4>(z1, h1);
    layer<4, 4>(wf + FWD_L2 * 64, bias + NH, h1, z2, lane);
    activate<4>(z2, h2);
    layer<4, 4>(wf + FWD_L3 * 64, bias + 2 * NH, h2, z3, lane);
    activate<4>(z3, h3);
    layer<1, 4>(wf + FWD_L4 * 64, bias + 3 * NH, h3, y, lane);
    return y[0];
}

}  // namespace mlp
}  // namespace nlfem
