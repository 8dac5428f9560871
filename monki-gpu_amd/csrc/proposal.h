// Learned scattering-direction proposal: a BF16 MLP (12 inputs, 4 hidden layers of 128 SiLU
// units, 40 outputs) whose outputs parameterize an 8-component von Mises-Fisher mixture over
// the outgoing direction. No reference counterpart: the reference samples directions from the
// phase function only (mc_routines_general_mod.f90:492-932).
//
// The network runs on gfx950's matrix cores with the batch on the N side of
// v_mfma_f32_32x32x16_bf16: D = W (M x K) * H (K x 32 samples), 16 reduction rows per
// instruction. Lane t of the wave holds, for sample column n = t % 32 and half h = t / 32,
//   A, B operands: 8 values of reduction slots 8 h + e, e = 0..7, of k-step s;
//   D (32 x 32):   rows 8 q + 4 h + i of register 4 q + i.
// Reduction slot 8 h + e of k-step s stands for feature row 16 s + 4 h + (e & 3) + 8 (e >> 2).
// With that order, registers 8 j .. 8 j + 7 of the accumulator of M-block mb are exactly the B
// operand of k-step 2 mb + j of the next layer: layers chain in registers, with no transpose.
// Weights are stored pre-swizzled in the same A-operand order (lane t holds
// W[32 mb + t % 32][feature row of slot 8 h + e]) so each operand load is one coalesced 16-byte
// load per lane.
#pragma once

#include <hip/hip_runtime.h>
#include <stdint.h>

namespace monki_net {

constexpr int kInputs = 12;
constexpr int kHidden = 128;
constexpr int kComponents = 8;
constexpr int kOutputs = 5 * kComponents;  // per component: mean direction (3), kappa, logit
constexpr int kLayers = 5;
constexpr float kKappaMin = 0.05f;
constexpr float kKappaMax = 2000.0f;

// True and padded shapes of layer l (rows = outputs, cols = inputs).
__host__ __device__ constexpr int true_rows(int l) { return l == 4 ? kOutputs : kHidden; }
__host__ __device__ constexpr int true_cols(int l) { return l == 0 ? kInputs : kHidden; }
__host__ __device__ constexpr int rows(int l) { return l == 4 ? 64 : kHidden; }
__host__ __device__ constexpr int cols(int l) { return l == 0 ? 16 : kHidden; }

// Master parameters (float32): per layer, W row-major [true_rows][true_cols], then b.
__host__ __device__ constexpr int param_offset(int l) {
  return l == 0 ? 0 : param_offset(l - 1) + true_rows(l - 1) * true_cols(l - 1) + true_rows(l - 1);
}
constexpr int kParams = param_offset(kLayers);

// Packed bf16 operands, in units of 8 bf16 (one lane's operand of one block).
__host__ __device__ constexpr int packed_offset(int l) {
  return l == 0 ? 0 : packed_offset(l - 1) + (rows(l - 1) / 32) * (cols(l - 1) / 16) * 64;
}
constexpr int kPacked = packed_offset(kLayers);
// Transposed operands W^T (for the backward pass) of layers 1..4: M' = cols(l), K' = rows(l).
__host__ __device__ constexpr int packed_t_offset(int l) {
  return l <= 1 ? 0 : packed_t_offset(l - 1) + (cols(l - 1) / 32) * (rows(l - 1) / 16) * 64;
}
constexpr int kPackedT = packed_t_offset(kLayers);
// Padded biases (float32).
__host__ __device__ constexpr int bias_offset(int l) { return l == 0 ? 0 : bias_offset(l - 1) + rows(l - 1); }
constexpr int kBiases = bias_offset(kLayers);

// Feature row of reduction slot `slot` (0..15) of a k-step: 4 h + (e & 3) + 8 (e >> 2).
__host__ __device__ constexpr int slot_row(int slot) { return 4 * (slot >> 3) + (slot & 3) + 8 * ((slot >> 2) & 1); }

typedef short bf16x8 __attribute__((ext_vector_type(8)));  // bf16 bit patterns
typedef __bf16 mfma_bf16x8 __attribute__((ext_vector_type(8)));
typedef float floatx16 __attribute__((ext_vector_type(16)));

// Device view of one network.
struct Net {
  const bf16x8 *packed;    // kPacked operands
  const bf16x8 *packed_t;  // kPackedT operands (training only)
  const float *bias;       // kBiases
};

// Loads through a global (address space 1) pointer: the network's weight and bias pointers are
// global memory, and a global load waits on the vector-memory counter only (a flat load also
// holds the LDS counter).
template <typename T> __device__ inline T load_global(const T *p) {
  return *(const __attribute__((address_space(1))) T *)p;
}

__device__ inline floatx16 mfma(bf16x8 a, bf16x8 b, floatx16 c) {
  return __builtin_amdgcn_mfma_f32_32x32x16_bf16(__builtin_bit_cast(mfma_bf16x8, a),
                                                 __builtin_bit_cast(mfma_bf16x8, b), c, 0, 0, 0);
}

// float32 -> bfloat16, round to nearest even (finite inputs).
__host__ __device__ inline uint16_t bf16_bits(float x) {
  uint32_t u;
  __builtin_memcpy(&u, &x, 4);
  u += 0x7fffu + ((u >> 16) & 1u);
  return uint16_t(u >> 16);
}

__host__ __device__ inline float bf16_value(uint16_t bits) {
  const uint32_t u = uint32_t(bits) << 16;
  float x;
  __builtin_memcpy(&x, &u, 4);
  return x;
}

// Eight accumulator values (registers 8 j .. 8 j + 7) as one operand: gfx950's
// v_cvt_pk_bf16_f32 rounds two values to nearest even per instruction.
__device__ inline bf16x8 pack8(const float (&v)[8]) {
  const mfma_bf16x8 out = {(__bf16)v[0], (__bf16)v[1], (__bf16)v[2], (__bf16)v[3],
                           (__bf16)v[4], (__bf16)v[5], (__bf16)v[6], (__bf16)v[7]};
  return __builtin_bit_cast(bf16x8, out);
}

// SiLU and its derivative with the hardware reciprocal (one instruction, ~1 ulp) instead of an
// IEEE division: the results are rounded to bf16 next.
__device__ inline float sigmoid(float z) { return __builtin_amdgcn_rcpf(1.0f + __expf(-z)); }
__device__ inline float silu(float z) { return z * sigmoid(z); }

__device__ inline float silu_grad(float z) {
  const float s = sigmoid(z);
  return s * (1.0f + z * (1.0f - s));
}

// One dense layer on one 32-sample block: x (Steps k-steps of B operands) -> next-layer B
// operands (SiLU) or raw outputs. Blocks = rows / 32. z_out, when given, receives the
// pre-activations in D order [block][register][lane].
template <int Steps, int Blocks, bool Activate>
__device__ inline void dense(const bf16x8 *__restrict__ w, const float *__restrict__ bias,
                             const bf16x8 (&x)[8], bf16x8 (&y)[8], floatx16 (&raw)[2],
                             float *__restrict__ z_out) {
  const int lane = __lane_id();
  const int h = lane >> 5;
#pragma unroll
  for (int mb = 0; mb < Blocks; ++mb) {
    // The accumulator starts at the biases of its rows.
    floatx16 c = {0, 0, 0, 0, 0, 0, 0, 0, 0, 0, 0, 0, 0, 0, 0, 0};
#pragma unroll
    for (int s = 0; s < Steps; ++s) c = mfma(load_global(w + (mb * Steps + s) * 64 + lane), x[s], c);
#pragma unroll
    for (int v = 0; v < 16; ++v) c[v] += load_global(bias + 32 * mb + 8 * (v >> 2) + 4 * h + (v & 3));
    if (z_out) {
#pragma unroll
      for (int v = 0; v < 16; ++v) z_out[(mb * 16 + v) * 64 + lane] = c[v];
    }
    if (Activate) {
#pragma unroll
      for (int j = 0; j < 2; ++j) {
        float v[8];
#pragma unroll
        for (int e = 0; e < 8; ++e) v[e] = silu(c[8 * j + e]);
        y[2 * mb + j] = pack8(v);
      }
    } else {
      raw[mb] = c;
    }
    // One M-block at a time: keeps the scheduler from hoisting the weight loads of every
    // block ahead, which would need all of them in registers at once.
    __builtin_amdgcn_sched_barrier(0);
  }
}

// Forward pass of one 32-sample block. x: input B operand (k-step 0; rows >= 12 zero).
// raw: the 64 output rows in D order (rows >= 40 are padding). z (optional): pre-activations
// of the four hidden layers, 4 x 4 x 16 x 64 floats.
__device__ inline void forward_block(const Net &net, const bf16x8 (&x)[8], floatx16 (&raw)[2],
                                     float *__restrict__ z) {
  bf16x8 a[8], b[8];
  dense<1, 4, true>(net.packed + packed_offset(0), net.bias + bias_offset(0), x, a, raw,
                    z ? z : nullptr);
  dense<8, 4, true>(net.packed + packed_offset(1), net.bias + bias_offset(1), a, b, raw,
                    z ? z + 1 * 4096 : nullptr);
  dense<8, 4, true>(net.packed + packed_offset(2), net.bias + bias_offset(2), b, a, raw,
                    z ? z + 2 * 4096 : nullptr);
  dense<8, 4, true>(net.packed + packed_offset(3), net.bias + bias_offset(3), a, b, raw,
                    z ? z + 3 * 4096 : nullptr);
  dense<8, 2, false>(net.packed + packed_offset(4), net.bias + bias_offset(4), b, a, raw, nullptr);
}

// Output m of sample column `column` (0..31) of the block, from the D registers of the wave.
// Every lane calls it with the same m.
template <int M> __device__ inline float output_row(const floatx16 (&raw)[2], int column) {
  constexpr int mb = M / 32, q = (M % 32) / 8, h = (M % 8) / 4, i = M % 4;
  return __shfl(raw[mb][4 * q + i], column + 32 * h, 64);
}

// Gather the 40 outputs of sample column `column` of the block into every lane of the wave;
// lanes keep them when `keep`.
template <int M = 0>
__device__ inline void gather_outputs(const floatx16 (&raw)[2], int column, bool keep, float (&out)[kOutputs]) {
  if constexpr (M < kOutputs) {
    const float value = output_row<M>(raw, column);
    if (keep) out[M] = value;
    gather_outputs<M + 1>(raw, column, keep, out);
  }
}

// Lane of the k-th set bit (k counted from 0) of a wave mask, k < popcount(mask).
__device__ inline int nth_set_lane(unsigned long long mask, int k) {
  int lane = 0;
#pragma unroll
  for (int step = 32; step > 0; step >>= 1) {
    const int count = __popcll((mask >> lane) & ((1ull << step) - 1ull));
    if (k >= count) {
      k -= count;
      lane += step;
    }
  }
  return lane;
}

// Input B operand of one block from per-lane features f[12]: sample column n = lane % 32 takes
// the features of lane source(n) (padding columns pass a negative source and get zeros); lane
// half h holds feature rows slot_row(8 h + e), e = 0..7.
__device__ inline void scatter_inputs(const float (&f)[kInputs], int source, bf16x8 (&x)[8]) {
  const bool upper = __lane_id() >= 32;
  const bool valid = source >= 0;
  float v[16];
#pragma unroll
  for (int k = 0; k < kInputs; ++k) {
    const float value = __shfl(f[k], valid ? source : 0, 64);
    v[k] = valid ? value : 0.0f;
  }
#pragma unroll
  for (int k = kInputs; k < 16; ++k) v[k] = 0.0f;
  float e[8];
#pragma unroll
  for (int i = 0; i < 8; ++i) e[i] = upper ? v[slot_row(8 + i)] : v[slot_row(i)];
  x[0] = pack8(e);
}

// ---- von Mises-Fisher mixture -----------------------------------------------------------
// Component k of the 40 outputs: o[5k..5k+2] unnormalized mean direction, o[5k+3] kappa
// (softplus, clamped to [kKappaMin, kKappaMax]), o[5k+4] mixture logit (softmax).
struct Mixture {
  float mx[kComponents], my[kComponents], mz[kComponents];
  float kappa[kComponents], weight[kComponents];
  // weight x vMF normalization kappa / (2 pi (1 - exp(-2 kappa))), so that the density is
  // sum_k scale_k exp(kappa_k (mu_k . w - 1)).
  float scale[kComponents];
};

__host__ __device__ inline float softplus(float x) { return x > 20.0f ? x : log1pf(expf(x)); }

// The mixture of one event's 40 outputs, with the hardware exp/log/reciprocal (the proposal
// only has to be one density, sampled and evaluated consistently; its form is free).
__device__ inline Mixture mixture_from(const float *o) {
  Mixture m;
  float top = -INFINITY;
#pragma unroll
  for (int k = 0; k < kComponents; ++k) top = fmaxf(top, o[5 * k + 4]);
  float total = 0.0f;
#pragma unroll
  for (int k = 0; k < kComponents; ++k) {
    const float x = o[5 * k], y = o[5 * k + 1], z = o[5 * k + 2];
    const float square = x * x + y * y + z * z;
    const bool ok = square > 1e-12f;
    const float inverse = ok ? rsqrtf(square) : 0.0f;
    m.mx[k] = ok ? x * inverse : 0.0f;
    m.my[k] = ok ? y * inverse : 0.0f;
    m.mz[k] = ok ? z * inverse : 1.0f;
    const float raw = o[5 * k + 3];
    const float soft = raw > 20.0f ? raw : __logf(1.0f + __expf(raw));
    m.kappa[k] = fminf(fmaxf(soft, kKappaMin), kKappaMax);
    m.weight[k] = __expf(o[5 * k + 4] - top);
    total += m.weight[k];
  }
  const float inverse = __builtin_amdgcn_rcpf(total);
#pragma unroll
  for (int k = 0; k < kComponents; ++k) {
    m.weight[k] *= inverse;
    const float kappa = m.kappa[k];
    m.scale[k] = m.weight[k] * kappa * 0.15915494309189533577f * __builtin_amdgcn_rcpf(1.0f - __expf(-2.0f * kappa));
  }
  return m;
}

// log of the vMF density kappa / (2 pi (1 - exp(-2 kappa))) exp(kappa (mu . w - 1)).
__host__ __device__ inline float log_vmf(float kappa, float dot) {
  return logf(kappa) - logf(6.28318530717958647692f) - log1pf(-expf(-2.0f * kappa)) +
         kappa * (dot - 1.0f);
}

__device__ inline float mixture_density(const Mixture &m, float wx, float wy, float wz) {
  float q = 0.0f;
#pragma unroll
  for (int k = 0; k < kComponents; ++k)
    q += m.scale[k] * __expf(m.kappa[k] * (m.mx[k] * wx + m.my[k] * wy + m.mz[k] * wz - 1.0f));
  return q;
}

// The direction at cosine w (sine s) from the unit axis (mx, my, mz) and azimuth phi about it,
// in an orthonormal frame around the axis (Frisvad's construction, branch-free in z sign).
__device__ inline void frame_direction(float mx, float my, float mz, float w, float s, float phi, float &wx,
                                       float &wy, float &wz) {
  const float sign = mz >= 0.0f ? 1.0f : -1.0f;
  const float a = -1.0f / (sign + mz);
  const float b = mx * my * a;
  const float e1x = 1.0f + sign * mx * mx * a, e1y = sign * b, e1z = -sign * mx;
  const float e2x = b, e2y = sign + my * my * a, e2z = -my;
  const float c = cosf(phi) * s, d = sinf(phi) * s;
  wx = w * mx + c * e1x + d * e2x;
  wy = w * my + c * e1y + d * e2y;
  wz = w * mz + c * e1z + d * e2z;
  const float norm = rsqrtf(wx * wx + wy * wy + wz * wz);
  wx *= norm, wy *= norm, wz *= norm;
}

// A direction of the vMF distribution of unit mean (mx, my, mz) and concentration kappa from
// two uniforms (Wood 1994 for the cosine to the mean).
__device__ inline void vmf_sample(float mx, float my, float mz, float kappa, float u1, float u2, float &wx,
                                  float &wy, float &wz) {
  // Cosine to the mean: w = 1 + log(u + (1 - u) exp(-2 kappa)) / kappa, in [-1, 1].
  const float w = fminf(1.0f, fmaxf(-1.0f, 1.0f + logf(u1 + (1.0f - u1) * expf(-2.0f * kappa)) / kappa));
  const float s = sqrtf(fmaxf(0.0f, 1.0f - w * w));
  frame_direction(mx, my, mz, w, s, 6.28318530717958647692f * u2, wx, wy, wz);
}

__device__ inline void mixture_sample(const Mixture &m, float u0, float u1, float u2, float &wx,
                                      float &wy, float &wz) {
  int k = kComponents - 1;
  float cumulative = 0.0f;
#pragma unroll
  for (int j = 0; j < kComponents - 1; ++j) {
    cumulative += m.weight[j];
    if (u0 < cumulative && k == kComponents - 1) k = j;
  }
  // Arithmetic selection of component k (a selected pointer into the arrays would keep the
  // whole mixture in scratch memory).
  float mx = 0.0f, my = 0.0f, mz = 0.0f, kappa = 0.0f;
#pragma unroll
  for (int j = 0; j < kComponents; ++j) {
    const float pick = j == k ? 1.0f : 0.0f;
    mx += pick * m.mx[j];
    my += pick * m.my[j];
    mz += pick * m.mz[j];
    kappa += pick * m.kappa[j];
  }
  vmf_sample(mx, my, mz, kappa, u1, u2, wx, wy, wz);
}

// Scattering events logged by the plain tracer for training (device pointers).
// A dispatch logs in two passes over the same histories: the counting pass writes each
// photon's number of logged events into starts[photon]; the host turns them into first slots
// (prefix sum from the samples already held, clamped to the capacity); the writing pass stores
// event k of a photon at starts[photon] + k. Samples end up in photon, then event order.
// With layer weights w (a path target), an estimate of detector weight C at an event where the
// history's layer paths are L adds C (w . L) to the photon's path contribution; the samples keep
// it and w . L at the event, the photons its total, so that the path the photon still travels
// before each later estimate, weighted by that estimate, is known after the dispatch.
struct EventLog {
  float *features;    // [capacity][kInputs]
  float *directions;  // [capacity][3] direction after the event
  float *cumulative;  // [capacity] detector weight of the history up to the event
  uint32_t *photon;   // [capacity] photon of the dispatch
  float *totals;      // [photons] detector weight of each whole history
  uint32_t *starts;   // [photons] events to log (counting pass), then first slots
  unsigned *count;    // samples held after the dispatch (set by the host)
  uint32_t capacity, photons;
  uint32_t counting;  // 1 in the counting pass
  const float *layer_weight;  // [layer_count] weights of the layer paths, or null (no path target)
  uint32_t layer_count;
  float *path_cumulative;     // [capacity] path contribution of the history up to the event
  float *event_path;          // [capacity] w . L at the event
  float *path_totals;         // [photons] path contribution of each whole history
};

// Host interface of a proposal handle, for the transport translation unit.
Net proposal_net(void *handle);
EventLog proposal_log(void *handle);
int proposal_device(void *handle);
uint32_t proposal_samples(void *handle);
// After a logging dispatch: training weights of events [first, logged) from the totals.
bool proposal_finish_log(void *handle, uint32_t first, hipStream_t stream, char *error, size_t size);

}  // namespace monki_net
