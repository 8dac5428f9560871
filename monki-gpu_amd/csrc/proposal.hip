// Training and probes of the learned scattering-direction proposal (network in proposal.h).
//
// Samples are scattering events logged by the plain tracer: features of the event, the
// direction the phase function gave the photon, and G = the detector weight (sum over views of
// C1, mc_routines_cartesian_mod.f90:1173-1194) the photon contributed after that event. The
// network is fitted by weighted maximum likelihood, loss = -sum G log q(direction | features) /
// sum G: drawn from the phase function and weighted by what followed, the fitted q approaches
// the product of the phase function and the photon's future contribution.
//
// A training step runs three kernels. train_kernel: one wave per 32-sample block runs the forward
// pass on the matrix cores (proposal.h), keeping the hidden pre-activations in a scratch buffer,
// the loss head, and the backward pass, which chains W^T products the same way; per layer it
// stores the block's dZ rows and H columns (bf16, transposed through LDS) in global memory in the
// matrix cores' operand order, and its loss. weight_gradients: dW = dZ H^T with the samples on
// the reduction side, one wave per 32 x 32 tile and 512 samples (split K); the waves of the first
// column tile also form db = dZ 1 the same way. sum_gradients: the split-K partials and the losses
// summed in a fixed order, so a training run is reproducible to the bit (no float atomics).

#include <hip/hip_runtime.h>

#include <math.h>
#include <stdint.h>
#include <stdio.h>
#include <string.h>

#include <random>
#include <vector>

#include "monki_hip.h"
#include "proposal.h"

using namespace monki_net;

namespace {

// Samples per training wave: one 32-sample block, the matrix cores' N.
constexpr int kWaveSamples = 32;
// bf16 per tile row: 32 samples + 8, so that rows start 16-byte aligned and the 16 lanes of a
// 16-byte LDS read phase hit distinct banks.
constexpr int kTileStride = kWaveSamples + 8;
constexpr int kZPerWave = 4 * 4096;  // pre-activations of the four hidden layers, D order
// Waves of one gradient launch (32,768 samples); larger batches run in chunks of this many.
constexpr uint32_t kChunkWaves = 1024;
// Weight-gradient operands and tiles. Layer l's dZ operands come in rows(l) / 32 blocks of 32
// rows, its H operands in col_tiles(l) blocks of 32 columns; per block, one operand (8 bf16 per
// lane) per 16-sample chunk. dW tiles: 32 x 32 blocks of every layer, layer by layer.
__host__ __device__ constexpr int row_tiles(int l) { return rows(l) / 32; }
__host__ __device__ constexpr int col_tiles(int l) { return (cols(l) + 31) / 32; }
__host__ __device__ constexpr int dz_block_offset(int l) { return l == 0 ? 0 : dz_block_offset(l - 1) + row_tiles(l - 1); }
__host__ __device__ constexpr int h_block_offset(int l) { return l == 0 ? 0 : h_block_offset(l - 1) + col_tiles(l - 1); }
__host__ __device__ constexpr int tile_offset(int l) {
  return l == 0 ? 0 : tile_offset(l - 1) + row_tiles(l - 1) * col_tiles(l - 1);
}
constexpr int kDzBlocks = dz_block_offset(kLayers);
constexpr int kHBlocks = h_block_offset(kLayers);
constexpr int kTiles = tile_offset(kLayers);
constexpr uint32_t kGroupChunks = 32;  // 16-sample chunks per weight-gradient wave (512 samples)

struct Proposal {
  int device = -1;
  float *params = nullptr, *grad = nullptr, *adam_m = nullptr, *adam_v = nullptr;
  bf16x8 *packed = nullptr, *packed_t = nullptr;
  float *bias = nullptr;
  float *z = nullptr;
  size_t z_waves = 0;
  bf16x8 *dz_ops = nullptr, *h_ops = nullptr;  // weight-gradient operands of a launch
  float *tile_part = nullptr;   // [kTiles][groups][64][16] split-K partial dW tiles, D order
  float *bias_part = nullptr;   // [kDzBlocks][groups][32] split-K partial db blocks
  double *loss_part = nullptr;  // [waves] loss of each wave
  // Samples: features [n][12], directions [n][3], weights [n].
  float *features = nullptr, *directions = nullptr, *weights = nullptr;
  size_t capacity = 0, count = 0;
  // Per-event contribution so far and per-photon totals of a logging dispatch.
  float *cumulative = nullptr, *totals = nullptr;
  uint32_t *photon = nullptr, *starts = nullptr;
  double *target_parts = nullptr;  // per-block sums of the training weights of a dispatch
  size_t target_blocks = 0;
  // Path target: layer weights, the share `path_mix` of the training weight that comes from the
  // weighted path, and the logged path contributions (see EventLog).
  float *layer_weight = nullptr, *path_cumulative = nullptr, *event_path = nullptr, *path_totals = nullptr;
  float path_mix = 0.0f;
  uint32_t layer_count = 0;
  unsigned *log_count = nullptr;
  size_t photon_capacity = 0;
  double *loss = nullptr;
  double weight_sum = 0.0;
  uint64_t step = 0;
  hipStream_t stream = nullptr;
};

bool ok(hipError_t status, char *error, size_t size, const char *what) {
  if (status == hipSuccess) return true;
  if (error && size) snprintf(error, size, "%s: %s", what, hipGetErrorString(status));
  return false;
}

void fail(char *error, size_t size, const char *text) {
  if (error && size) snprintf(error, size, "%s", text);
}

// ---- Kernels -----------------------------------------------------------------------------

// Master parameters -> padded bf16 operands (A order and transposed) and padded biases.
__global__ void pack_kernel(const float *__restrict__ params, bf16x8 *__restrict__ packed,
                            bf16x8 *__restrict__ packed_t, float *__restrict__ bias) {
  const int index = blockIdx.x * blockDim.x + threadIdx.x;
  const int lane = index & 63;
  const int block = index >> 6;
  auto weight = [&](int l, int r, int c) -> float {
    return r < true_rows(l) && c < true_cols(l) ? params[param_offset(l) + r * true_cols(l) + c] : 0.0f;
  };
  // Forward operands: block (mb, s) of layer l, lane -> W[32 mb + lane % 32][16 s + slot_row(8 h + e)].
  if (index < kPacked) {
    int l = 0;
    while (l + 1 < kLayers && index >= packed_offset(l + 1)) ++l;
    const int local = (index - packed_offset(l)) >> 6;
    const int steps = cols(l) / 16;
    const int mb = local / steps, s = local % steps;
    const int r = 32 * mb + (lane & 31);
    float v[8];
    for (int e = 0; e < 8; ++e) v[e] = weight(l, r, 16 * s + slot_row(8 * (lane >> 5) + e));
    packed[index] = pack8(v);
  }
  // Transposed operands of layers 1..4: A'[m'][k'] = W[k'][m'], M' = cols(l), K' = rows(l).
  if (index < kPackedT) {
    int l = 1;
    while (l + 1 < kLayers && index >= packed_t_offset(l + 1)) ++l;
    const int local = (index - packed_t_offset(l)) >> 6;
    const int steps = rows(l) / 16;
    const int mb = local / steps, s = local % steps;
    const int m = 32 * mb + (lane & 31);
    float v[8];
    for (int e = 0; e < 8; ++e) v[e] = weight(l, 16 * s + slot_row(8 * (lane >> 5) + e), m);
    packed_t[index] = pack8(v);
  }
  (void)block;
  if (index < kBiases) {
    int l = 0;
    while (l + 1 < kLayers && index >= bias_offset(l + 1)) ++l;
    const int r = index - bias_offset(l);
    bias[index] = r < true_rows(l) ? params[param_offset(l) + true_rows(l) * true_cols(l) + r] : 0.0f;
  }
}

__device__ inline uint32_t mix32(uint64_t x) {
  x ^= x >> 33;
  x *= 0xff51afd7ed558ccdull;
  x ^= x >> 33;
  x *= 0xc4ceb9fe1a85ec53ull;
  x ^= x >> 33;
  return uint32_t(x);
}

// Loss -G log q(w) of the mixture and its gradient with respect to the 40 outputs.
__device__ inline float head_loss(const float *o, float wx, float wy, float wz, float g, float *dy) {
  float logit_top = -INFINITY;
  for (int k = 0; k < kComponents; ++k) logit_top = fmaxf(logit_top, o[5 * k + 4]);
  float logit_sum = 0.0f;
  for (int k = 0; k < kComponents; ++k) logit_sum += expf(o[5 * k + 4] - logit_top);
  const float log_norm = logit_top + logf(logit_sum);
  float ell[kComponents], kappa[kComponents], dot[kComponents], mu[kComponents][3], norm[kComponents];
  float top = -INFINITY;
  for (int k = 0; k < kComponents; ++k) {
    const float x = o[5 * k], y = o[5 * k + 1], z = o[5 * k + 2];
    norm[k] = sqrtf(x * x + y * y + z * z);
    const bool valid = norm[k] > 1e-6f;
    mu[k][0] = valid ? x / norm[k] : 0.0f;
    mu[k][1] = valid ? y / norm[k] : 0.0f;
    mu[k][2] = valid ? z / norm[k] : 1.0f;
    kappa[k] = fminf(fmaxf(softplus(o[5 * k + 3]), kKappaMin), kKappaMax);
    dot[k] = mu[k][0] * wx + mu[k][1] * wy + mu[k][2] * wz;
    ell[k] = o[5 * k + 4] - log_norm + log_vmf(kappa[k], dot[k]);
    top = fmaxf(top, ell[k]);
  }
  float sum = 0.0f;
  for (int k = 0; k < kComponents; ++k) sum += expf(ell[k] - top);
  const float log_q = top + logf(sum);
  for (int k = 0; k < kComponents; ++k) {
    const float r = expf(ell[k] - log_q);
    const float pi = expf(o[5 * k + 4] - log_norm);
    const float dell = -g * r;
    dy[5 * k + 4] = -g * (r - pi);
    const float soft = softplus(o[5 * k + 3]);
    const float dlog_kappa = 1.0f / kappa[k] - 2.0f / expm1f(2.0f * kappa[k]) + dot[k] - 1.0f;
    const bool free_kappa = soft > kKappaMin && soft < kKappaMax;
    dy[5 * k + 3] = free_kappa ? dell * dlog_kappa / (1.0f + expf(-o[5 * k + 3])) : 0.0f;
    // d log vmf / d mu = kappa w, projected onto the tangent of the normalization.
    const float gx = dell * kappa[k] * wx, gy = dell * kappa[k] * wy, gz = dell * kappa[k] * wz;
    const float along = gx * mu[k][0] + gy * mu[k][1] + gz * mu[k][2];
    const bool valid = norm[k] > 1e-6f;
    dy[5 * k + 0] = valid ? (gx - along * mu[k][0]) / norm[k] : 0.0f;
    dy[5 * k + 1] = valid ? (gy - along * mu[k][1]) / norm[k] : 0.0f;
    dy[5 * k + 2] = valid ? (gz - along * mu[k][2]) / norm[k] : 0.0f;
  }
  return -g * log_q;
}

// Row `low` of the per-lane array v held by lane src for lanes 0..31, row `high` for lanes
// 32..63 (rows fixed at compile time): two shuffles, one per half.
template <int Low, int High, int Count>
__device__ inline float shuffle_rows(const float (&v)[Count], int src) {
  const float low = Low < Count ? __shfl(v[Low < Count ? Low : 0], src, 64) : 0.0f;
  const float high = High < Count ? __shfl(v[High < Count ? High : 0], src, 64) : 0.0f;
  return __lane_id() < 32 ? low : high;
}

// dY of sample column 32 nb + n as the B operand of k-step S: lane half h holds rows
// 16 S + slot_row(8 h + e).
template <int S, int E = 0>
__device__ inline void dy_operand(const float (&dy)[kOutputs], int src, float (&e)[8]) {
  if constexpr (E < 8) {
    e[E] = shuffle_rows<16 * S + slot_row(E), 16 * S + slot_row(8 + E), kOutputs>(dy, src);
    dy_operand<S, E + 1>(dy, src, e);
  }
}

template <int S = 0>
__device__ inline void dy_operands(const float (&dy)[kOutputs], int src, bf16x8 (&d)[8]) {
  if constexpr (S < 4) {
    float e[8];
    dy_operand<S>(dy, src, e);
    d[S] = pack8(e);
    dy_operands<S + 1>(dy, src, d);
  }
}

// Input rows K + h (h = lane half) of sample src for the layer-0 H tile, K even.
template <int K = 0>
__device__ inline void input_rows(const float (&f)[kInputs], int src, int column, uint16_t *h_tile) {
  if constexpr (K < 32) {
    const int row = K + (__lane_id() >> 5);
    const float value = shuffle_rows<K, K + 1, kInputs>(f, src);
    h_tile[row * kTileStride + column] = row < kInputs ? bf16_bits(value) : 0;
    input_rows<K + 2>(f, src, column, h_tile);
  }
}

// One training wave: 32 samples (sample indices from `indices`, or hashed from (seed, step)), one
// per sample column n = lane % 32; both lane halves hold the same sample (the matrix cores' B and
// D layouts split each column over the halves), and only the lower half counts its loss.
__global__ void __launch_bounds__(64)
    train_kernel(Net net, const float *__restrict__ features, const float *__restrict__ directions,
                 const float *__restrict__ weights, const uint32_t *__restrict__ indices, uint32_t count,
                 uint32_t base, uint32_t batch, uint64_t seed, float scale, bf16x8 *__restrict__ dz_ops,
                 bf16x8 *__restrict__ h_ops, uint32_t chunks, float *__restrict__ z_scratch,
                 double *__restrict__ loss_part) {
  __shared__ __attribute__((aligned(16))) uint16_t dz_tile[128 * kTileStride];
  __shared__ __attribute__((aligned(16))) uint16_t h_tile[128 * kTileStride];
  const int lane = __lane_id();
  const int h = lane >> 5, n = lane & 31;
  const uint32_t sample = base + blockIdx.x * uint32_t(kWaveSamples) + uint32_t(n);
  const bool valid = sample < batch;
  const uint32_t index = !valid ? 0u : indices ? indices[sample] : mix32(seed * 0x9E3779B97F4A7C15ull + sample) % count;
  float f[kInputs];
#pragma unroll
  for (int k = 0; k < kInputs; ++k) f[k] = valid ? features[size_t(index) * kInputs + k] : 0.0f;
  const float wx = valid ? directions[size_t(index) * 3] : 0.0f;
  const float wy = valid ? directions[size_t(index) * 3 + 1] : 0.0f;
  const float wz = valid ? directions[size_t(index) * 3 + 2] : 1.0f;
  const float g = valid ? weights[index] * scale : 0.0f;
  float *z = z_scratch + size_t(blockIdx.x) * kZPerWave;

  float o[kOutputs];
  {
    bf16x8 x[8];
    scatter_inputs(f, n, x);
    floatx16 raw[2];
    forward_block(net, x, raw, z);
    gather_outputs(raw, n, true, o);
  }
  float dy[kOutputs];
  const float loss = valid ? head_loss(o, wx, wy, wz, g, dy) : 0.0f;
  if (!valid) {
#pragma unroll
    for (int k = 0; k < kOutputs; ++k) dy[k] = 0.0f;
  }
  double total = h == 0 ? loss : 0.0;
  for (int offset = 32; offset > 0; offset >>= 1) total += __shfl_xor(total, offset, 64);
  if (lane == 0) loss_part[blockIdx.x] = total;

  // dY as B operands (rows 16 s + slot_row(8 h + e) of 64, of sample column n).
  bf16x8 d[8];
  dy_operands(dy, n, d);

#pragma unroll
  for (int l = kLayers - 1; l >= 0; --l) {
    const int r_steps = rows(l) / 16;
    // dZ_{l+1} tile [row][sample] from the B operands.
#pragma unroll
    for (int s = 0; s < r_steps; ++s) {
      const bf16x8 v = d[s];
#pragma unroll
      for (int j = 0; j < 8; ++j) dz_tile[(16 * s + slot_row(8 * h + j)) * kTileStride + n] = uint16_t(v[j]);
    }
    // H_l tile [feature][sample]: inputs for l = 0, else SiLU of the saved pre-activations.
    const int c_blocks = (cols(l) + 31) / 32;
    if (l == 0) {
      input_rows(f, n, n, h_tile);
    } else {
      const float *zl = z + (l - 1) * 4096;
#pragma unroll
      for (int mb = 0; mb < 4; ++mb) {
#pragma unroll
        for (int v = 0; v < 16; ++v) {
          const int row = 32 * mb + 8 * (v >> 2) + 4 * h + (v & 3);
          h_tile[row * kTileStride + n] = bf16_bits(silu(zl[(mb * 16 + v) * 64 + lane]));
        }
      }
    }
    __syncthreads();
    // The operands of dW_l = dZ_{l+1} H_l^T for this block's two 16-sample chunks (any sample
    // order serves, as long as both operands use the same one).
#pragma unroll
    for (int s = 0; s < kWaveSamples / 16; ++s) {
      const size_t chunk = size_t(blockIdx.x) * (kWaveSamples / 16) + s;
      for (int mb = 0; mb < rows(l) / 32; ++mb)
        dz_ops[(size_t(dz_block_offset(l) + mb) * chunks + chunk) * 64 + lane] =
            *reinterpret_cast<const bf16x8 *>(dz_tile + (32 * mb + n) * kTileStride + 16 * s + 8 * h);
      for (int cb = 0; cb < c_blocks; ++cb)
        h_ops[(size_t(h_block_offset(l) + cb) * chunks + chunk) * 64 + lane] =
            *reinterpret_cast<const bf16x8 *>(h_tile + (32 * cb + n) * kTileStride + 16 * s + 8 * h);
    }
    __syncthreads();
    if (l == 0) break;
    // dH_l = W_l^T dZ_{l+1}, then dZ_l = dH_l * silu'(Z_l): chained like the forward pass.
    const float *zl = z + (l - 1) * 4096;
    bf16x8 next[8];
#pragma unroll
    for (int mb = 0; mb < 4; ++mb) {
      floatx16 c = {0, 0, 0, 0, 0, 0, 0, 0, 0, 0, 0, 0, 0, 0, 0, 0};
#pragma unroll
      for (int s = 0; s < r_steps; ++s)
        c = mfma(net.packed_t[packed_t_offset(l) + (mb * r_steps + s) * 64 + lane], d[s], c);
#pragma unroll
      for (int j = 0; j < 2; ++j) {
        float e[8];
#pragma unroll
        for (int k = 0; k < 8; ++k) e[k] = c[8 * j + k] * silu_grad(zl[(mb * 16 + 8 * j + k) * 64 + lane]);
        next[2 * mb + j] = pack8(e);
      }
      __builtin_amdgcn_sched_barrier(0);
    }
#pragma unroll
    for (int s = 0; s < 8; ++s) d[s] = next[s];
  }
}

// Forward probe: 40 outputs per sample, the per-lane path that transport uses.
__global__ void __launch_bounds__(64) forward_kernel(Net net, const float *__restrict__ features,
                                                     uint32_t count, float *__restrict__ out) {
  const int lane = __lane_id();
  const uint32_t sample = blockIdx.x * 64u + uint32_t(lane);
  float f[kInputs];
#pragma unroll
  for (int k = 0; k < kInputs; ++k) f[k] = sample < count ? features[size_t(sample) * kInputs + k] : 0.0f;
  float o[kOutputs];
#pragma unroll
  for (int nb = 0; nb < 2; ++nb) {
    bf16x8 x[8];
    scatter_inputs(f, 32 * nb + (lane & 31), x);
    floatx16 raw[2];
    forward_block(net, x, raw, nullptr);
    gather_outputs(raw, lane & 31, (lane >> 5) == nb, o);
  }
  if (sample < count) {
    for (int k = 0; k < kOutputs; ++k) out[size_t(sample) * kOutputs + k] = o[k];
  }
}

__global__ void adam_kernel(float *__restrict__ params, const float *__restrict__ grad,
                            float *__restrict__ m, float *__restrict__ v, int count, float lr,
                            float beta1, float beta2, float correction1, float correction2) {
  const int i = blockIdx.x * blockDim.x + threadIdx.x;
  if (i >= count) return;
  const float g = grad[i];
  const float mi = beta1 * m[i] + (1.0f - beta1) * g;
  const float vi = beta2 * v[i] + (1.0f - beta2) * g * g;
  m[i] = mi;
  v[i] = vi;
  params[i] -= lr * (mi / correction1) / (sqrtf(vi / correction2) + 1e-8f);
}

// dW tiles, split K: one wave per (tile, group of kGroupChunks 16-sample chunks), the partial
// tile in D order [lane][register]; waves of column tile 0 also sum their dZ rows (B = ones).
__global__ void __launch_bounds__(64)
    weight_gradients(const bf16x8 *__restrict__ dz_ops, const bf16x8 *__restrict__ h_ops, uint32_t chunks,
                     uint32_t groups, float *__restrict__ tile_part, float *__restrict__ bias_part) {
  const int lane = __lane_id();
  const uint32_t work = blockIdx.x;
  const int tile = int(work / groups);
  const uint32_t group = work % groups;
  int l = 0;
  while (l + 1 < kLayers && tile >= tile_offset(l + 1)) ++l;
  const int local = tile - tile_offset(l);
  const int mb = local / col_tiles(l), cb = local % col_tiles(l);
  const bf16x8 *a = dz_ops + size_t(dz_block_offset(l) + mb) * chunks * 64 + lane;
  const bf16x8 *b = h_ops + size_t(h_block_offset(l) + cb) * chunks * 64 + lane;
  const uint32_t first = group * kGroupChunks;
  const uint32_t last = first + kGroupChunks < chunks ? first + kGroupChunks : chunks;
  floatx16 c = {0, 0, 0, 0, 0, 0, 0, 0, 0, 0, 0, 0, 0, 0, 0, 0};
  if (cb == 0) {
    const short one = 0x3F80;  // bf16 1.0
    const bf16x8 ones = {one, one, one, one, one, one, one, one};
    floatx16 rows = {0, 0, 0, 0, 0, 0, 0, 0, 0, 0, 0, 0, 0, 0, 0, 0};
    for (uint32_t chunk = first; chunk < last; ++chunk) {
      const bf16x8 x = a[size_t(chunk) * 64];
      c = mfma(x, b[size_t(chunk) * 64], c);
      rows = mfma(x, ones, rows);
    }
    // Column 0 of the row sums: lanes 0 and 32, rows 8 (v / 4) + 4 (lane / 32) + v % 4.
    if ((lane & 31) == 0) {
      float *out = bias_part + (size_t(dz_block_offset(l) + mb) * groups + group) * 32;
#pragma unroll
      for (int v = 0; v < 16; ++v) out[8 * (v >> 2) + 4 * (lane >> 5) + (v & 3)] = rows[v];
    }
  } else {
    for (uint32_t chunk = first; chunk < last; ++chunk) c = mfma(a[size_t(chunk) * 64], b[size_t(chunk) * 64], c);
  }
  *reinterpret_cast<floatx16 *>(tile_part + (size_t(work) * 64 + lane) * 16) = c;
}

// Gradient of a launch (into grad, or added to it for the later chunks of a large batch), 256
// threads a block: blocks [0, 4 kTiles) sum a quarter of one dW tile's split-K partials in group
// order (block 4 tile + j, thread t: D element 256 j + t), blocks [4 kTiles, 4 kTiles + kDzBlocks)
// one db block's, and the last block the waves' losses (strided partial sums, then a fixed tree).
constexpr uint32_t kSumBlocks = 4 * kTiles + kDzBlocks + 1;
__global__ void __launch_bounds__(256)
    sum_gradients(const float *__restrict__ tile_part, const float *__restrict__ bias_part,
                  const double *__restrict__ loss_part, uint32_t waves, uint32_t groups, bool add,
                  float *__restrict__ grad, double *__restrict__ loss) {
  const int t = threadIdx.x;
  const int block = blockIdx.x;
  if (block < 4 * kTiles) {
    const int tile = block >> 2, e = 256 * (block & 3) + t;
    int l = 0;
    while (l + 1 < kLayers && tile >= tile_offset(l + 1)) ++l;
    const int mb = (tile - tile_offset(l)) / col_tiles(l), cb = (tile - tile_offset(l)) % col_tiles(l);
    float sum = 0.0f;
    for (uint32_t g = 0; g < groups; ++g) sum += tile_part[(size_t(tile) * groups + g) * 1024 + e];
    // D order: element 16 lane + v holds row 8 (v / 4) + 4 (lane / 32) + v % 4, column lane % 32.
    const int lane = e >> 4, v = e & 15;
    const int row = 32 * mb + 8 * (v >> 2) + 4 * (lane >> 5) + (v & 3), col = 32 * cb + (lane & 31);
    if (row < true_rows(l) && col < true_cols(l)) {
      float &out = grad[param_offset(l) + row * true_cols(l) + col];
      out = add ? out + sum : sum;
    }
  } else if (block < 4 * kTiles + kDzBlocks) {
    const int b = block - 4 * kTiles;
    int l = 0;
    while (l + 1 < kLayers && b >= dz_block_offset(l + 1)) ++l;
    const int row = 32 * (b - dz_block_offset(l)) + t;
    if (t < 32 && row < true_rows(l)) {
      float sum = 0.0f;
      for (uint32_t g = 0; g < groups; ++g) sum += bias_part[(size_t(b) * groups + g) * 32 + t];
      float &out = grad[param_offset(l) + true_rows(l) * true_cols(l) + row];
      out = add ? out + sum : sum;
    }
  } else {
    __shared__ double partial[256];
    double sum = 0.0;
    for (uint32_t w = t; w < waves; w += 256) sum += loss_part[w];
    partial[t] = sum;
    __syncthreads();
    for (int width = 128; width > 0; width >>= 1) {
      if (t < width) partial[t] += partial[t + width];
      __syncthreads();
    }
    if (t == 0) *loss = add ? *loss + partial[0] : partial[0];
  }
}

// Training weights of the events logged in one dispatch: G = photon total - contribution so far
// (the detector weight the photon contributed after the event) and, with a path target,
// P = (path total - path contribution so far) - (w . L at the event) G >= 0: the weighted path
// the photon still travelled before each later estimate, times that estimate's weight. P goes
// into event_path in place. One sum of each per 256-event block (lanes, then the block's four
// waves, in a fixed order).
__global__ void __launch_bounds__(256)
    targets_kernel(const float *__restrict__ totals, const uint32_t *__restrict__ photon,
                   const float *__restrict__ cumulative, const float *__restrict__ path_totals,
                   const float *__restrict__ path_cumulative, float *__restrict__ event_path,
                   float *__restrict__ weights, uint32_t first, uint32_t last, bool path,
                   double *__restrict__ block_sums) {
  __shared__ double waves[2][4];
  const uint32_t i = first + blockIdx.x * blockDim.x + threadIdx.x;
  double value = 0.0, path_value = 0.0;
  if (i < last) {
    const float g = fmaxf(0.0f, totals[photon[i]] - cumulative[i]);
    weights[i] = g;
    value = g;
    if (path) {
      const float p = fmaxf(0.0f, (path_totals[photon[i]] - path_cumulative[i]) - event_path[i] * g);
      event_path[i] = p;
      path_value = p;
    }
  }
  for (int offset = 32; offset > 0; offset >>= 1) {
    value += __shfl_xor(value, offset, 64);
    path_value += __shfl_xor(path_value, offset, 64);
  }
  if (__lane_id() == 0) {
    waves[0][threadIdx.x >> 6] = value;
    waves[1][threadIdx.x >> 6] = path_value;
  }
  __syncthreads();
  if (threadIdx.x < 2)
    block_sums[2 * blockIdx.x + threadIdx.x] =
        ((waves[threadIdx.x][0] + waves[threadIdx.x][1]) + waves[threadIdx.x][2]) + waves[threadIdx.x][3];
}

// Training weight with a path target: (1 - mix) G + mix scale P, scale = sum G / sum P over the
// dispatch, so both parts carry the same total.
__global__ void mix_targets(float *__restrict__ weights, const float *__restrict__ path_weights, uint32_t first,
                            uint32_t last, float detector_share, float path_share) {
  const uint32_t i = first + blockIdx.x * blockDim.x + threadIdx.x;
  if (i < last) weights[i] = detector_share * weights[i] + path_share * path_weights[i];
}

Net view(const Proposal *p) { return Net{p->packed, p->packed_t, p->bias}; }

bool repack(Proposal *p, char *error, size_t size) {
  const int threads = 256;
  const int work = kPacked > kPackedT ? kPacked : kPackedT;
  hipLaunchKernelGGL(pack_kernel, dim3((work + threads - 1) / threads), dim3(threads), 0, p->stream,
                     p->params, p->packed, p->packed_t, p->bias);
  return ok(hipGetLastError(), error, size, "pack proposal");
}

void release(Proposal *p) {
  if (!p) return;
  if (p->device >= 0) (void)hipSetDevice(p->device);
  if (p->stream) (void)hipStreamDestroy(p->stream);
  void *buffers[] = {p->params, p->grad, p->adam_m, p->adam_v, p->packed, p->packed_t, p->bias, p->z,
                     p->dz_ops, p->h_ops, p->bias_part, p->tile_part, p->loss_part, p->features,
                     p->directions, p->weights, p->cumulative,
                     p->totals, p->photon, p->starts, p->target_parts, p->log_count, p->loss,
                     p->layer_weight, p->path_cumulative, p->event_path, p->path_totals};
  for (void *b : buffers) (void)hipFree(b);
  delete p;
}

// Per-wave buffers of a gradient launch (pre-activations, weight-gradient operands, bias rows,
// losses) and its split-K tiles, kept for later launches.
bool ensure_scratch(Proposal *p, size_t waves, char *error, size_t size) {
  if (waves <= p->z_waves) return true;
  void *buffers[] = {p->z, p->dz_ops, p->h_ops, p->tile_part, p->bias_part, p->loss_part};
  for (void *b : buffers) (void)hipFree(b);
  p->z = nullptr;
  p->dz_ops = p->h_ops = nullptr;
  p->tile_part = p->bias_part = nullptr;
  p->loss_part = nullptr;
  p->z_waves = 0;
  const size_t chunks = waves * (kWaveSamples / 16);
  const size_t groups = (chunks + kGroupChunks - 1) / kGroupChunks;
  if (!ok(hipMalloc(&p->z, waves * kZPerWave * sizeof(float)), error, size, "allocate proposal scratch") ||
      !ok(hipMalloc(&p->dz_ops, size_t(kDzBlocks) * chunks * 64 * sizeof(bf16x8)), error, size,
          "allocate gradient operands") ||
      !ok(hipMalloc(&p->h_ops, size_t(kHBlocks) * chunks * 64 * sizeof(bf16x8)), error, size,
          "allocate gradient operands") ||
      !ok(hipMalloc(&p->tile_part, size_t(kTiles) * groups * 64 * 16 * sizeof(float)), error, size,
          "allocate gradient tiles") ||
      !ok(hipMalloc(&p->bias_part, size_t(kDzBlocks) * groups * 32 * sizeof(float)), error, size,
          "allocate bias tiles") ||
      !ok(hipMalloc(&p->loss_part, waves * sizeof(double)), error, size, "allocate partial losses"))
    return false;
  p->z_waves = waves;
  return true;
}

// Launches the gradient of the mean loss over `batch` samples into p->grad and the loss into
// device_loss, on p->stream (no synchronization). Batches past kChunkWaves waves run in chunks
// whose sums add up in chunk order.
bool gradients(Proposal *p, const uint32_t *indices, uint32_t batch, uint64_t seed, double *device_loss,
               char *error, size_t size) {
  const uint32_t total_waves = (batch + kWaveSamples - 1) / kWaveSamples;
  const uint32_t chunk = total_waves < kChunkWaves ? total_waves : kChunkWaves;
  if (!ensure_scratch(p, chunk, error, size)) return false;
  const float scale = p->weight_sum > 0.0 ? float(double(p->count) / p->weight_sum / batch) : 0.0f;
  for (uint32_t first = 0; first < total_waves; first += chunk) {
    const uint32_t waves = total_waves - first < chunk ? total_waves - first : chunk;
    const uint32_t chunks = waves * uint32_t(kWaveSamples / 16);
    const uint32_t groups = (chunks + kGroupChunks - 1) / kGroupChunks;
    hipLaunchKernelGGL(train_kernel, dim3(waves), dim3(64), 0, p->stream, view(p), p->features,
                       p->directions, p->weights, indices, uint32_t(p->count), first * uint32_t(kWaveSamples), batch,
                       seed, scale, p->dz_ops, p->h_ops, chunks, p->z, p->loss_part);
    hipLaunchKernelGGL(weight_gradients, dim3(uint32_t(kTiles) * groups), dim3(64), 0, p->stream, p->dz_ops,
                       p->h_ops, chunks, groups, p->tile_part, p->bias_part);
    hipLaunchKernelGGL(sum_gradients, dim3(kSumBlocks), dim3(256), 0, p->stream, p->tile_part,
                       p->bias_part, p->loss_part, waves, groups, first > 0, p->grad, device_loss);
    if (!ok(hipGetLastError(), error, size, "train launch")) return false;
  }
  return true;
}

bool train_steps(Proposal *p, uint32_t steps, uint32_t batch, float learning_rate, uint64_t seed,
                 double *device_losses, char *error, size_t error_size) {
  const float beta1 = 0.9f, beta2 = 0.999f;
  for (uint32_t i = 0; i < steps; ++i) {
    if (!gradients(p, nullptr, batch, seed * 1000003ull + p->step, device_losses + i, error, error_size))
      return false;
    ++p->step;
    const float c1 = 1.0f - powf(beta1, float(p->step));
    const float c2 = 1.0f - powf(beta2, float(p->step));
    hipLaunchKernelGGL(adam_kernel, dim3((kParams + 255) / 256), dim3(256), 0, p->stream, p->params, p->grad,
                       p->adam_m, p->adam_v, kParams, learning_rate, beta1, beta2, c1, c2);
    if (!ok(hipGetLastError(), error, error_size, "adam launch") || !repack(p, error, error_size)) return false;
  }
  return true;
}

}  // namespace

// Internal interface for the transport translation unit (same library).
namespace monki_net {

Net proposal_net(void *handle) { return view(static_cast<Proposal *>(handle)); }

EventLog proposal_log(void *handle) {
  Proposal *p = static_cast<Proposal *>(handle);
  const bool path = p->path_mix > 0.0f && p->layer_weight;
  return EventLog{p->features, p->directions, p->cumulative, p->photon, p->totals, p->starts, p->log_count,
                  uint32_t(p->capacity), uint32_t(p->photon_capacity), 0u, path ? p->layer_weight : nullptr,
                  path ? p->layer_count : 0u, p->path_cumulative, p->event_path, p->path_totals};
}

int proposal_device(void *handle) { return static_cast<Proposal *>(handle)->device; }

uint32_t proposal_samples(void *handle) { return uint32_t(static_cast<Proposal *>(handle)->count); }

bool proposal_finish_log(void *handle, uint32_t first, hipStream_t stream, char *error, size_t size) {
  Proposal *p = static_cast<Proposal *>(handle);
  unsigned logged = 0;
  if (!ok(hipMemcpyAsync(&logged, p->log_count, sizeof(unsigned), hipMemcpyDeviceToHost, stream), error,
          size, "copy log count") ||
      !ok(hipStreamSynchronize(stream), error, size, "log count"))
    return false;
  const uint32_t last = logged < p->capacity ? logged : uint32_t(p->capacity);
  if (last > first) {
    const size_t blocks = (last - first + 255) / 256;
    if (blocks > p->target_blocks) {
      (void)hipFree(p->target_parts);
      p->target_parts = nullptr;
      p->target_blocks = 0;
      if (!ok(hipMalloc(&p->target_parts, 2 * blocks * sizeof(double)), error, size, "allocate target sums"))
        return false;
      p->target_blocks = blocks;
    }
    const bool path = p->path_mix > 0.0f && p->layer_weight;
    hipLaunchKernelGGL(targets_kernel, dim3(unsigned(blocks)), dim3(256), 0, stream, p->totals, p->photon,
                       p->cumulative, p->path_totals, p->path_cumulative, p->event_path, p->weights, first, last,
                       path, p->target_parts);
    std::vector<double> sums(2 * blocks);
    if (!ok(hipGetLastError(), error, size, "targets launch") ||
        !ok(hipMemcpyAsync(sums.data(), p->target_parts, 2 * blocks * sizeof(double), hipMemcpyDeviceToHost,
                           stream),
            error, size, "copy target sums") ||
        !ok(hipStreamSynchronize(stream), error, size, "targets"))
      return false;
    double detector = 0.0, path_sum = 0.0;
    for (size_t b = 0; b < blocks; ++b) {  // block order: reproducible
      detector += sums[2 * b];
      path_sum += sums[2 * b + 1];
    }
    if (path && path_sum > 0.0) {
      const float share = p->path_mix;
      hipLaunchKernelGGL(mix_targets, dim3(unsigned(blocks)), dim3(256), 0, stream, p->weights, p->event_path,
                         first, last, 1.0f - share, float(share * detector / path_sum));
      if (!ok(hipGetLastError(), error, size, "mix targets launch")) return false;
    }
    p->weight_sum += detector;
  }
  p->count = last;
  return true;
}

}  // namespace monki_net

extern "C" {

size_t monki_hip_proposal_parameter_count(void) { return size_t(kParams); }

void *monki_hip_proposal_create(uint32_t device, uint32_t seed, char *error, size_t error_size) {
  int devices = 0;
  if (hipGetDeviceCount(&devices) != hipSuccess || device >= uint32_t(devices)) {
    fail(error, error_size, "HIP device unavailable");
    return nullptr;
  }
  Proposal *p = new Proposal();
  p->device = int(device);
  // He-style initialization for SiLU layers; small output layer so the first mixture is broad.
  std::vector<float> host(kParams, 0.0f);
  std::mt19937 generator(seed);
  for (int l = 0; l < kLayers; ++l) {
    const float scale = (l == kLayers - 1 ? 0.1f : 1.0f) * sqrtf(2.0f / float(true_cols(l)));
    std::normal_distribution<float> normal(0.0f, scale);
    for (int i = 0; i < true_rows(l) * true_cols(l); ++i) host[param_offset(l) + i] = normal(generator);
  }
  for (int k = 0; k < kComponents; ++k)  // kappa = softplus(1) ~ 1.3
    host[param_offset(4) + true_rows(4) * true_cols(4) + 5 * k + 3] = 1.0f;
  if (!ok(hipSetDevice(int(device)), error, error_size, "hipSetDevice") ||
      !ok(hipStreamCreate(&p->stream), error, error_size, "create stream") ||
      !ok(hipMalloc(&p->params, kParams * sizeof(float)), error, error_size, "allocate parameters") ||
      !ok(hipMalloc(&p->grad, kParams * sizeof(float)), error, error_size, "allocate gradient") ||
      !ok(hipMalloc(&p->adam_m, kParams * sizeof(float)), error, error_size, "allocate moments") ||
      !ok(hipMalloc(&p->adam_v, kParams * sizeof(float)), error, error_size, "allocate moments") ||
      !ok(hipMalloc(&p->packed, kPacked * sizeof(bf16x8)), error, error_size, "allocate operands") ||
      !ok(hipMalloc(&p->packed_t, kPackedT * sizeof(bf16x8)), error, error_size, "allocate operands") ||
      !ok(hipMalloc(&p->bias, kBiases * sizeof(float)), error, error_size, "allocate biases") ||
      !ok(hipMalloc(&p->log_count, sizeof(unsigned)), error, error_size, "allocate log counter") ||
      !ok(hipMalloc(&p->loss, sizeof(double)), error, error_size, "allocate loss") ||
      !ok(hipMemcpy(p->params, host.data(), kParams * sizeof(float), hipMemcpyHostToDevice), error,
          error_size, "upload parameters") ||
      !ok(hipMemset(p->adam_m, 0, kParams * sizeof(float)), error, error_size, "clear moments") ||
      !ok(hipMemset(p->adam_v, 0, kParams * sizeof(float)), error, error_size, "clear moments") ||
      !ok(hipMemset(p->log_count, 0, sizeof(unsigned)), error, error_size, "clear log counter") ||
      !repack(p, error, error_size) || !ok(hipStreamSynchronize(p->stream), error, error_size, "pack")) {
    release(p);
    return nullptr;
  }
  return p;
}

int monki_hip_proposal_set_parameters(void *handle, const float *params, size_t count, char *error,
                                      size_t error_size) {
  Proposal *p = static_cast<Proposal *>(handle);
  if (!p || !params || count != size_t(kParams)) {
    fail(error, error_size, "Invalid proposal parameters");
    return 1;
  }
  if (!ok(hipSetDevice(p->device), error, error_size, "hipSetDevice") ||
      !ok(hipMemcpy(p->params, params, kParams * sizeof(float), hipMemcpyHostToDevice), error, error_size,
          "upload parameters") ||
      !repack(p, error, error_size) || !ok(hipStreamSynchronize(p->stream), error, error_size, "pack"))
    return 1;
  return 0;
}

int monki_hip_proposal_get_parameters(void *handle, float *params, size_t count) {
  Proposal *p = static_cast<Proposal *>(handle);
  if (!p || !params || count != size_t(kParams)) return 1;
  if (hipSetDevice(p->device) != hipSuccess) return 1;
  return hipMemcpy(params, p->params, kParams * sizeof(float), hipMemcpyDeviceToHost) == hipSuccess ? 0 : 1;
}

int monki_hip_proposal_reserve(void *handle, size_t events, size_t photons, char *error, size_t error_size) {
  Proposal *p = static_cast<Proposal *>(handle);
  if (!p || events == 0 || events > (size_t(1) << 31) || photons > (size_t(1) << 31)) {
    fail(error, error_size, "Invalid proposal sample capacity");
    return 1;
  }
  if (!ok(hipSetDevice(p->device), error, error_size, "hipSetDevice")) return 1;
  void **buffers[] = {(void **)&p->features, (void **)&p->directions, (void **)&p->weights,
                      (void **)&p->cumulative, (void **)&p->photon, (void **)&p->totals, (void **)&p->starts,
                      (void **)&p->path_cumulative, (void **)&p->event_path, (void **)&p->path_totals};
  for (void **b : buffers) {
    (void)hipFree(*b);
    *b = nullptr;
  }
  p->capacity = p->count = p->photon_capacity = 0;
  p->weight_sum = 0.0;
  if (!ok(hipMalloc(&p->features, events * kInputs * sizeof(float)), error, error_size, "allocate samples") ||
      !ok(hipMalloc(&p->directions, events * 3 * sizeof(float)), error, error_size, "allocate samples") ||
      !ok(hipMalloc(&p->weights, events * sizeof(float)), error, error_size, "allocate samples") ||
      !ok(hipMalloc(&p->cumulative, events * sizeof(float)), error, error_size, "allocate samples") ||
      !ok(hipMalloc(&p->photon, events * sizeof(uint32_t)), error, error_size, "allocate samples") ||
      (photons && !ok(hipMalloc(&p->totals, photons * sizeof(float)), error, error_size, "allocate totals")) ||
      (photons && !ok(hipMalloc(&p->starts, photons * sizeof(uint32_t)), error, error_size, "allocate log starts")) ||
      (photons && !ok(hipMalloc(&p->path_totals, photons * sizeof(float)), error, error_size, "allocate path totals")) ||
      !ok(hipMalloc(&p->path_cumulative, events * sizeof(float)), error, error_size, "allocate path samples") ||
      !ok(hipMalloc(&p->event_path, events * sizeof(float)), error, error_size, "allocate path samples") ||
      !ok(hipMemset(p->log_count, 0, sizeof(unsigned)), error, error_size, "clear log counter"))
    return 1;
  p->capacity = events;
  p->photon_capacity = photons;
  return 0;
}

int monki_hip_proposal_set_samples(void *handle, const float *features, const float *directions,
                                   const float *weights, size_t count, char *error, size_t error_size) {
  Proposal *p = static_cast<Proposal *>(handle);
  if (!p || !features || !directions || !weights || count == 0) {
    fail(error, error_size, "Invalid proposal samples");
    return 1;
  }
  if (count > p->capacity && monki_hip_proposal_reserve(handle, count, 0, error, error_size)) return 1;
  if (!ok(hipSetDevice(p->device), error, error_size, "hipSetDevice") ||
      !ok(hipMemcpy(p->features, features, count * kInputs * sizeof(float), hipMemcpyHostToDevice), error,
          error_size, "upload samples") ||
      !ok(hipMemcpy(p->directions, directions, count * 3 * sizeof(float), hipMemcpyHostToDevice), error,
          error_size, "upload samples") ||
      !ok(hipMemcpy(p->weights, weights, count * sizeof(float), hipMemcpyHostToDevice), error, error_size,
          "upload samples"))
    return 1;
  double sum = 0.0;
  for (size_t i = 0; i < count; ++i) sum += weights[i];
  p->count = count;
  p->weight_sum = sum;
  return 0;
}

int monki_hip_proposal_set_path_target(void *handle, const float *layer_weights, size_t layers, float mix,
                                       char *error, size_t error_size) {
  Proposal *p = static_cast<Proposal *>(handle);
  if (!p || !(mix >= 0.0f && mix <= 1.0f) || (mix > 0.0f && (!layer_weights || layers == 0 || layers > 256))) {
    fail(error, error_size, "Invalid path target (mix in [0, 1], 1-256 layer weights)");
    return 1;
  }
  if (!ok(hipSetDevice(p->device), error, error_size, "hipSetDevice")) return 1;
  (void)hipFree(p->layer_weight);
  p->layer_weight = nullptr;
  p->path_mix = mix;
  p->layer_count = mix > 0.0f ? uint32_t(layers) : 0u;
  if (mix > 0.0f &&
      (!ok(hipMalloc(&p->layer_weight, layers * sizeof(float)), error, error_size, "allocate layer weights") ||
       !ok(hipMemcpy(p->layer_weight, layer_weights, layers * sizeof(float), hipMemcpyHostToDevice), error,
           error_size, "upload layer weights")))
    return 1;
  return 0;
}

int monki_hip_proposal_samples(void *handle, uint64_t *count, double *weight_sum) {
  Proposal *p = static_cast<Proposal *>(handle);
  if (!p) return 1;
  if (count) *count = p->count;
  if (weight_sum) *weight_sum = p->weight_sum;
  return 0;
}

int monki_hip_proposal_read_samples(void *handle, float *features, float *directions, float *weights,
                                    size_t count) {
  Proposal *p = static_cast<Proposal *>(handle);
  if (!p || count > p->count || hipSetDevice(p->device) != hipSuccess) return 1;
  if (features && hipMemcpy(features, p->features, count * kInputs * sizeof(float), hipMemcpyDeviceToHost))
    return 1;
  if (directions && hipMemcpy(directions, p->directions, count * 3 * sizeof(float), hipMemcpyDeviceToHost))
    return 1;
  if (weights && hipMemcpy(weights, p->weights, count * sizeof(float), hipMemcpyDeviceToHost)) return 1;
  return 0;
}

int monki_hip_proposal_gradients(void *handle, size_t count, float *grad, double *loss, char *error,
                                 size_t error_size) {
  Proposal *p = static_cast<Proposal *>(handle);
  if (!p || !grad || count == 0 || count > p->count) {
    fail(error, error_size, "Invalid gradient request");
    return 1;
  }
  uint32_t *indices = nullptr;
  std::vector<uint32_t> host(count);
  for (size_t i = 0; i < count; ++i) host[i] = uint32_t(i);
  double value = 0.0;
  const bool good =
      ok(hipSetDevice(p->device), error, error_size, "hipSetDevice") &&
      ok(hipMalloc(&indices, count * sizeof(uint32_t)), error, error_size, "allocate indices") &&
      ok(hipMemcpy(indices, host.data(), count * sizeof(uint32_t), hipMemcpyHostToDevice), error, error_size,
         "upload indices") &&
      gradients(p, indices, uint32_t(count), 0, p->loss, error, error_size) &&
      ok(hipMemcpyAsync(&value, p->loss, sizeof(double), hipMemcpyDeviceToHost, p->stream), error, error_size,
         "copy loss") &&
      ok(hipMemcpyAsync(grad, p->grad, kParams * sizeof(float), hipMemcpyDeviceToHost, p->stream), error,
         error_size, "copy gradient") &&
      ok(hipStreamSynchronize(p->stream), error, error_size, "gradient");
  (void)hipFree(indices);
  if (loss) *loss = value;
  return good ? 0 : 1;
}

int monki_hip_proposal_train(void *handle, uint32_t steps, uint32_t batch, float learning_rate,
                             uint64_t seed, double *losses, char *error, size_t error_size) {
  Proposal *p = static_cast<Proposal *>(handle);
  if (!p || steps == 0 || batch == 0 || batch > (1u << 24) || p->count == 0 || !(learning_rate > 0.0f)) {
    fail(error, error_size, "Invalid training request (no samples?)");
    return 1;
  }
  if (!ok(hipSetDevice(p->device), error, error_size, "hipSetDevice")) return 1;
  // Each step's loss lands in device memory; one copy after the last step (no per-step wait).
  double *device_losses = nullptr;
  if (!ok(hipMalloc(&device_losses, size_t(steps) * sizeof(double)), error, error_size, "allocate losses"))
    return 1;
  const int status = train_steps(p, steps, batch, learning_rate, seed, device_losses, error, error_size) &&
                             (!losses || ok(hipMemcpyAsync(losses, device_losses, size_t(steps) * sizeof(double),
                                                           hipMemcpyDeviceToHost, p->stream),
                                            error, error_size, "copy losses")) &&
                             ok(hipStreamSynchronize(p->stream), error, error_size, "train")
                         ? 0
                         : 1;
  (void)hipStreamSynchronize(p->stream);
  (void)hipFree(device_losses);
  return status;
}

int monki_hip_proposal_forward(void *handle, const float *features, size_t count, float *out, char *error,
                               size_t error_size) {
  Proposal *p = static_cast<Proposal *>(handle);
  if (!p || !features || !out || count == 0 || count > (1u << 24)) {
    fail(error, error_size, "Invalid forward request");
    return 1;
  }
  float *din = nullptr, *dout = nullptr;
  const bool good =
      ok(hipSetDevice(p->device), error, error_size, "hipSetDevice") &&
      ok(hipMalloc(&din, count * kInputs * sizeof(float)), error, error_size, "allocate inputs") &&
      ok(hipMalloc(&dout, count * kOutputs * sizeof(float)), error, error_size, "allocate outputs") &&
      ok(hipMemcpy(din, features, count * kInputs * sizeof(float), hipMemcpyHostToDevice), error, error_size,
         "upload inputs") &&
      [&] {
        hipLaunchKernelGGL(forward_kernel, dim3((count + 63) / 64), dim3(64), 0, p->stream, view(p), din,
                           uint32_t(count), dout);
        return ok(hipGetLastError(), error, error_size, "forward launch");
      }() &&
      ok(hipStreamSynchronize(p->stream), error, error_size, "forward") &&
      ok(hipMemcpy(out, dout, count * kOutputs * sizeof(float), hipMemcpyDeviceToHost), error, error_size,
         "copy outputs");
  (void)hipFree(din);
  (void)hipFree(dout);
  return good ? 0 : 1;
}

void monki_hip_proposal_close(void *handle) { release(static_cast<Proposal *>(handle)); }

}  // extern "C"
