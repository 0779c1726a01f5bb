// pt_mlp.hip -- the radiance-cache network on gfx950 BF16 MFMA (no
// counterpart in the reference).
//
// Network: 96 inputs -> 64 -> 64 -> 64 -> 64 -> 3, ReLU on the hidden
// layers, 18,883 parameters.  FP32 master weights; BF16 operands of
// v_mfma_f32_16x16x32_bf16 with FP32 accumulation.
//
// Operand layouts of the 16x16x32 MFMA (lane l, q = l % 16, g = l / 16):
//   A (16 x 32): lane holds A[q][8g .. 8g+7]       (one 16-byte register set)
//   B (32 x 16): lane holds B[8g .. 8g+7][q]
//   C (16 x 16): lane holds C[4g .. 4g+3][q]
// The network runs transposed: Z = W X with one sample per column, so a
// lane owns one sample q.  A hidden layer's output tile t holds units
// 16t + 4g + r in lane (q, g); the next layer's B operand at k-step s takes
// slot j of lane (q, g) from tile 2s + j/4, row 4g + j%4.  Packing the next
// weight matrix with its input (k) index permuted the same way,
//   k = 32s + 8g + j  <->  unit 16 (2s + j/4) + 4g + j%4,
// makes every layer feed the next from registers, with no LDS transposes.
// The backward pass runs W^T the same way.  Weight gradients are
// dZ X^T over the batch: a separate split-K MFMA GEMM over unit-major BF16
// copies of the activations and output gradients.
#include <hip/hip_runtime.h>

#include <algorithm>

#include "pt_device.hpp"
#include "pt_kernels.hpp"

namespace pt {

typedef __attribute__((ext_vector_type(8))) __bf16 bf16x8;
typedef __attribute__((ext_vector_type(4))) float f32x4;

constexpr int MB = 256;  // block size of the network kernels (4 waves)
constexpr int W_LANES = 64;

// FP32 -> BF16, round to nearest even (NaN kept quiet)
PTD uint16_t to_bf16(float f) {
  uint32_t u = __float_as_uint(f);
  if ((u & 0x7FFFFFFFu) > 0x7F800000u) return (uint16_t)((u >> 16) | 0x40u);
  u += 0x7FFFu + ((u >> 16) & 1u);
  return (uint16_t)(u >> 16);
}
PTD float from_bf16(uint16_t h) { return __uint_as_float((uint32_t)h << 16); }

PTD bf16x8 as_bf16x8(uint4 v) { return __builtin_bit_cast(bf16x8, v); }
PTD uint4 pack8(const float* v) {
  uint4 r;
  r.x = (uint32_t)to_bf16(v[0]) | ((uint32_t)to_bf16(v[1]) << 16);
  r.y = (uint32_t)to_bf16(v[2]) | ((uint32_t)to_bf16(v[3]) << 16);
  r.z = (uint32_t)to_bf16(v[4]) | ((uint32_t)to_bf16(v[5]) << 16);
  r.w = (uint32_t)to_bf16(v[6]) | ((uint32_t)to_bf16(v[7]) << 16);
  return r;
}
PTD f32x4 mfma(uint4 a, uint4 b, f32x4 c) {
  return __builtin_amdgcn_mfma_f32_16x16x32_bf16(as_bf16x8(a), as_bf16x8(b), c, 0, 0, 0);
}

// ------------------------------------------------------------ packing
// Packed operand arrays, one uint4 (8 BF16) per lane per (tile, k-step):
//   forward  layer L: [T_L][S_L][64 lanes]  T = 4,4,4,4,1  S = 3,2,2,2,2
//   backward layer L = 2..5 (W^T): [4][S'][64], S' = 2,2,2,1
constexpr int FW_OFF[5] = {0, 768, 1280, 1792, 2304};  // uint4 offsets, total 2432
constexpr int FW_TOTAL = 2432;
constexpr int BW_OFF[5] = {0, 0, 512, 1024, 1536};     // layer 2 at 0, ..., layer 5 at 1536
constexpr int BW_TOTAL = 1792;
constexpr int L_IN[5] = {MLP_IN, MLP_H, MLP_H, MLP_H, MLP_H};
constexpr int L_OUT[5] = {MLP_H, MLP_H, MLP_H, MLP_H, MLP_OUT};
constexpr int W_OFF[5] = {0, 6208, 10368, 14528, 18688};  // weights; bias follows each
constexpr int B_OFF[5] = {6144, 10304, 14464, 18624, 18880};

// k-step position of input unit u of a hidden-input layer
PTD void hidden_k(int u, int& s, int& g, int& j) {
  int t = u >> 4, w = u & 15;
  s = t >> 1;
  j = ((t & 1) << 2) | (w & 3);
  g = w >> 2;
}

// Writes the packed BF16 operands and the FP32 biases of one parameter.
PTD void pack_param(int p, float val, uint16_t* fw, uint16_t* bw, float* bias) {
  int L = 0;
  while (L < 4 && p >= W_OFF[L + 1]) L++;
  if (p >= B_OFF[L]) {  // bias
    bias[L * 64 + (p - B_OFF[L])] = val;
    return;
  }
  int o = (p - W_OFF[L]) / L_IN[L], i = (p - W_OFF[L]) % L_IN[L];
  uint16_t h = to_bf16(val);
  // forward A = W: row o (tile o/16, m = o%16), k from input unit i
  int s, g, j;
  if (L == 0) {
    s = i >> 5;
    g = (i & 31) >> 3;
    j = i & 7;
  } else {
    hidden_k(i, s, g, j);
  }
  int S = (L == 0) ? 3 : 2;
  int lane = (o & 15) + 16 * g;
  fw[8 * (FW_OFF[L] + ((o >> 4) * S + s) * W_LANES + lane) + j] = h;
  if (L == 0) return;
  // backward A = W^T: row i (tile i/16, m = i%16), k from output unit o
  int sb, gb, jb, Sb;
  if (L == 4) {
    sb = 0;
    gb = o >> 2;
    jb = o & 3;
    Sb = 1;
  } else {
    hidden_k(o, sb, gb, jb);
    Sb = 2;
  }
  int laneb = (i & 15) + 16 * gb;
  bw[8 * (BW_OFF[L] + ((i >> 4) * Sb + sb) * W_LANES + laneb) + jb] = h;
}

__global__ void k_mlp_pack(const float* params, uint16_t* fw, uint16_t* bw, float* bias) {
  for (int p = blockIdx.x * MB + threadIdx.x; p < MLP_PARAMS; p += gridDim.x * MB) pack_param(p, params[p], fw, bw, bias);
}

// ------------------------------------------------------------ forward
struct Fwd {
  uint4 x[3];      // input, B format
  uint4 h[4][2];   // hidden outputs 1..4, B format (post ReLU)
  f32x4 y;         // output tile (lanes g = 0 hold rows 0..3)
};

// accumulators of one layer -> ReLU -> next B operand (2 k-steps)
PTD void relu_pack(const f32x4* acc, uint4* out) {
  for (int s = 0; s < 2; s++) {
    float v[8];
    for (int r = 0; r < 4; r++) {
      v[r] = fmaxf(acc[2 * s][r], 0.f);
      v[4 + r] = fmaxf(acc[2 * s + 1][r], 0.f);
    }
    out[s] = pack8(v);
  }
}

PTD void forward(const uint4* fw, const float* bias, int g, Fwd& f) {
  const int lane = threadIdx.x & 63;
  f32x4 acc[4];
  for (int t = 0; t < 4; t++) {
    for (int r = 0; r < 4; r++) acc[t][r] = bias[16 * t + 4 * g + r];
    for (int s = 0; s < 3; s++) acc[t] = mfma(fw[FW_OFF[0] + (t * 3 + s) * W_LANES + lane], f.x[s], acc[t]);
  }
  relu_pack(acc, f.h[0]);
  for (int L = 1; L < 4; L++) {
    for (int t = 0; t < 4; t++) {
      for (int r = 0; r < 4; r++) acc[t][r] = bias[L * 64 + 16 * t + 4 * g + r];
      for (int s = 0; s < 2; s++)
        acc[t] = mfma(fw[FW_OFF[L] + (t * 2 + s) * W_LANES + lane], f.h[L - 1][s], acc[t]);
    }
    relu_pack(acc, f.h[L]);
  }
  for (int r = 0; r < 4; r++) f.y[r] = (4 * g + r < MLP_OUT) ? bias[4 * 64 + 4 * g + r] : 0.f;
  for (int s = 0; s < 2; s++) f.y = mfma(fw[FW_OFF[4] + s * W_LANES + lane], f.h[3][s], f.y);
}

// weights and biases into LDS (cooperative)
PTD void stage(uint4* dst, const uint4* src, int n) {
  for (int i = threadIdx.x; i < n; i += MB) dst[i] = src[i];
}

// Inference on n rows (BF16 rows of 96 features = 12 uint4 each).  With
// `seg`, the rows are positions of a wave-segmented record set (segments of
// a grid whose stride is `stride` positions, counts seg[w]): a 16-row tile
// lies in one 64-position run of one segment and only rows below the
// segment's count are evaluated.
__global__ __launch_bounds__(MB) void k_mlp_infer(const uint4* feats, uint32_t n, const uint4* fw_g, const float* bias_g,
                                                  float* out, const uint32_t* seg, uint32_t stride) {
  __shared__ uint4 fw[FW_TOTAL];
  __shared__ float bias[4 * 64 + 16];
  stage(fw, fw_g, FW_TOTAL);
  for (int i = threadIdx.x; i < 4 * 64 + 16; i += MB) bias[i] = bias_g[i];
  __syncthreads();
  const int lane = threadIdx.x & 63, q = lane & 15, g = lane >> 4;
  const int wave = blockIdx.x * (MB / 64) + (threadIdx.x >> 6);
  for (uint32_t base = (uint32_t)wave * 16; base < n; base += gridDim.x * (MB / 64) * 16) {
    uint32_t row = base + q;
    bool valid = row < n;
    if (seg) {
      const uint32_t run = base / stride, rem = base - run * stride;
      const uint32_t live = __builtin_amdgcn_readfirstlane(seg[rem >> 6]);
      const uint32_t j0 = run * 64u + (rem & 63u);
      if (j0 >= live) continue;  // the whole tile is past the segment's count
      valid = valid && j0 + (uint32_t)q < live;
    }
    Fwd f;
    for (int s = 0; s < 3; s++) f.x[s] = valid ? feats[(size_t)row * 12 + 4 * s + g] : make_uint4(0, 0, 0, 0);
    forward(fw, bias, g, f);
    if (valid && g == 0)
      for (int r = 0; r < MLP_OUT; r++) out[(size_t)row * MLP_OUT + r] = f.y[r];
  }
}

// ------------------------------------------------------------ backward
// unit-major BF16 copies: X_L[unit][sample] of the inputs of layer L and
// dZ_L[unit][sample] of its output gradients (layer 5 padded to 16 rows)
constexpr int X_ROWS[5] = {96, 64, 64, 64, 64};
constexpr int X_ROW0[5] = {0, 96, 160, 224, 288};  // total 352 rows
constexpr int DZ_ROWS[5] = {64, 64, 64, 64, 16};
constexpr int DZ_ROW0[5] = {0, 64, 128, 192, 256};  // total 272 rows
constexpr int DB_WORDS = 4 * 64 + MLP_OUT;

PTD void store_b(uint16_t* buf, size_t ld, int row0, uint4 v, int s, int g, uint32_t col, bool hidden) {
  // the 8 BF16 of a B-format register are units (s, g, j) -> unit index
  uint16_t h[8] = {(uint16_t)(v.x & 0xFFFF), (uint16_t)(v.x >> 16), (uint16_t)(v.y & 0xFFFF), (uint16_t)(v.y >> 16),
                   (uint16_t)(v.z & 0xFFFF), (uint16_t)(v.z >> 16), (uint16_t)(v.w & 0xFFFF), (uint16_t)(v.w >> 16)};
  for (int j = 0; j < 8; j++) {
    int u = hidden ? (16 * (2 * s + (j >> 2)) + 4 * g + (j & 3)) : (32 * s + 8 * g + j);
    buf[(size_t)(row0 + u) * ld + col] = h[j];
  }
}

// sum over the 16 lanes that share g (lanes q = 0..15 of a row group)
PTD float sum16(float v) {
  for (int off = 1; off < 16; off <<= 1) v += __shfl_xor(v, off);
  return v;
}

// Forward, loss and backward through the activations for 16 samples per
// wave; writes X and dZ (unit-major BF16, ld = padded batch), per-wave bias
// gradient sums and loss sums.
__global__ __launch_bounds__(MB) void k_mlp_fwd_bwd(const uint4* feats, const float* targets, uint32_t n, uint32_t ld,
                                                    const uint4* fw_g, const uint4* bw_g, const float* bias_g,
                                                    float inv_n, uint16_t* X, uint16_t* dZ, float* db_partial,
                                                    float* loss_partial) {
  __shared__ uint4 fw[FW_TOTAL];
  __shared__ uint4 bw[BW_TOTAL];
  __shared__ float bias[4 * 64 + 16];
  stage(fw, fw_g, FW_TOTAL);
  stage(bw, bw_g, BW_TOTAL);
  for (int i = threadIdx.x; i < 4 * 64 + 16; i += MB) bias[i] = bias_g[i];
  __syncthreads();
  const int lane = threadIdx.x & 63, q = lane & 15, g = lane >> 4;
  const int wave = blockIdx.x * (MB / 64) + (threadIdx.x >> 6);
  const uint32_t base = (uint32_t)wave * 16;
  if (base >= ld) return;  // whole wave past the padded batch (no barrier follows)
  const uint32_t col = base + q;
  const bool valid = col < n;
  Fwd f;
  for (int s = 0; s < 3; s++) f.x[s] = valid ? feats[(size_t)col * 12 + 4 * s + g] : make_uint4(0, 0, 0, 0);
  forward(fw, bias, g, f);
  for (int s = 0; s < 3; s++) store_b(X, ld, X_ROW0[0], f.x[s], s, g, col, false);
  for (int L = 0; L < 4; L++)
    for (int s = 0; s < 2; s++) store_b(X, ld, X_ROW0[L + 1], f.h[L][s], s, g, col, true);
  // relative L2 loss and its gradient (cache_loss_terms)
  float dz5[4] = {0.f, 0.f, 0.f, 0.f};
  float loss = 0.f;
  for (int r = 0; r < 4; r++) {
    if (g == 0 && r < MLP_OUT && valid) {
      float l, d;
      cache_loss_terms(f.y[r], targets[(size_t)col * MLP_OUT + r], l, d);
      loss += l * inv_n;
      dz5[r] = d * inv_n;
    }
  }
  for (int r = 0; r < 4; r++) dZ[(size_t)(DZ_ROW0[4] + 4 * g + r) * ld + col] = to_bf16(dz5[r]);
  float lsum = sum16(loss);
  float db[4][4][4];  // [layer 1..4 -> 0..3][tile][r]
  // layer 5 backward: dH4 = W5^T dZ5 (k: units 4g + j, j < 4)
  float b5[8] = {dz5[0], dz5[1], dz5[2], dz5[3], 0.f, 0.f, 0.f, 0.f};
  uint4 B = pack8(b5);
  f32x4 acc[4];
  for (int t = 0; t < 4; t++) {
    f32x4 z = {0.f, 0.f, 0.f, 0.f};
    acc[t] = mfma(bw[BW_OFF[4] + t * W_LANES + lane], B, z);
  }
  for (int L = 3; L >= 0; L--) {
    // dZ_L = dH_L * relu'(h_L) (h_L = output of layer L, B format)
    float v[2][8];
    for (int s = 0; s < 2; s++) {
      uint4 hv = f.h[L][s];
      uint32_t w[4] = {hv.x, hv.y, hv.z, hv.w};
      for (int j = 0; j < 8; j++) {
        uint16_t hb = (uint16_t)((w[j >> 1] >> (16 * (j & 1))) & 0xFFFF);
        float dh = acc[2 * s + (j >> 2)][j & 3];
        v[s][j] = (from_bf16(hb) > 0.f) ? dh : 0.f;
      }
    }
    uint4 dzp[2] = {pack8(v[0]), pack8(v[1])};
    for (int s = 0; s < 2; s++) store_b(dZ, ld, DZ_ROW0[L], dzp[s], s, g, col, true);
    // bias gradient: sum over the 16 samples, kept per wave
    for (int t = 0; t < 4; t++)
      for (int r = 0; r < 4; r++) db[L][t][r] = sum16(from_bf16(to_bf16(v[t >> 1][4 * (t & 1) + r])));
    if (L == 0) break;
    // dH_{L-1} = W_L^T dZ_L
    for (int t = 0; t < 4; t++) {
      f32x4 z = {0.f, 0.f, 0.f, 0.f};
      for (int s = 0; s < 2; s++) z = mfma(bw[BW_OFF[L] + (t * 2 + s) * W_LANES + lane], dzp[s], z);
      acc[t] = z;
    }
  }
  // layer 5 bias gradient: the lanes g = 0 hold dZ5 rows 0..2
  float b5s[3];
  for (int r = 0; r < MLP_OUT; r++) b5s[r] = sum16(g == 0 ? from_bf16(to_bf16(dz5[r])) : 0.f);
  float* out = db_partial + (size_t)wave * DB_WORDS;
  if (q == 0)
    for (int L = 0; L < 4; L++)
      for (int t = 0; t < 4; t++)
        for (int r = 0; r < 4; r++) out[L * 64 + 16 * t + 4 * g + r] = db[L][t][r];
  if (lane == 0) {
    for (int r = 0; r < MLP_OUT; r++) out[4 * 64 + r] = b5s[r];
    loss_partial[wave] = lsum;
  }
}

// ------------------------------------------------------------ weight grads
// One wave per (layer, output tile, input tile, k chunk): dW tile =
// sum over the chunk's samples of dZ[16 to + m][k] X[16 ti + n][k].
struct TileTask {
  int L, to, ti;
};
constexpr int N_TILES = 24 + 16 * 3 + 4;  // 76

PTD TileTask tile_task(int i) {
  TileTask t;
  if (i < 24) { t.L = 0; t.to = i / 6; t.ti = i % 6; return t; }
  i -= 24;
  if (i < 48) { t.L = 1 + i / 16; t.to = (i % 16) / 4; t.ti = i % 4; return t; }
  i -= 48;
  t.L = 4; t.to = 0; t.ti = i;
  return t;
}

__global__ __launch_bounds__(MB) void k_mlp_wgrad(const uint16_t* X, const uint16_t* dZ, uint32_t ld, uint32_t chunk,
                                                  float* dw_partial) {
  const int lane = threadIdx.x & 63, q = lane & 15, g = lane >> 4;
  const int wave = blockIdx.x * (MB / 64) + (threadIdx.x >> 6);
  const uint32_t nchunks = (ld + chunk - 1) / chunk;
  if ((uint32_t)wave >= N_TILES * nchunks) return;
  const int tile = wave % N_TILES;
  const uint32_t c = wave / N_TILES;
  const TileTask T = tile_task(tile);
  const uint16_t* a_row = dZ + (size_t)(DZ_ROW0[T.L] + 16 * T.to + q) * ld;
  const uint16_t* b_row = X + (size_t)(X_ROW0[T.L] + 16 * T.ti + q) * ld;
  f32x4 acc = {0.f, 0.f, 0.f, 0.f};
  const uint32_t k0 = c * chunk, k1 = min(ld, k0 + chunk);
  for (uint32_t k = k0; k < k1; k += 32) {
    uint4 a = *(const uint4*)(a_row + k + 8 * g);
    uint4 b = *(const uint4*)(b_row + k + 8 * g);
    acc = mfma(a, b, acc);
  }
  // C[m = 4g + r][n = q]: output unit 16 to + 4g + r, input unit 16 ti + q
  float* out = dw_partial + (size_t)c * MLP_PARAMS;
  for (int r = 0; r < 4; r++) {
    int o = 16 * T.to + 4 * g + r, i = 16 * T.ti + q;
    if (o < L_OUT[T.L] && i < L_IN[T.L]) out[W_OFF[T.L] + o * L_IN[T.L] + i] = acc[r];
  }
}

// ------------------------------------------------------------ Adam
// Bias gradients: one wave per bias unit sums the per-wave partials (lane l
// takes partials l, l + 64, ...; then a fixed shuffle tree), into db_sum.
__global__ __launch_bounds__(MB) void k_mlp_bias_reduce(const float* db_partial, uint32_t nwaves, float* db_sum) {
  const int unit = blockIdx.x * (MB / 64) + (threadIdx.x >> 6), lane = threadIdx.x & 63;
  if (unit >= DB_WORDS) return;
  float v = 0.f;
  for (uint32_t w = lane; w < nwaves; w += 64) v += db_partial[(size_t)w * DB_WORDS + unit];
  for (int off = 32; off > 0; off >>= 1) v += __shfl_xor(v, off);
  if (lane == 0) db_sum[unit] = v;
}
// Sums the weight partials in a fixed order, then one Adam step (Kingma &
// Ba 2015) on the FP32 master weights, and repacks the operands.
__global__ __launch_bounds__(MB) void k_mlp_adam(const float* dw_partial, uint32_t nchunks, const float* db_sum,
                                                 float* params, float* m1, float* m2, float* grad_out, float lr,
                                                 float beta1, float beta2, float eps, float bc1, float bc2,
                                                 uint16_t* fw, uint16_t* bw, float* bias) {
  for (int p = blockIdx.x * MB + threadIdx.x; p < MLP_PARAMS; p += gridDim.x * MB) {
    int L = 0;
    while (L < 4 && p >= W_OFF[L + 1]) L++;
    float gsum = 0.f;
    if (p >= B_OFF[L]) {
      gsum = db_sum[L * 64 + (p - B_OFF[L])];
    } else {
      for (uint32_t c = 0; c < nchunks; c++) gsum += dw_partial[(size_t)c * MLP_PARAMS + p];
    }
    grad_out[p] = gsum;
    float a = beta1 * m1[p] + (1.f - beta1) * gsum;
    float b = beta2 * m2[p] + (1.f - beta2) * (gsum * gsum);
    m1[p] = a;
    m2[p] = b;
    float mh = a / bc1, vh = b / bc2;
    float val = params[p] - lr * mh / (sqrtf(vh) + eps);
    params[p] = val;
    pack_param(p, val, fw, bw, bias);
  }
}

// ------------------------------------------------------------ launchers
void launch_mlp_pack(const float* params, uint16_t* fw, uint16_t* bw, float* bias, hipStream_t st) {
  hipLaunchKernelGGL(k_mlp_pack, dim3((MLP_PARAMS + MB - 1) / MB), dim3(MB), 0, st, params, fw, bw, bias);
}
void launch_mlp_infer(const uint16_t* feats, uint32_t n, const uint16_t* fw, const float* bias, float* out, int grid,
                      hipStream_t st, const uint32_t* seg, uint32_t stride) {
  hipLaunchKernelGGL(k_mlp_infer, dim3(grid), dim3(MB), 0, st, (const uint4*)feats, n, (const uint4*)fw, bias, out, seg,
                     stride);
}
uint32_t mlp_padded(uint32_t n) { return (n + 31) & ~31u; }
uint32_t mlp_fwd_bwd_waves(uint32_t n) { return mlp_padded(n) / 16; }
uint32_t mlp_chunks(uint32_t n) { return (mlp_padded(n) + MLP_CHUNK - 1) / MLP_CHUNK; }
void launch_mlp_train(const uint16_t* feats, const float* targets, uint32_t n, const uint16_t* fw, const uint16_t* bw,
                      const float* bias_in, uint16_t* X, uint16_t* dZ, float* db_partial, float* loss_partial,
                      float* dw_partial, float* params, float* m1, float* m2, float* grad_out, float lr,
                      uint32_t step, uint16_t* fw_out, uint16_t* bw_out, float* bias_out, hipStream_t st) {
  const uint32_t ld = mlp_padded(n);
  const uint32_t waves = mlp_fwd_bwd_waves(n);
  hipLaunchKernelGGL(k_mlp_fwd_bwd, dim3((waves + 3) / 4), dim3(MB), 0, st, (const uint4*)feats, targets, n, ld,
                     (const uint4*)fw, (const uint4*)bw, bias_in, 1.f / (float)n, X, dZ, db_partial, loss_partial);
  const uint32_t nch = mlp_chunks(n);
  const uint32_t wg = N_TILES * nch;
  hipLaunchKernelGGL(k_mlp_wgrad, dim3((wg + 3) / 4), dim3(MB), 0, st, (const uint16_t*)X, (const uint16_t*)dZ, ld,
                     (uint32_t)MLP_CHUNK, dw_partial);
  // the bias sums go after the per-wave partials (db_partial has a spare
  // row of DB_WORDS floats past the last wave)
  float* db_sum = db_partial + (size_t)waves * DB_WORDS;
  hipLaunchKernelGGL(k_mlp_bias_reduce, dim3((DB_WORDS + MB / 64 - 1) / (MB / 64)), dim3(MB), 0, st,
                     (const float*)db_partial, waves, db_sum);
  const float b1 = 0.9f, b2 = 0.999f;
  const float bc1 = 1.f - powf(b1, (float)step), bc2 = 1.f - powf(b2, (float)step);
  hipLaunchKernelGGL(k_mlp_adam, dim3((MLP_PARAMS + MB - 1) / MB), dim3(MB), 0, st, (const float*)dw_partial, nch,
                     (const float*)db_sum, params, m1, m2, grad_out, lr, b1, b2, 1e-8f, bc1, bc2, fw_out, bw_out,
                     bias_out);
}

}  // namespace pt
