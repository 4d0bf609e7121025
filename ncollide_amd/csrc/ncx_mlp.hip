// ncx_mlp.hip -- learned initial search direction for GJK (no reference counterpart).
//
// A 44 -> 64 -> 64 -> 3 MLP (ReLU) predicts contact_support_map_support_map_with_params'
// `init_dir` (src/query/contact/contact_support_map_support_map.rs:38-61) from
//   * the relative pose of shape 2 in shape 1's frame: R1^T R2 (9) and R1^T (t2 - t1) (3),
//   * 16 support-function samples h(u_k) = max_p p . u_k of each shape along fixed
//     Fibonacci directions of its local frame (16 + 16).
// The prediction is in shape 1's frame; it is rotated to world and normalised.
//
// Arithmetic: FP16 operands, FP32 accumulation on the matrix cores
// (v_mfma_f32_16x16x16f16).  Every layer computes transposed activations
// H^T = W . X^T, so the accumulator layout of one layer (lane l holds units
// 4*(l/16)+t of sample l%16) is exactly the B-operand layout of the next and the
// forward pass chains in registers.  Weight gradients contract over samples and
// go through one LDS transpose per tile.  Master weights are FP32 (Adam).
#include "ncx_internal.h"
#include <random>

#define MLP_IN 44
#define MLP_INP 48  // padded input width (3 K-steps of 16)
#define MLP_H 64
#define MLP_OUT 3
#define MLP_OUTP 16
#define MLP_DESC 16

// flat FP32 parameter layout (padded entries stay zero)
#define OFF_W1 0                              // [64][48]
#define OFF_B1 (OFF_W1 + MLP_H * MLP_INP)     // [64]
#define OFF_W2 (OFF_B1 + MLP_H)               // [64][64]
#define OFF_B2 (OFF_W2 + MLP_H * MLP_H)       // [64]
#define OFF_W3 (OFF_B2 + MLP_H)               // [16][64]
#define OFF_B3 (OFF_W3 + MLP_OUTP * MLP_H)    // [16]
#define MLP_NP (OFF_B3 + MLP_OUTP)            // 8336

typedef _Float16 h4 __attribute__((ext_vector_type(4)));
typedef float f4 __attribute__((ext_vector_type(4)));

struct ncx_mlp {
  ncx_ctx* ctx;
  DevBuf master;   // FP32 params
  DevBuf m, v;     // Adam moments
  DevBuf half;     // FP16 copies: W1h [64][48], W2h [64][64], W3h [16][64], W2Th [64][64], W3Th [64][16]
  DevBuf partial;  // per-block gradient partial sums
  DevBuf grad;     // reduced gradient
  DevBuf loss;     // [0] = loss sum
  int64_t step = 0;
  float beta1 = 0.9f, beta2 = 0.999f, eps = 1e-8f;
};

#define HOFF_W1 0
#define HOFF_W2 (HOFF_W1 + MLP_H * MLP_INP)
#define HOFF_W3 (HOFF_W2 + MLP_H * MLP_H)
#define HOFF_W2T (HOFF_W3 + MLP_OUTP * MLP_H)
#define HOFF_W3T (HOFF_W2T + MLP_H * MLP_H)
#define HALF_N (HOFF_W3T + MLP_H * MLP_OUTP)

__device__ __forceinline__ f4 mfma16(h4 a, h4 b, f4 c) { return __builtin_amdgcn_mfma_f32_16x16x16f16(a, b, c, 0, 0, 0); }

// ------------------------------------------------------------------ descriptors and features
// Fibonacci directions u_k, k = 0..15: z = 1 - (2k+1)/16, phi = k * pi (3 - sqrt 5).
__device__ __forceinline__ f3 fib_dir(int k) {
  float z = 1.f - (2.f * k + 1.f) / 16.f;
  float r = sqrtf(fmaxf(0.f, 1.f - z * z));
  float ph = (float)k * 2.39996322972865332f;
  return f3{r * cosf(ph), r * sinf(ph), z};
}

// h(u) = max over the shape of x . u in its local frame (support function of the local shape).
__global__ __launch_bounds__(256) void k_descriptors(const ncx_shape* __restrict__ shapes, const float* __restrict__ hx,
                                                      const float* __restrict__ hy, const float* __restrict__ hz,
                                                      size_t n, float* __restrict__ desc) {
  size_t gid = (size_t)blockIdx.x * blockDim.x + threadIdx.x;
  size_t i = gid / MLP_DESC;
  int k = (int)(gid % MLP_DESC);
  if (i >= n) return;
  ncx_shape s = shapes[i];
  f3 u = fib_dir(k);
  float h;
  switch (s.type) {
    case NCX_BALL: h = s.p0; break;
    case NCX_CUBOID: h = (s.p0 * fabsf(u.x) + s.p1 * fabsf(u.y)) + s.p2 * fabsf(u.z); break;
    case NCX_CAPSULE: h = s.p0 * fabsf(u.y) + s.p1; break;
    default: {
      h = -NCX_FLT_MAX;
      for (uint32_t v = 0; v < s.hull_count; ++v) {
        uint32_t o = s.hull_offset + v;
        h = fmaxf(h, dot(f3{hx[o], hy[o], hz[o]}, u));
      }
    }
  }
  desc[i * MLP_DESC + k] = h;
}

__device__ __forceinline__ void quat_mat(quat q, float M[3][3]) {
  float i = q.i, j = q.j, k = q.k, w = q.w;
  M[0][0] = 1.f - 2.f * (j * j + k * k); M[0][1] = 2.f * (i * j - w * k); M[0][2] = 2.f * (i * k + w * j);
  M[1][0] = 2.f * (i * j + w * k); M[1][1] = 1.f - 2.f * (i * i + k * k); M[1][2] = 2.f * (j * k - w * i);
  M[2][0] = 2.f * (i * k - w * j); M[2][1] = 2.f * (j * k + w * i); M[2][2] = 1.f - 2.f * (i * i + j * j);
}

// Features of pair k (FP16, 48 per pair, last 4 zero).
__device__ __forceinline__ void pair_features(const ncx_pose* poses, const float* desc, uint32_t a, uint32_t b,
                                              float f[MLP_INP]) {
  iso m1 = load_pose(poses, a), m2 = load_pose(poses, b);
  float R1[3][3], R2[3][3];
  quat_mat(m1.q, R1);
  quat_mat(m2.q, R2);
  for (int r = 0; r < 3; ++r)
    for (int c = 0; c < 3; ++c) f[3 * r + c] = R1[0][r] * R2[0][c] + R1[1][r] * R2[1][c] + R1[2][r] * R2[2][c];
  f3 d = m2.t - m1.t;
  for (int r = 0; r < 3; ++r) f[9 + r] = R1[0][r] * d.x + R1[1][r] * d.y + R1[2][r] * d.z;
  for (int k = 0; k < MLP_DESC; ++k) {
    f[12 + k] = desc[(size_t)a * MLP_DESC + k];
    f[28 + k] = desc[(size_t)b * MLP_DESC + k];
  }
  for (int k = MLP_IN; k < MLP_INP; ++k) f[k] = 0.f;
}

__global__ __launch_bounds__(256) void k_features(const ncx_pose* __restrict__ poses, const float* __restrict__ desc,
                                                   const uint32_t* __restrict__ ia, const uint32_t* __restrict__ ib,
                                                   size_t n, _Float16* __restrict__ out) {
  size_t k = (size_t)blockIdx.x * blockDim.x + threadIdx.x;
  if (k >= n) return;
  float f[MLP_INP];
  pair_features(poses, desc, ia[k], ib[k], f);
  for (int c = 0; c < MLP_INP; ++c) out[k * MLP_INP + c] = (_Float16)f[c];
}

// ------------------------------------------------------------------ forward tile
// One wave, 16 samples.  X rows are FP16 [sample][48].  Returns Z3^T fragment (units 4*(l/16)+t of
// sample l%16), and keeps the post-ReLU activations (FP16, accumulator layout) when `keep`.
struct Acts {
  h4 h1[4];  // H1^T tiles: units 16*mt + 4*(l/16) + t
  h4 h2[4];
  f4 z1[4];
  f4 z2[4];
};

__device__ __forceinline__ f4 mlp_tile_forward(const _Float16* __restrict__ Wh, const float* __restrict__ P,
                                               const _Float16* __restrict__ X, size_t row0, size_t nrows, Acts& A) {
  const int l = __lane_id();
  const int r = l & 15, g = l >> 4;
  // B operand of layer 1: X^T[k][sample], lane holds k = 16 s + 4 g + t of sample r
  h4 xb[3];
  size_t row = row0 + r;
  for (int s = 0; s < 3; ++s) {
    if (row < nrows) {
      const h4* p = (const h4*)(X + row * MLP_INP + 16 * s + 4 * g);
      xb[s] = *p;
    } else {
      xb[s] = h4{(_Float16)0, (_Float16)0, (_Float16)0, (_Float16)0};
    }
  }
  // layer 1
  for (int mt = 0; mt < 4; ++mt) {
    f4 acc = {0.f, 0.f, 0.f, 0.f};
    for (int s = 0; s < 3; ++s) {
      h4 a = *(const h4*)(Wh + HOFF_W1 + (16 * mt + r) * MLP_INP + 16 * s + 4 * g);
      acc = mfma16(a, xb[s], acc);
    }
    f4 z;
    h4 hh;
    for (int t = 0; t < 4; ++t) {
      float zz = acc[t] + P[OFF_B1 + 16 * mt + 4 * g + t];
      z[t] = zz;
      hh[t] = (_Float16)fmaxf(zz, 0.f);
    }
    A.z1[mt] = z;
    A.h1[mt] = hh;
  }
  // layer 2
  for (int mt = 0; mt < 4; ++mt) {
    f4 acc = {0.f, 0.f, 0.f, 0.f};
    for (int s = 0; s < 4; ++s) {
      h4 a = *(const h4*)(Wh + HOFF_W2 + (16 * mt + r) * MLP_H + 16 * s + 4 * g);
      acc = mfma16(a, A.h1[s], acc);
    }
    f4 z;
    h4 hh;
    for (int t = 0; t < 4; ++t) {
      float zz = acc[t] + P[OFF_B2 + 16 * mt + 4 * g + t];
      z[t] = zz;
      hh[t] = (_Float16)fmaxf(zz, 0.f);
    }
    A.z2[mt] = z;
    A.h2[mt] = hh;
  }
  // layer 3 (16 padded outputs)
  f4 acc = {0.f, 0.f, 0.f, 0.f};
  for (int s = 0; s < 4; ++s) {
    h4 a = *(const h4*)(Wh + HOFF_W3 + r * MLP_H + 16 * s + 4 * g);
    acc = mfma16(a, A.h2[s], acc);
  }
  for (int t = 0; t < 4; ++t) acc[t] += P[OFF_B3 + 4 * g + t];
  return acc;
}

// Forward over n rows (FP16 features); out = n x 3 FP32.  One wave per 16 rows, grid-stride.
__global__ __launch_bounds__(256) void k_mlp_forward(const _Float16* __restrict__ Wh, const float* __restrict__ P,
                                                      const _Float16* __restrict__ X, size_t n,
                                                      float* __restrict__ out) {
  const int l = __lane_id();
  const size_t wave = ((size_t)blockIdx.x * blockDim.x + threadIdx.x) / 64;
  const size_t nw = ((size_t)gridDim.x * blockDim.x) / 64;
  for (size_t row0 = wave * 16; row0 < n; row0 += nw * 16) {
    Acts A;
    f4 z3 = mlp_tile_forward(Wh, P, X, row0, n, A);
    size_t row = row0 + (l & 15);
    if ((l >> 4) == 0 && row < n) {
      out[row * 3 + 0] = z3[0];
      out[row * 3 + 1] = z3[1];
      out[row * 3 + 2] = z3[2];
    }
  }
}

// ------------------------------------------------------------------ backward
// Per-wave LDS tile: [unit 0..63][sample 0..15] FP16, used to transpose accumulator-layout tiles.
__device__ __forceinline__ void lds_put(_Float16* T, int unit0, h4 v) {
  const int l = __lane_id();
  const int r = l & 15, g = l >> 4;
  for (int t = 0; t < 4; ++t) T[(unit0 + 4 * g + t) * 16 + r] = v[t];
}
// A/B fragment with K = samples: lane reads units (unit0 + l%16), samples 4*(l/16)+j.
__device__ __forceinline__ h4 lds_get_k_samples(const _Float16* T, int unit0) {
  const int l = __lane_id();
  return *(const h4*)(T + (unit0 + (l & 15)) * 16 + 4 * (l >> 4));
}

#define BW_BLOCK 256
#define BW_WAVES (BW_BLOCK / 64)

// Gradient of sum_k |z3_k - y_k|^2 over the rows (not yet divided by n).  Each block writes its
// partial gradient (MLP_NP floats) to partial[block]; loss sum to partial_loss[block].
__global__ __launch_bounds__(BW_BLOCK) void k_mlp_backward(const _Float16* __restrict__ Wh,
                                                            const float* __restrict__ P,
                                                            const _Float16* __restrict__ X,
                                                            const float* __restrict__ Y, size_t n,
                                                            float* __restrict__ partial,
                                                            float* __restrict__ partial_loss) {
  __shared__ _Float16 tile[BW_WAVES][3][MLP_H * 16];  // H1^T, H2^T/dZ, X^T staging per wave
  __shared__ float red[MLP_NP];
  __shared__ float lred[BW_WAVES];
  const int l = __lane_id();
  const int w = threadIdx.x / 64;
  const int r = l & 15, g = l >> 4;
  const size_t wave = ((size_t)blockIdx.x * BW_BLOCK + threadIdx.x) / 64;
  const size_t nw = ((size_t)gridDim.x * BW_BLOCK) / 64;
  for (int i = threadIdx.x; i < MLP_NP; i += BW_BLOCK) red[i] = 0.f;
  // accumulators of weight gradients, accumulator layout
  f4 gW1[4][3], gW2[4][4], gW3[4];
  float gb1[4][4], gb2[4][4], gb3[4];
  for (int a = 0; a < 4; ++a) {
    for (int b = 0; b < 3; ++b) gW1[a][b] = f4{0.f, 0.f, 0.f, 0.f};
    for (int b = 0; b < 4; ++b) gW2[a][b] = f4{0.f, 0.f, 0.f, 0.f};
    gW3[a] = f4{0.f, 0.f, 0.f, 0.f};
    for (int t = 0; t < 4; ++t) { gb1[a][t] = 0.f; gb2[a][t] = 0.f; }
    gb3[a] = 0.f;
  }
  float lsum = 0.f;
  _Float16* T0 = tile[w][0];
  _Float16* T1 = tile[w][1];
  _Float16* T2 = tile[w][2];
  for (size_t row0 = wave * 16; row0 < n; row0 += nw * 16) {
    Acts A;
    f4 z3 = mlp_tile_forward(Wh, P, X, row0, n, A);
    size_t row = row0 + r;
    bool valid = row < n;
    // dZ3^T (accumulator layout): units 4g+t of sample r; only units 0..2 are outputs
    f4 d3;
    for (int t = 0; t < 4; ++t) {
      int u = 4 * g + t;
      float dd = 0.f;
      if (valid && u < MLP_OUT) {
        float e = z3[t] - Y[row * 3 + u];
        lsum += e * e;
        dd = 2.f * e;
      }
      d3[t] = dd;
    }
    h4 d3h;
    for (int t = 0; t < 4; ++t) d3h[t] = (_Float16)d3[t];
    // dH2^T = W3^T . dZ3^T ; dZ2 = dH2 * (Z2 > 0)
    h4 d2h[4];
    f4 d2[4];
    for (int mt = 0; mt < 4; ++mt) {
      h4 a = *(const h4*)(Wh + HOFF_W3T + (16 * mt + r) * MLP_OUTP + 4 * g);
      f4 acc = mfma16(a, d3h, f4{0.f, 0.f, 0.f, 0.f});
      for (int t = 0; t < 4; ++t) {
        float dd = A.z2[mt][t] > 0.f ? acc[t] : 0.f;
        d2[mt][t] = dd;
        d2h[mt][t] = (_Float16)dd;
      }
    }
    // dH1^T = W2^T . dZ2^T ; dZ1 = dH1 * (Z1 > 0)
    h4 d1h[4];
    f4 d1[4];
    for (int mt = 0; mt < 4; ++mt) {
      f4 acc = {0.f, 0.f, 0.f, 0.f};
      for (int s = 0; s < 4; ++s) {
        h4 a = *(const h4*)(Wh + HOFF_W2T + (16 * mt + r) * MLP_H + 16 * s + 4 * g);
        acc = mfma16(a, d2h[s], acc);
      }
      for (int t = 0; t < 4; ++t) {
        float dd = A.z1[mt][t] > 0.f ? acc[t] : 0.f;
        d1[mt][t] = dd;
        d1h[mt][t] = (_Float16)dd;
      }
    }
    // bias gradients: sum over samples (16 lanes sharing g)
    for (int t = 0; t < 4; ++t) {
      float v3 = d3[t];
      for (int o = 1; o < 16; o <<= 1) v3 += __shfl_xor(v3, o, 64);
      gb3[t] += v3;
      for (int mt = 0; mt < 4; ++mt) {
        float v2 = d2[mt][t], v1 = d1[mt][t];
        for (int o = 1; o < 16; o <<= 1) {
          v2 += __shfl_xor(v2, o, 64);
          v1 += __shfl_xor(v1, o, 64);
        }
        gb2[mt][t] += v2;
        gb1[mt][t] += v1;
      }
    }
    // weight gradients contract over the 16 samples: transpose through LDS
    // dW3[16][64] = dZ3^T . H2 ; A = dZ3^T (units x samples), B = H2 (samples x hidden)
    lds_put(T0, 0, d3h);
    for (int mt = 0; mt < 4; ++mt) lds_put(T1, 16 * mt, A.h2[mt]);
    __builtin_amdgcn_wave_barrier();
    __builtin_amdgcn_fence(__ATOMIC_SEQ_CST, "wavefront");
    {
      h4 a = lds_get_k_samples(T0, 0);
      for (int nt = 0; nt < 4; ++nt) gW3[nt] = mfma16(a, lds_get_k_samples(T1, 16 * nt), gW3[nt]);
    }
    __builtin_amdgcn_wave_barrier();
    __builtin_amdgcn_fence(__ATOMIC_SEQ_CST, "wavefront");
    // dW2[64][64] = dZ2^T . H1
    for (int mt = 0; mt < 4; ++mt) {
      lds_put(T0, 16 * mt, d2h[mt]);
      lds_put(T1, 16 * mt, A.h1[mt]);
    }
    __builtin_amdgcn_wave_barrier();
    __builtin_amdgcn_fence(__ATOMIC_SEQ_CST, "wavefront");
    for (int mt = 0; mt < 4; ++mt) {
      h4 a = lds_get_k_samples(T0, 16 * mt);
      for (int nt = 0; nt < 4; ++nt) gW2[mt][nt] = mfma16(a, lds_get_k_samples(T1, 16 * nt), gW2[mt][nt]);
    }
    __builtin_amdgcn_wave_barrier();
    __builtin_amdgcn_fence(__ATOMIC_SEQ_CST, "wavefront");
    // dW1[64][48] = dZ1^T . X ; X^T staged as [input][sample]
    for (int mt = 0; mt < 4; ++mt) lds_put(T0, 16 * mt, d1h[mt]);
    for (int k = l; k < MLP_INP * 16; k += 64) {
      int sample = k & 15, in = k >> 4;
      size_t rr = row0 + sample;
      T2[in * 16 + sample] = rr < n ? X[rr * MLP_INP + in] : (_Float16)0;
    }
    __builtin_amdgcn_wave_barrier();
    __builtin_amdgcn_fence(__ATOMIC_SEQ_CST, "wavefront");
    for (int mt = 0; mt < 4; ++mt) {
      h4 a = lds_get_k_samples(T0, 16 * mt);
      for (int nt = 0; nt < 3; ++nt) gW1[mt][nt] = mfma16(a, lds_get_k_samples(T2, 16 * nt), gW1[mt][nt]);
    }
    __builtin_amdgcn_wave_barrier();
    __builtin_amdgcn_fence(__ATOMIC_SEQ_CST, "wavefront");
  }
  __syncthreads();
  // reduce the block's waves into LDS (accumulator layout: rows 16*mt + 4g + t, column 16*nt + r)
  for (int mt = 0; mt < 4; ++mt)
    for (int nt = 0; nt < 3; ++nt)
      for (int t = 0; t < 4; ++t) atomicAdd(&red[OFF_W1 + (16 * mt + 4 * g + t) * MLP_INP + 16 * nt + r], gW1[mt][nt][t]);
  for (int mt = 0; mt < 4; ++mt)
    for (int nt = 0; nt < 4; ++nt)
      for (int t = 0; t < 4; ++t) atomicAdd(&red[OFF_W2 + (16 * mt + 4 * g + t) * MLP_H + 16 * nt + r], gW2[mt][nt][t]);
  for (int nt = 0; nt < 4; ++nt)
    for (int t = 0; t < 4; ++t) atomicAdd(&red[OFF_W3 + (4 * g + t) * MLP_H + 16 * nt + r], gW3[nt][t]);
  if (r == 0) {
    for (int mt = 0; mt < 4; ++mt)
      for (int t = 0; t < 4; ++t) {
        atomicAdd(&red[OFF_B1 + 16 * mt + 4 * g + t], gb1[mt][t]);
        atomicAdd(&red[OFF_B2 + 16 * mt + 4 * g + t], gb2[mt][t]);
      }
    for (int t = 0; t < 4; ++t) atomicAdd(&red[OFF_B3 + 4 * g + t], gb3[t]);
  }
  for (int o = 1; o < 64; o <<= 1) lsum += __shfl_xor(lsum, o, 64);
  if (l == 0) lred[w] = lsum;
  __syncthreads();
  for (int i = threadIdx.x; i < MLP_NP; i += BW_BLOCK) partial[(size_t)blockIdx.x * MLP_NP + i] = red[i];
  if (threadIdx.x == 0) {
    float s = 0.f;
    for (int i = 0; i < BW_WAVES; ++i) s += lred[i];
    partial_loss[blockIdx.x] = s;
  }
}

// Sums block partials: grad[i] = sum_b partial[b][i] * scale; loss likewise.
__global__ void k_reduce_partials(const float* __restrict__ partial, const float* __restrict__ partial_loss,
                                  int nblocks, float scale, float* __restrict__ grad, float* __restrict__ loss) {
  int i = blockIdx.x * blockDim.x + threadIdx.x;
  if (i < MLP_NP) {
    float s = 0.f;
    for (int b = 0; b < nblocks; ++b) s += partial[(size_t)b * MLP_NP + i];
    grad[i] = s * scale;
  }
  if (i == 0) {
    float s = 0.f;
    for (int b = 0; b < nblocks; ++b) s += partial_loss[b];
    loss[0] = s * scale;
  }
}

// Adam on the FP32 master weights, then refresh the FP16 copies (and their transposes).
__global__ void k_adam(float* __restrict__ P, float* __restrict__ m, float* __restrict__ v,
                       const float* __restrict__ g, float lr, float b1, float b2, float eps, float c1, float c2) {
  int i = blockIdx.x * blockDim.x + threadIdx.x;
  if (i >= MLP_NP) return;
  float gi = g[i];
  float mi = b1 * m[i] + (1.f - b1) * gi;
  float vi = b2 * v[i] + (1.f - b2) * gi * gi;
  m[i] = mi;
  v[i] = vi;
  float mh = mi / c1, vh = vi / c2;
  P[i] -= lr * mh / (sqrtf(vh) + eps);
}

__global__ void k_to_half(const float* __restrict__ P, _Float16* __restrict__ H) {
  int i = blockIdx.x * blockDim.x + threadIdx.x;
  if (i < MLP_H * MLP_INP) H[HOFF_W1 + i] = (_Float16)P[OFF_W1 + i];
  if (i < MLP_H * MLP_H) {
    H[HOFF_W2 + i] = (_Float16)P[OFF_W2 + i];
    int o = i / MLP_H, in = i % MLP_H;
    H[HOFF_W2T + in * MLP_H + o] = (_Float16)P[OFF_W2 + i];
  }
  if (i < MLP_OUTP * MLP_H) {
    H[HOFF_W3 + i] = (_Float16)P[OFF_W3 + i];
    int o = i / MLP_H, in = i % MLP_H;
    H[HOFF_W3T + in * MLP_OUTP + o] = (_Float16)P[OFF_W3 + i];
  }
}

// Fused inference: features -> MLP -> world direction, one wave per 16 pairs.
// `shapes` (nullable): pairs that are not polytope-polytope (cuboid / convex hull on both sides)
// get an all-zero direction, i.e. the reference's default.  `n_dev` (nullable): device-side count
// capping n.
__global__ __launch_bounds__(256) void k_predict_dirs(const _Float16* __restrict__ Wh, const float* __restrict__ P,
                                                       const ncx_pose* __restrict__ poses,
                                                       const float* __restrict__ desc,
                                                       const uint32_t* __restrict__ ia,
                                                       const uint32_t* __restrict__ ib, size_t ncap,
                                                       float* __restrict__ dirs, const ncx_shape* __restrict__ shapes,
                                                       const uint32_t* __restrict__ n_dev) {
  __shared__ _Float16 xs[4][16 * MLP_INP];
  const int l = __lane_id();
  const int w = threadIdx.x / 64;
  const size_t wave = ((size_t)blockIdx.x * blockDim.x + threadIdx.x) / 64;
  const size_t nw = ((size_t)gridDim.x * blockDim.x) / 64;
  const size_t n = n_dev ? min(ncap, (size_t)*n_dev) : ncap;
  for (size_t row0 = wave * 16; row0 < n; row0 += nw * 16) {
    // 4 lanes per pair build the 48 features (12 each)
    size_t row = row0 + (l >> 2);
    int part = l & 3;
    if (row < n) {
      float f[MLP_INP];
      pair_features(poses, desc, ia[row], ib[row], f);
      for (int c = 0; c < 12; ++c) xs[w][(l >> 2) * MLP_INP + 12 * part + c] = (_Float16)f[12 * part + c];
    } else {
      for (int c = 0; c < 12; ++c) xs[w][(l >> 2) * MLP_INP + 12 * part + c] = (_Float16)0;
    }
    __builtin_amdgcn_wave_barrier();
    __builtin_amdgcn_fence(__ATOMIC_SEQ_CST, "wavefront");
    Acts A;
    f4 z3 = mlp_tile_forward(Wh, P, xs[w], 0, 16, A);
    __builtin_amdgcn_wave_barrier();
    size_t r2 = row0 + (l & 15);
    if ((l >> 4) == 0 && r2 < n) {
      bool poly = true;
      if (shapes) {
        uint32_t t1 = shapes[ia[r2]].type, t2 = shapes[ib[r2]].type;
        poly = (t1 == NCX_CUBOID || t1 == NCX_CONVEX_HULL) && (t2 == NCX_CUBOID || t2 == NCX_CONVEX_HULL);
      }
      iso m1 = load_pose(poses, ia[r2]);
      f3 d = xform_vec(m1, f3{z3[0], z3[1], z3[2]});
      float nl = sqrtf(norm_sq(d));
      if (!(nl > 1e-6f)) d = f3{1.f, 0.f, 0.f}, nl = 1.f;
      if (!poly) d = f3{0.f, 0.f, 0.f}, nl = 1.f;
      dirs[3 * r2] = d.x / nl;
      dirs[3 * r2 + 1] = d.y / nl;
      dirs[3 * r2 + 2] = d.z / nl;
    }
  }
}

// ------------------------------------------------------------------ host side
static int refresh_half(ncx_mlp* m) {
  hipLaunchKernelGGL(k_to_half, dim3(ceil_div(MLP_H * MLP_H, 256)), dim3(256), 0, m->ctx->stream,
                     m->master.as<float>(), m->half.as<_Float16>());
  NCX_HIP_CHECK(hipGetLastError());
  return NCX_OK;
}

extern "C" int ncx_mlp_create(ncx_ctx* ctx, int in, int hidden, int out, uint64_t seed, ncx_mlp** res) {
  NCX_REQUIRE(ctx && res, "ncx_mlp_create: null argument");
  NCX_REQUIRE(in == MLP_IN && hidden == MLP_H && out == MLP_OUT, "ncx_mlp_create: the kernels are built for 44-64-64-3");
  ncx_mlp* m = new ncx_mlp;
  m->ctx = ctx;
  int rc;
  if ((rc = m->master.reserve(MLP_NP * 4)) || (rc = m->m.reserve(MLP_NP * 4)) || (rc = m->v.reserve(MLP_NP * 4)) ||
      (rc = m->half.reserve(HALF_N * 2)) || (rc = m->grad.reserve(MLP_NP * 4)) || (rc = m->loss.reserve(16))) {
    delete m;
    return rc;
  }
  // He-uniform initialisation, padded entries zero.
  std::vector<float> p(MLP_NP, 0.f);
  std::mt19937_64 g(seed);
  auto fill = [&](int off, int rows, int cols, int valid_rows, int valid_cols, int fan_in) {
    float lim = sqrtf(6.f / (float)fan_in);
    std::uniform_real_distribution<float> u(-lim, lim);
    for (int r = 0; r < rows; ++r)
      for (int c = 0; c < cols; ++c) p[off + r * cols + c] = (r < valid_rows && c < valid_cols) ? u(g) : 0.f;
  };
  fill(OFF_W1, MLP_H, MLP_INP, MLP_H, MLP_IN, MLP_IN);
  fill(OFF_W2, MLP_H, MLP_H, MLP_H, MLP_H, MLP_H);
  fill(OFF_W3, MLP_OUTP, MLP_H, MLP_OUT, MLP_H, MLP_H);
  std::vector<float> zero(MLP_NP, 0.f);
  NCX_HIP_CHECK(hipMemcpyAsync(m->master.p, p.data(), MLP_NP * 4, hipMemcpyHostToDevice, ctx->stream));
  NCX_HIP_CHECK(hipMemcpyAsync(m->m.p, zero.data(), MLP_NP * 4, hipMemcpyHostToDevice, ctx->stream));
  NCX_HIP_CHECK(hipMemcpyAsync(m->v.p, zero.data(), MLP_NP * 4, hipMemcpyHostToDevice, ctx->stream));
  if ((rc = refresh_half(m))) return rc;
  NCX_HIP_CHECK(hipStreamSynchronize(ctx->stream));
  *res = m;
  return NCX_OK;
}

extern "C" int ncx_mlp_destroy(ncx_mlp* m) {
  if (!m) return NCX_OK;
  m->master.release(); m->m.release(); m->v.release(); m->half.release(); m->partial.release();
  m->grad.release(); m->loss.release();
  delete m;
  return NCX_OK;
}

extern "C" int ncx_mlp_num_params(ncx_mlp* m, size_t* n) {
  NCX_REQUIRE(m && n, "ncx_mlp_num_params: null argument");
  *n = MLP_NP;
  return NCX_OK;
}

extern "C" int ncx_mlp_get_params(ncx_mlp* m, float* host) {
  NCX_REQUIRE(m && host, "ncx_mlp_get_params: null argument");
  NCX_HIP_CHECK(hipMemcpyAsync(host, m->master.p, MLP_NP * 4, hipMemcpyDeviceToHost, m->ctx->stream));
  NCX_HIP_CHECK(hipStreamSynchronize(m->ctx->stream));
  return NCX_OK;
}

extern "C" int ncx_mlp_set_params(ncx_mlp* m, const float* host) {
  NCX_REQUIRE(m && host, "ncx_mlp_set_params: null argument");
  NCX_HIP_CHECK(hipMemcpyAsync(m->master.p, host, MLP_NP * 4, hipMemcpyHostToDevice, m->ctx->stream));
  int rc = refresh_half(m);
  if (rc) return rc;
  NCX_HIP_CHECK(hipStreamSynchronize(m->ctx->stream));
  return NCX_OK;
}

extern "C" int ncx_mlp_forward(ncx_mlp* m, const void* feat_f16_dev, size_t n, float* out_dev) {
  NCX_REQUIRE(m && feat_f16_dev && out_dev, "ncx_mlp_forward: null argument");
  if (n == 0) return NCX_OK;
  unsigned blocks = ceil_div(ceil_div(n, 16), 4);
  if (blocks > (unsigned)m->ctx->num_cus * 16) blocks = (unsigned)m->ctx->num_cus * 16;
  hipLaunchKernelGGL(k_mlp_forward, dim3(blocks), dim3(256), 0, m->ctx->stream, m->half.as<_Float16>(),
                     m->master.as<float>(), (const _Float16*)feat_f16_dev, n, out_dev);
  NCX_HIP_CHECK(hipGetLastError());
  return NCX_OK;
}

static int mlp_grad_impl(ncx_mlp* m, const void* feat, const float* target, size_t n) {
  NCX_REQUIRE(n > 0, "ncx_mlp: empty batch");
  unsigned blocks = ceil_div(ceil_div(n, 16), BW_WAVES);
  unsigned maxb = (unsigned)m->ctx->num_cus * 2;
  if (blocks > maxb) blocks = maxb;
  int rc;
  if ((rc = m->partial.reserve((size_t)blocks * (MLP_NP + 1) * 4))) return rc;
  float* part = m->partial.as<float>();
  float* ploss = part + (size_t)blocks * MLP_NP;
  hipLaunchKernelGGL(k_mlp_backward, dim3(blocks), dim3(BW_BLOCK), 0, m->ctx->stream, m->half.as<_Float16>(),
                     m->master.as<float>(), (const _Float16*)feat, target, n, part, ploss);
  NCX_HIP_CHECK(hipGetLastError());
  hipLaunchKernelGGL(k_reduce_partials, dim3(ceil_div(MLP_NP, 256)), dim3(256), 0, m->ctx->stream, part, ploss,
                     (int)blocks, 1.f / (float)n, m->grad.as<float>(), m->loss.as<float>());
  NCX_HIP_CHECK(hipGetLastError());
  return NCX_OK;
}

// Gradient of the mean squared error (1/n) sum |f(x) - y|^2 and the loss, for checking.
extern "C" int ncx_mlp_grad(ncx_mlp* m, const void* feat_f16_dev, const float* target_dev, size_t n,
                            float* grad_dev, float* loss_dev) {
  NCX_REQUIRE(m && feat_f16_dev && target_dev && grad_dev && loss_dev, "ncx_mlp_grad: null argument");
  int rc = mlp_grad_impl(m, feat_f16_dev, target_dev, n);
  if (rc) return rc;
  NCX_HIP_CHECK(hipMemcpyAsync(grad_dev, m->grad.p, MLP_NP * 4, hipMemcpyDeviceToDevice, m->ctx->stream));
  NCX_HIP_CHECK(hipMemcpyAsync(loss_dev, m->loss.p, 4, hipMemcpyDeviceToDevice, m->ctx->stream));
  return NCX_OK;
}

// One Adam step on a batch; *loss_host (optional) receives the batch loss (synchronises).
extern "C" int ncx_mlp_train_step(ncx_mlp* m, const void* feat_f16_dev, const float* target_dev, size_t n, float lr,
                                  float* loss_host) {
  NCX_REQUIRE(m && feat_f16_dev && target_dev, "ncx_mlp_train_step: null argument");
  int rc = mlp_grad_impl(m, feat_f16_dev, target_dev, n);
  if (rc) return rc;
  m->step += 1;
  float c1 = 1.f - powf(m->beta1, (float)m->step), c2 = 1.f - powf(m->beta2, (float)m->step);
  hipLaunchKernelGGL(k_adam, dim3(ceil_div(MLP_NP, 256)), dim3(256), 0, m->ctx->stream, m->master.as<float>(),
                     m->m.as<float>(), m->v.as<float>(), m->grad.as<float>(), lr, m->beta1, m->beta2, m->eps, c1, c2);
  NCX_HIP_CHECK(hipGetLastError());
  if ((rc = refresh_half(m))) return rc;
  if (loss_host) {
    NCX_HIP_CHECK(hipMemcpyAsync(loss_host, m->loss.p, 4, hipMemcpyDeviceToHost, m->ctx->stream));
    NCX_HIP_CHECK(hipStreamSynchronize(m->ctx->stream));
  }
  return NCX_OK;
}

extern "C" int ncx_shape_descriptors(ncx_shapes* s, float* desc_dev) {
  NCX_REQUIRE(s && desc_dev, "ncx_shape_descriptors: null argument");
  size_t n = s->n;
  if (n == 0) return NCX_OK;
  hipLaunchKernelGGL(k_descriptors, dim3(ceil_div(n * MLP_DESC, 256)), dim3(256), 0, s->ctx->stream,
                     s->shapes.as<ncx_shape>(), s->hx.as<float>(), s->hy.as<float>(), s->hz.as<float>(), n, desc_dev);
  NCX_HIP_CHECK(hipGetLastError());
  return NCX_OK;
}

extern "C" int ncx_features(ncx_shapes* s, const ncx_pose* poses_dev, const float* desc_dev, const uint32_t* ia_dev,
                            const uint32_t* ib_dev, size_t n, void* feat_f16_dev) {
  NCX_REQUIRE(s && poses_dev && desc_dev && ia_dev && ib_dev && feat_f16_dev, "ncx_features: null argument");
  ncx_ctx* ctx = s->ctx;
  if (n == 0) return NCX_OK;
  hipLaunchKernelGGL(k_features, dim3(ceil_div(n, 256)), dim3(256), 0, ctx->stream, poses_dev, desc_dev, ia_dev, ib_dev,
                     n, (_Float16*)feat_f16_dev);
  NCX_HIP_CHECK(hipGetLastError());
  return NCX_OK;
}

extern "C" int ncx_mlp_predict_dirs(ncx_mlp* m, const ncx_pose* poses_dev, const float* desc_dev,
                                    const uint32_t* ia_dev, const uint32_t* ib_dev, size_t n, float* dirs_dev) {
  NCX_REQUIRE(m && poses_dev && desc_dev && ia_dev && ib_dev && dirs_dev, "ncx_mlp_predict_dirs: null argument");
  if (n == 0) return NCX_OK;
  unsigned blocks = ceil_div(ceil_div(n, 16), 4);
  if (blocks > (unsigned)m->ctx->num_cus * 16) blocks = (unsigned)m->ctx->num_cus * 16;
  hipLaunchKernelGGL(k_predict_dirs, dim3(blocks), dim3(256), 0, m->ctx->stream, m->half.as<_Float16>(),
                     m->master.as<float>(), poses_dev, desc_dev, ia_dev, ib_dev, n, dirs_dev,
                     (const ncx_shape*)nullptr, (const uint32_t*)nullptr);
  NCX_HIP_CHECK(hipGetLastError());
  return NCX_OK;
}

int mlp_predict_device_count(ncx_mlp* m, ncx_ctx* ctx, const ncx_shape* shapes, const ncx_pose* poses,
                             const float* desc, const uint32_t* ia, const uint32_t* ib, const uint32_t* n_dev,
                             size_t cap, float* dirs) {
  unsigned blocks = (unsigned)ctx->num_cus * 8;
  hipLaunchKernelGGL(k_predict_dirs, dim3(blocks), dim3(256), 0, ctx->stream, m->half.as<_Float16>(),
                     m->master.as<float>(), poses, desc, ia, ib, cap, dirs, shapes, n_dev);
  NCX_HIP_CHECK(hipGetLastError());
  return NCX_OK;
}

int shape_descriptors_on(ncx_shapes* s, ncx_ctx* ctx, float* desc) {
  if (s->n == 0) return NCX_OK;
  hipLaunchKernelGGL(k_descriptors, dim3(ceil_div(s->n * MLP_DESC, 256)), dim3(256), 0, ctx->stream,
                     s->shapes.as<ncx_shape>(), s->hx.as<float>(), s->hy.as<float>(), s->hz.as<float>(), s->n, desc);
  NCX_HIP_CHECK(hipGetLastError());
  return NCX_OK;
}
