// Learned initial-pose proposal: a multilayer perceptron trained and run
// with hand-written bf16 MFMA kernels (gfx950), and the bird's-eye-view
// descriptor of a scan that feeds it.
//
// Network: sizes s0 -> s1 -> ... -> sL; hidden layers use GELU (exact erf
// form), the last layer is linear.  Every layer's weights are kept as FP32
// masters (Adam state alongside) and as two bf16 copies: W^T (out x in, for
// the forward product) and W (in x out, for the backward product with
// respect to the layer input).  Products run on MFMA 32x32x16 bf16 -> fp32.
//
// All products are of the form C = A B^T with A (M x K) and B (N x K) both
// row-major, so every MFMA operand fragment is one 16-byte load along K;
// the backward products use transposed copies of activations and gradients
// made by a small transpose kernel.
#include <hip/hip_runtime.h>

#include <cmath>
#include <cstdint>
#include <cstring>
#include <string>
#include <vector>

#include "../../include/ndt_hip.h"

namespace {

typedef __attribute__((__vector_size__(8 * sizeof(__bf16)))) __bf16 bf16x8;
typedef __attribute__((__vector_size__(16 * sizeof(float)))) float floatx16;

__host__ __device__ inline uint16_t f2bf(float f) {
  // round to nearest even (inputs here are finite)
  uint32_t u;
  std::memcpy(&u, &f, 4);
  u += 0x7FFFu + ((u >> 16) & 1u);
  return (uint16_t)(u >> 16);
}
__host__ __device__ inline float bf2f(uint16_t b) {
  const uint32_t u = (uint32_t)b << 16;
  float f;
  std::memcpy(&f, &u, 4);
  return f;
}

constexpr int kTile = 32;      // MFMA output tile
constexpr int kKStep = 16;     // K per MFMA
constexpr int kWavesPerWg = 4; // four 32x32 tiles along N per workgroup

enum Epilogue { kStore = 0, kBias = 1, kPartial = 2 };

// C[M x N] (f32) = A[M x K] B[N x K]^T (+ bias[N]).  M, N multiples of 32,
// K a multiple of 16.  One wave per 32x32 tile; grid (N / 128, M / 32).
//
// MFMA 32x32x16 bf16 operand layout (lane l, h = l / 32, r = l % 32): the A
// fragment is A[r][8h .. 8h+7] of the 32x16 slice, the B fragment is
// B[8h .. 8h+7][r] of the 16x32 slice, i.e. row r of the N x K operand; the
// 16 results of lane l are D[8i + 4h + j][r] for i, j in 0..3.
__global__ __launch_bounds__(64 * kWavesPerWg) void gemm_nt(const uint16_t* __restrict__ A,
                                                            const uint16_t* __restrict__ B, float* __restrict__ C,
                                                            const float* __restrict__ bias, int M, int N, int K,
                                                            int epi) {
  const int lane = threadIdx.x & 63, wave = threadIdx.x >> 6;
  const int tn = (blockIdx.x * kWavesPerWg + wave) * kTile, tm = blockIdx.y * kTile;
  if (tn >= N || tm >= M) return;
  const int r = lane & 31, h = lane >> 5;
  const uint16_t* a = A + (size_t)(tm + r) * K + 8 * h;
  const uint16_t* b = B + (size_t)(tn + r) * K + 8 * h;
  floatx16 acc = {};
  for (int k = 0; k < K; k += kKStep) {
    const bf16x8 av = *reinterpret_cast<const bf16x8*>(a + k);
    const bf16x8 bv = *reinterpret_cast<const bf16x8*>(b + k);
    acc = __builtin_amdgcn_mfma_f32_32x32x16_bf16(av, bv, acc, 0, 0, 0);
  }
  const float bb = (epi == kBias) ? bias[tn + r] : 0.0f;
#pragma unroll
  for (int i = 0; i < 4; ++i)
#pragma unroll
    for (int j = 0; j < 4; ++j) C[(size_t)(tm + 8 * i + 4 * h + j) * N + tn + r] = acc[4 * i + j] + bb;
}

// Split-K form of gemm_nt for long K: grid (N / 128, M / 32, S); slice s
// covers K range [s * kchunk, min(K, (s + 1) * kchunk)) and writes its
// partial tile to P + s * M * N.  Two MFMA chains per wave (even and odd
// 16-wide K steps) keep two products in flight.  kchunk and K are multiples
// of 32.
__global__ __launch_bounds__(64 * kWavesPerWg) void gemm_nt_splitk(const uint16_t* __restrict__ A,
                                                                   const uint16_t* __restrict__ B,
                                                                   float* __restrict__ P, int M, int N, int K,
                                                                   int kchunk) {
  const int lane = threadIdx.x & 63, wave = threadIdx.x >> 6;
  const int tn = (blockIdx.x * kWavesPerWg + wave) * kTile, tm = blockIdx.y * kTile;
  if (tn >= N || tm >= M) return;
  const int k0 = blockIdx.z * kchunk, k1 = min(K, k0 + kchunk);
  const int r = lane & 31, h = lane >> 5;
  const uint16_t* a = A + (size_t)(tm + r) * K + 8 * h;
  const uint16_t* b = B + (size_t)(tn + r) * K + 8 * h;
  floatx16 acc0 = {}, acc1 = {};
  for (int k = k0; k < k1; k += 2 * kKStep) {
    const bf16x8 a0 = *reinterpret_cast<const bf16x8*>(a + k);
    const bf16x8 b0 = *reinterpret_cast<const bf16x8*>(b + k);
    const bf16x8 a1 = *reinterpret_cast<const bf16x8*>(a + k + kKStep);
    const bf16x8 b1 = *reinterpret_cast<const bf16x8*>(b + k + kKStep);
    acc0 = __builtin_amdgcn_mfma_f32_32x32x16_bf16(a0, b0, acc0, 0, 0, 0);
    acc1 = __builtin_amdgcn_mfma_f32_32x32x16_bf16(a1, b1, acc1, 0, 0, 0);
  }
  float* out = P + (size_t)blockIdx.z * M * N;
#pragma unroll
  for (int i = 0; i < 4; ++i)
#pragma unroll
    for (int j = 0; j < 4; ++j)
      out[(size_t)(tm + 8 * i + 4 * h + j) * N + tn + r] = acc0[4 * i + j] + acc1[4 * i + j];
}

// Macro-tiled form for the large products: a workgroup computes a 128x128
// tile of C (each wave a 64x64 quadrant: 2x2 MFMA tiles), K in steps of 32
// staged through double-buffered LDS with the next step's global loads in
// flight during the current step's products.  Operand traffic per product is
// a quarter of one 32x32 tile per wave.  grid (S, tiles): blockIdx.x is the K
// slice [x * kchunk, min(K, (x + 1) * kchunk)), blockIdx.y the tile (row-major
// over N / 128 columns); with S a multiple of 8 the slice index picks the XCD,
// so workgroups that read the same K range share one L2.  M, N multiples of
// 128; K and kchunk multiples of 32.  epi as for gemm_nt (kPartial: tile to
// C + x * M * N).
constexpr int kMacro = 128;
constexpr int kKStage = 32;
constexpr int kLdsRow = kKStage + 8;  // bf16 per LDS row (80 bytes: staggers the banks)
__global__ __launch_bounds__(256) void gemm_nt_tiled(const uint16_t* __restrict__ A, const uint16_t* __restrict__ B,
                                                     float* __restrict__ C, const float* __restrict__ bias, int M,
                                                     int N, int K, int kchunk, int epi) {
  __shared__ __attribute__((aligned(16))) uint16_t sA[2][kMacro * kLdsRow];
  __shared__ __attribute__((aligned(16))) uint16_t sB[2][kMacro * kLdsRow];
  const int tiles_n = N / kMacro;
  const int tm = (int)(blockIdx.y / tiles_n) * kMacro, tn = (int)(blockIdx.y % tiles_n) * kMacro;
  const int k0 = blockIdx.x * kchunk, k1 = min(K, k0 + kchunk);
  const int steps = (k1 - k0) / kKStage;
  const int t = threadIdx.x, lane = t & 63, wave = t >> 6;
  const int wm = wave >> 1, wn = wave & 1;
  const int r = lane & 31, h = lane >> 5;
  // copy assignment: 512 chunks of 8 bf16 per operand and stage; thread t
  // takes chunks t and t + 256: row c / 4, k offset (c % 4) * 8
  const int row0 = t >> 2, kseg = (t & 3) * 8;
  const uint16_t* ga0 = A + (size_t)(tm + row0) * K + k0 + kseg;
  const uint16_t* ga1 = A + (size_t)(tm + row0 + 64) * K + k0 + kseg;
  const uint16_t* gb0 = B + (size_t)(tn + row0) * K + k0 + kseg;
  const uint16_t* gb1 = B + (size_t)(tn + row0 + 64) * K + k0 + kseg;
  const int l0 = row0 * kLdsRow + kseg, l1 = (row0 + 64) * kLdsRow + kseg;
  bf16x8 pa0, pa1, pb0, pb1;
  auto fetch = [&](int s) {
    const int o = s * kKStage;
    pa0 = *reinterpret_cast<const bf16x8*>(ga0 + o);
    pa1 = *reinterpret_cast<const bf16x8*>(ga1 + o);
    pb0 = *reinterpret_cast<const bf16x8*>(gb0 + o);
    pb1 = *reinterpret_cast<const bf16x8*>(gb1 + o);
  };
  auto stash = [&](int buf) {
    *reinterpret_cast<bf16x8*>(&sA[buf][l0]) = pa0;
    *reinterpret_cast<bf16x8*>(&sA[buf][l1]) = pa1;
    *reinterpret_cast<bf16x8*>(&sB[buf][l0]) = pb0;
    *reinterpret_cast<bf16x8*>(&sB[buf][l1]) = pb1;
  };
  floatx16 acc[2][2] = {};
  if (steps > 0) {
    fetch(0);
    stash(0);
  }
  __syncthreads();
  for (int s = 0; s < steps; ++s) {
    const int buf = s & 1;
    if (s + 1 < steps) fetch(s + 1);
#pragma unroll
    for (int kh = 0; kh < 2; ++kh) {
      bf16x8 fa[2], fb[2];
#pragma unroll
      for (int i = 0; i < 2; ++i) {
        fa[i] = *reinterpret_cast<const bf16x8*>(&sA[buf][(64 * wm + 32 * i + r) * kLdsRow + 16 * kh + 8 * h]);
        fb[i] = *reinterpret_cast<const bf16x8*>(&sB[buf][(64 * wn + 32 * i + r) * kLdsRow + 16 * kh + 8 * h]);
      }
#pragma unroll
      for (int i = 0; i < 2; ++i)
#pragma unroll
        for (int j = 0; j < 2; ++j) acc[i][j] = __builtin_amdgcn_mfma_f32_32x32x16_bf16(fa[i], fb[j], acc[i][j], 0, 0, 0);
    }
    if (s + 1 < steps) stash(buf ^ 1);
    __syncthreads();
  }
  float* out = C + (epi == kPartial ? (size_t)blockIdx.x * M * N : 0);
#pragma unroll
  for (int i = 0; i < 2; ++i)
#pragma unroll
    for (int j = 0; j < 2; ++j) {
      const int col = tn + 64 * wn + 32 * j + r;
      const float bb = (epi == kBias) ? bias[col] : 0.0f;
#pragma unroll
      for (int a = 0; a < 4; ++a)
#pragma unroll
        for (int b = 0; b < 4; ++b)
          out[(size_t)(tm + 64 * wm + 32 * i + 8 * a + 4 * h + b) * N + col] = acc[i][j][4 * a + b] + bb;
    }
}

// C[i] = sum over s = 0 .. S-1 of P[s * MN + i] (in that order) + bias[i % N].
__global__ void splitk_reduce(const float* __restrict__ P, float* __restrict__ C, const float* __restrict__ bias,
                              size_t MN, int N, int S) {
  for (size_t i = blockIdx.x * (size_t)blockDim.x + threadIdx.x; i < MN; i += (size_t)gridDim.x * blockDim.x) {
    float v = 0.0f;
    for (int s = 0; s < S; ++s) v += P[(size_t)s * MN + i];
    C[i] = bias ? v + bias[i % N] : v;
  }
}

__device__ inline float gelu(float z) { return 0.5f * z * (1.0f + erff(z * 0.70710678118654752f)); }
__device__ inline float gelu_grad(float z) {
  const float cdf = 0.5f * (1.0f + erff(z * 0.70710678118654752f));
  const float pdf = 0.39894228040143268f * expf(-0.5f * z * z);
  return cdf + z * pdf;
}

// act = bf16(gelu(z)) (hidden) or bf16(z) (pass-through); n elements.
__global__ void activate(const float* __restrict__ z, uint16_t* __restrict__ act, size_t n, int apply_gelu) {
  for (size_t i = blockIdx.x * (size_t)blockDim.x + threadIdx.x; i < n; i += (size_t)gridDim.x * blockDim.x)
    act[i] = f2bf(apply_gelu ? gelu(z[i]) : z[i]);
}

// dz = dy * gelu'(z) (f32), n elements.
__global__ void gelu_backward(const float* __restrict__ dy, const float* __restrict__ z, float* __restrict__ dz,
                              size_t n) {
  for (size_t i = blockIdx.x * (size_t)blockDim.x + threadIdx.x; i < n; i += (size_t)gridDim.x * blockDim.x)
    dz[i] = dy[i] * gelu_grad(z[i]);
}

// out[c][r] = bf16(in[r][c]) for an R x C f32 or bf16 matrix (tiled through LDS).
template <class T>
__global__ void transpose_to_bf16(const T* __restrict__ in, uint16_t* __restrict__ out, int R, int C) {
  __shared__ float tile[32][33];
  const int c0 = blockIdx.x * 32, r0 = blockIdx.y * 32;
  const int tx = threadIdx.x & 31, ty = threadIdx.x >> 5;  // 256 threads: 32 x 8
  for (int y = ty; y < 32; y += 8) {
    const int r = r0 + y, c = c0 + tx;
    float v = 0.0f;
    if (r < R && c < C) {
      if constexpr (sizeof(T) == 4)
        v = in[(size_t)r * C + c];
      else
        v = bf2f(in[(size_t)r * C + c]);
    }
    tile[y][tx] = v;
  }
  __syncthreads();
  for (int y = ty; y < 32; y += 8) {
    const int c = c0 + y, r = r0 + tx;
    if (r < R && c < C) out[(size_t)c * R + r] = f2bf(tile[tx][y]);
  }
}

__global__ void to_bf16(const float* __restrict__ in, uint16_t* __restrict__ out, size_t n) {
  for (size_t i = blockIdx.x * (size_t)blockDim.x + threadIdx.x; i < n; i += (size_t)gridDim.x * blockDim.x)
    out[i] = f2bf(in[i]);
}

// db[n] = sum over rows of dz (R x N), fixed order.
// db[c] = sum over rows of dz[r][c] (R x N row-major).  Workgroup of 1024
// threads per 64 columns: thread (c, g) sums rows g, g + 16, ... of column c,
// then the 16 partials of a column are added in order g = 0 .. 15.
__global__ __launch_bounds__(1024) void column_sums(const float* __restrict__ dz, float* __restrict__ db, int R,
                                                    int N) {
  __shared__ float part[16][64];
  const int cl = threadIdx.x & 63, g = threadIdx.x >> 6;
  const int c = blockIdx.x * 64 + cl;
  float v = 0.0f;
  if (c < N)
    for (int r = g; r < R; r += 16) v += dz[(size_t)r * N + c];
  part[g][cl] = v;
  __syncthreads();
  if (g == 0 && c < N) {
    float t = 0.0f;
    for (int k = 0; k < 16; ++k) t += part[k][cl];
    db[c] = t;
  }
}

__global__ void adam_step(float* __restrict__ w, const float* __restrict__ g, float* __restrict__ m,
                          float* __restrict__ v, uint16_t* __restrict__ wt_bf, size_t n, float lr, float b1, float b2,
                          float c1, float c2, float eps) {
  for (size_t i = blockIdx.x * (size_t)blockDim.x + threadIdx.x; i < n; i += (size_t)gridDim.x * blockDim.x) {
    const float gi = g[i];
    const float mi = b1 * m[i] + (1.0f - b1) * gi;
    const float vi = b2 * v[i] + (1.0f - b2) * gi * gi;
    m[i] = mi;
    v[i] = vi;
    const float wi = w[i] - lr * (mi / c1) / (sqrtf(vi / c2) + eps);
    w[i] = wi;
    if (wt_bf) wt_bf[i] = f2bf(wi);
  }
}

// Winner-takes-all proposal loss, one thread per sample.  out: B x (4H)
// (x, y, yaw, logit per hypothesis), label: B x 3 (x, y, yaw).  The closest
// hypothesis (squared x/y error + 2 (1 - cos dyaw)) gets the regression
// loss, every hypothesis a small share eps of it, and the logits a
// cross-entropy toward the winner.  Writes dout and per-sample loss.
__global__ void wta_loss(const float* __restrict__ out, const float* __restrict__ label, float* __restrict__ dout,
                         float* __restrict__ loss, int B, int H, float eps) {
  const int s = blockIdx.x * blockDim.x + threadIdx.x;
  if (s >= B) return;
  const float* o = out + (size_t)s * 4 * H;
  float* d = dout + (size_t)s * 4 * H;
  const float lx = label[3 * s], ly = label[3 * s + 1], lt = label[3 * s + 2];
  int best = 0;
  float best_e = 3.0e38f, lmax = -3.0e38f;
  for (int k = 0; k < H; ++k) {
    const float dx = o[4 * k] - lx, dy = o[4 * k + 1] - ly;
    const float e = dx * dx + dy * dy + 2.0f * (1.0f - cosf(o[4 * k + 2] - lt));
    if (e < best_e) {
      best_e = e;
      best = k;
    }
    lmax = fmaxf(lmax, o[4 * k + 3]);
  }
  float z = 0.0f;
  for (int k = 0; k < H; ++k) z += expf(o[4 * k + 3] - lmax);
  float total = 0.0f;
  const float inv_b = 1.0f / (float)B;
  for (int k = 0; k < H; ++k) {
    const float wk = (k == best ? 1.0f : 0.0f) + eps / (float)H;
    const float dx = o[4 * k] - lx, dy = o[4 * k + 1] - ly, dt = o[4 * k + 2] - lt;
    total += wk * (dx * dx + dy * dy + 2.0f * (1.0f - cosf(dt)));
    d[4 * k] = inv_b * wk * 2.0f * dx;
    d[4 * k + 1] = inv_b * wk * 2.0f * dy;
    d[4 * k + 2] = inv_b * wk * 2.0f * sinf(dt);
    const float p = expf(o[4 * k + 3] - lmax) / z;
    d[4 * k + 3] = inv_b * (p - (k == best ? 1.0f : 0.0f));
  }
  loss[s] = total - (o[4 * best + 3] - lmax - logf(z));
}

// Bird's-eye-view descriptor: row j describes scan q = index ? index[j] : j
// (points [offsets[q], offsets[q+1]) of xyz, sensor frame), optionally
// rotated about z by (c, s) = rot[j] (x' = c x - s y, y' = s x + c y, each
// product and difference rounded on its own), as log(1 + count) / 4 per cell
// of an nx x ny x nz grid of `cell` metres centred on the sensor in x/y,
// heights from zmin in steps of zcell; one workgroup per row, counts in LDS
// (nx*ny*nz <= 32768).
__global__ void bev_kernel(const float* __restrict__ xyz, const uint32_t* __restrict__ offsets,
                           const uint32_t* __restrict__ index, const float2* __restrict__ rot,
                           uint16_t* __restrict__ bev, int nx, int ny, int nz, float cell, float zmin, float zcell) {
  extern __shared__ uint32_t counts[];
  const int cells = nx * ny * nz;
  for (int c = threadIdx.x; c < cells; c += blockDim.x) counts[c] = 0;
  __syncthreads();
  const uint32_t q = index ? index[blockIdx.x] : blockIdx.x;
  const uint32_t b = offsets[q], e = offsets[q + 1];
  const float2 cs = rot ? rot[blockIdx.x] : make_float2(1.0f, 0.0f);
  for (uint32_t i = b + threadIdx.x; i < e; i += blockDim.x) {
    float x = xyz[3 * i], y = xyz[3 * i + 1];
    if (rot) {
      const float xr = __fsub_rn(__fmul_rn(cs.x, x), __fmul_rn(cs.y, y));
      y = __fadd_rn(__fmul_rn(cs.y, x), __fmul_rn(cs.x, y));
      x = xr;
    }
    const int ix = (int)floorf(x / cell) + nx / 2;
    const int iy = (int)floorf(y / cell) + ny / 2;
    const int iz = (int)floorf((xyz[3 * i + 2] - zmin) / zcell);
    if (ix >= 0 && ix < nx && iy >= 0 && iy < ny && iz >= 0 && iz < nz) atomicAdd(&counts[(iy * nx + ix) * nz + iz], 1u);
  }
  __syncthreads();
  uint16_t* out = bev + (size_t)blockIdx.x * cells;
  for (int c = threadIdx.x; c < cells; c += blockDim.x) out[c] = f2bf(log1pf((float)counts[c]) * 0.25f);
}

int grid1(size_t n) { return (int)std::min<size_t>(4096, (n + 255) / 256); }

}  // namespace

struct ndt_mlp {
  std::vector<int> sizes;
  int layers = 0;
  hipStream_t st = nullptr;
  std::string err;
  // per layer
  std::vector<float*> w, b, gw, gb, mw, vw, mb, vb;  // f32 (w is N x K, i.e. W^T)
  std::vector<uint16_t*> wt_bf, w_bf;                // bf16 W^T (N x K) and W (K x N)
  // activations for the current batch
  int cap = 0;
  std::vector<float*> z;        // pre-activations per layer (B x N)
  std::vector<uint16_t*> act;   // bf16 inputs per layer (B x K): act[0] = x, act[l] = gelu(z[l-1])
  float *dz = nullptr, *dy = nullptr, *dout = nullptr, *loss = nullptr, *label = nullptr;
  uint16_t *dz_bf = nullptr, *dzt_bf = nullptr, *actt_bf = nullptr;
  float* out_host = nullptr;
  float* splitk = nullptr;  // split-K partial tiles
  size_t splitk_cap = 0;    // floats
  long step = 0;
  std::vector<void*> allocs;
};

namespace {

int mlp_err(ndt_mlp* m, hipError_t e, const char* what) {
  if (m) m->err = std::string(what) + ": " + hipGetErrorString(e);
  return NDT_ERR_HIP;
}
#define MLP_HIP(call, what)                       \
  do {                                            \
    hipError_t e_ = (call);                       \
    if (e_ != hipSuccess) return mlp_err(m, e_, what); \
  } while (0)

template <class T>
hipError_t mlp_alloc(ndt_mlp* m, T** p, size_t n) {
  hipError_t e = hipMalloc(reinterpret_cast<void**>(p), std::max<size_t>(n, 1) * sizeof(T));
  if (e == hipSuccess) m->allocs.push_back(*p);
  return e;
}

// Split-K plan for an M x N x K product: slices of kchunk (a multiple of 32)
// so that the tiles times the slices reach about 2048 waves; one slice when
// K is short or the tiles already fill the GPU.
constexpr int kSplitMinK = 4096;
void splitk_plan(int M, int N, int K, int* S, int* kchunk) {
  const int tiles = (M / kTile) * (N / kTile);
  int s = 1;
  if (K >= kSplitMinK && tiles < 2048) s = std::min((2048 + tiles - 1) / tiles, K / 512);
  int kc = (K + s - 1) / s;
  kc = (kc + 2 * kKStep - 1) / (2 * kKStep) * (2 * kKStep);
  *kchunk = kc;
  *S = (K + kc - 1) / kc;
}

// Macro-tile plan: shapes that fit 128x128 tiles, with K split into slices
// (a multiple of 8 of them when there are at least 8) until about 256
// workgroups run; false when even then fewer than 128 would (the small
// products stay on gemm_nt).
bool tiled_plan(int M, int N, int K, int* S, int* kchunk) {
  if (M % kMacro || N % kMacro || K % kKStage) return false;
  const int tiles = (M / kMacro) * (N / kMacro);
  int s = 1;
  if (tiles < 256) s = std::min((256 + tiles - 1) / tiles, std::max(1, K / 512));
  if (s >= 8) s = s / 8 * 8;
  int kc = (K + s - 1) / s;
  kc = (kc + kKStage - 1) / kKStage * kKStage;
  s = (K + kc - 1) / kc;
  if (tiles * s < 128) return false;
  *S = s;
  *kchunk = kc;
  return true;
}

size_t splitk_floats(int M, int N, int K) {
  int S, kc;
  if (tiled_plan(M, N, K, &S, &kc)) return S > 1 ? (size_t)S * M * N : 0;
  splitk_plan(M, N, K, &S, &kc);
  return S > 1 ? (size_t)S * M * N : 0;
}

int ensure_batch(ndt_mlp* m, int B) {
  if (B <= m->cap) return NDT_OK;
  for (int l = 0; l < m->layers; ++l) {
    if (m->z[l]) (void)hipFree(m->z[l]);
    if (m->act[l]) (void)hipFree(m->act[l]);
  }
  for (void* p : {(void*)m->dz, (void*)m->dy, (void*)m->dout, (void*)m->loss, (void*)m->label, (void*)m->dz_bf,
                  (void*)m->dzt_bf, (void*)m->actt_bf})
    if (p) (void)hipFree(p);
  if (m->out_host) (void)hipHostFree(m->out_host);
  size_t widest = 0;
  for (int s : m->sizes) widest = std::max<size_t>(widest, s);
  for (int l = 0; l < m->layers; ++l) {
    MLP_HIP(hipMalloc(&m->z[l], (size_t)B * m->sizes[l + 1] * 4), "alloc z");
    MLP_HIP(hipMalloc(&m->act[l], (size_t)B * m->sizes[l] * 2), "alloc act");
  }
  MLP_HIP(hipMalloc(&m->dz, (size_t)B * widest * 4), "alloc dz");
  MLP_HIP(hipMalloc(&m->dy, (size_t)B * widest * 4), "alloc dy");
  MLP_HIP(hipMalloc(&m->dout, (size_t)B * m->sizes.back() * 4), "alloc dout");
  MLP_HIP(hipMalloc(&m->loss, (size_t)B * 4), "alloc loss");
  MLP_HIP(hipMalloc(&m->label, (size_t)B * 3 * 4), "alloc label");
  MLP_HIP(hipMalloc(&m->dz_bf, (size_t)B * widest * 2), "alloc dz");
  MLP_HIP(hipMalloc(&m->dzt_bf, (size_t)B * widest * 2), "alloc dz");
  MLP_HIP(hipMalloc(&m->actt_bf, (size_t)B * widest * 2), "alloc act");
  MLP_HIP(hipHostMalloc(&m->out_host, (size_t)B * std::max<size_t>(m->sizes.back(), 1) * 4), "alloc out");
  // split-K partials: the forward products (B x N x K) are the long-K ones
  size_t sk = 0;
  for (int l = 0; l < m->layers; ++l) sk = std::max(sk, splitk_floats(B, m->sizes[l + 1], m->sizes[l]));
  if (sk > m->splitk_cap) {
    if (m->splitk) (void)hipFree(m->splitk);
    m->splitk = nullptr;
    m->splitk_cap = 0;
    MLP_HIP(hipMalloc(&m->splitk, sk * 4), "alloc split-K");
    m->splitk_cap = sk;
  }
  m->cap = B;
  return NDT_OK;
}

void launch_gemm(ndt_mlp* m, const uint16_t* A, const uint16_t* B, float* C, const float* bias, int M, int N, int K) {
  const dim3 grid((unsigned)((N + kTile * kWavesPerWg - 1) / (kTile * kWavesPerWg)), (unsigned)(M / kTile));
  int S, kc;
  if (tiled_plan(M, N, K, &S, &kc) && (S == 1 || (size_t)S * M * N <= m->splitk_cap)) {
    const dim3 tg((unsigned)S, (unsigned)((M / kMacro) * (N / kMacro)));
    if (S == 1) {
      hipLaunchKernelGGL(gemm_nt_tiled, tg, dim3(256), 0, m->st, A, B, C, bias, M, N, K, kc,
                         bias ? (int)kBias : (int)kStore);
    } else {
      hipLaunchKernelGGL(gemm_nt_tiled, tg, dim3(256), 0, m->st, A, B, m->splitk, nullptr, M, N, K, kc,
                         (int)kPartial);
      const size_t mn = (size_t)M * N;
      hipLaunchKernelGGL(splitk_reduce, dim3(grid1(mn)), dim3(256), 0, m->st, m->splitk, C, bias, mn, N, S);
    }
    return;
  }
  splitk_plan(M, N, K, &S, &kc);
  if (S > 1 && (size_t)S * M * N <= m->splitk_cap && K % (2 * kKStep) == 0) {
    hipLaunchKernelGGL(gemm_nt_splitk, dim3(grid.x, grid.y, (unsigned)S), dim3(64 * kWavesPerWg), 0, m->st, A, B,
                       m->splitk, M, N, K, kc);
    const size_t mn = (size_t)M * N;
    hipLaunchKernelGGL(splitk_reduce, dim3(grid1(mn)), dim3(256), 0, m->st, m->splitk, C, bias, mn, N, S);
    return;
  }
  hipLaunchKernelGGL(gemm_nt, grid, dim3(64 * kWavesPerWg), 0, m->st, A, B, C, bias, M, N, K, bias ? kBias : kStore);
}

// Forward over act[0] (already holding the bf16 input, B rows); z[L-1] = output.
void forward(ndt_mlp* m, int B) {
  for (int l = 0; l < m->layers; ++l) {
    const int K = m->sizes[l], N = m->sizes[l + 1];
    launch_gemm(m, m->act[l], m->wt_bf[l], m->z[l], m->b[l], B, N, K);
    if (l + 1 < m->layers) {
      const size_t n = (size_t)B * N;
      hipLaunchKernelGGL(activate, dim3(grid1(n)), dim3(256), 0, m->st, m->z[l], m->act[l + 1], n, 1);
    }
  }
}

// Backward from m->dout (B x out, f32): gradients into gw / gb.
void backward(ndt_mlp* m, int B) {
  const float* dy = m->dout;
  for (int l = m->layers - 1; l >= 0; --l) {
    const int K = m->sizes[l], N = m->sizes[l + 1];
    const size_t n = (size_t)B * N;
    const float* dz = dy;
    if (l + 1 < m->layers) {
      hipLaunchKernelGGL(gelu_backward, dim3(grid1(n)), dim3(256), 0, m->st, dy, m->z[l], m->dz, n);
      dz = m->dz;
    }
    // gb = column sums of dz; gw (N x K) = dz^T act[l]
    hipLaunchKernelGGL(column_sums, dim3((N + 63) / 64), dim3(1024), 0, m->st, dz, m->gb[l], B, N);
    hipLaunchKernelGGL(transpose_to_bf16<float>, dim3((N + 31) / 32, (B + 31) / 32), dim3(256), 0, m->st, dz,
                       m->dzt_bf, B, N);
    hipLaunchKernelGGL(transpose_to_bf16<uint16_t>, dim3((K + 31) / 32, (B + 31) / 32), dim3(256), 0, m->st,
                       m->act[l], m->actt_bf, B, K);
    launch_gemm(m, m->dzt_bf, m->actt_bf, m->gw[l], nullptr, N, K, B);
    if (l > 0) {
      // d act[l] = dz W^T... as C = A B^T with A = dz (B x N), B = W (K x N)
      hipLaunchKernelGGL(to_bf16, dim3(grid1(n)), dim3(256), 0, m->st, dz, m->dz_bf, n);
      launch_gemm(m, m->dz_bf, m->w_bf[l], m->dy, nullptr, B, K, N);
      dy = m->dy;
    }
  }
}

void refresh_w_copy(ndt_mlp* m, int l) {
  const int K = m->sizes[l], N = m->sizes[l + 1];
  // w is N x K (W^T); w_bf is K x N
  hipLaunchKernelGGL(transpose_to_bf16<float>, dim3((K + 31) / 32, (N + 31) / 32), dim3(256), 0, m->st, m->w[l],
                     m->w_bf[l], N, K);
  hipLaunchKernelGGL(to_bf16, dim3(grid1((size_t)N * K)), dim3(256), 0, m->st, m->w[l], m->wt_bf[l], (size_t)N * K);
}

bool shape_ok(const ndt_mlp* m, size_t batch) {
  if (batch == 0 || batch > (1u << 24) || batch % kTile) return false;
  for (int l = 0; l < m->layers; ++l)
    if (m->sizes[l] % kTile || m->sizes[l + 1] % kTile || m->sizes[l] < kKStep) return false;
  return true;
}

}  // namespace

extern "C" {

int ndt_hip_mlp_create(const int32_t* sizes, int32_t n_sizes, uint64_t seed, ndt_mlp** out) {
  if (!sizes || !out || n_sizes < 2) return NDT_ERR_INVALID_ARGUMENT;
  for (int i = 0; i < n_sizes; ++i)
    if (sizes[i] <= 0 || sizes[i] % kTile) return NDT_ERR_INVALID_ARGUMENT;
  auto* m = new ndt_mlp();
  m->sizes.assign(sizes, sizes + n_sizes);
  m->layers = n_sizes - 1;
  const int L = m->layers;
  for (auto* v : {&m->w, &m->b, &m->gw, &m->gb, &m->mw, &m->vw, &m->mb, &m->vb}) v->assign(L, nullptr);
  m->wt_bf.assign(L, nullptr);
  m->w_bf.assign(L, nullptr);
  m->z.assign(L, nullptr);
  m->act.assign(L, nullptr);
  hipError_t e = hipStreamCreateWithFlags(&m->st, hipStreamNonBlocking);
  // He-style initialisation from a SplitMix64 stream (uniform, variance 2/fan_in
  // for hidden layers, 1/fan_in for the output layer).
  uint64_t x = seed;
  auto next = [&]() {
    x += 0x9E3779B97F4A7C15ull;
    uint64_t z = x;
    z = (z ^ (z >> 30)) * 0xBF58476D1CE4E5B9ull;
    z = (z ^ (z >> 27)) * 0x94D049BB133111EBull;
    z ^= z >> 31;
    return (double)(z >> 11) * (1.0 / 9007199254740992.0);
  };
  for (int l = 0; l < L && e == hipSuccess; ++l) {
    const size_t K = sizes[l], N = sizes[l + 1];
    std::vector<float> w(N * K), zero(std::max(N * K, N), 0.0f);
    const double a = std::sqrt(3.0 * (l + 1 < L ? 2.0 : 1.0) / (double)K);
    for (auto& v : w) v = (float)((2.0 * next() - 1.0) * a);
    for (auto* p : {&m->w[l], &m->gw[l], &m->mw[l], &m->vw[l]})
      if (e == hipSuccess) e = mlp_alloc(m, p, N * K);
    for (auto* p : {&m->b[l], &m->gb[l], &m->mb[l], &m->vb[l]})
      if (e == hipSuccess) e = mlp_alloc(m, p, N);
    if (e == hipSuccess) e = mlp_alloc(m, &m->wt_bf[l], N * K);
    if (e == hipSuccess) e = mlp_alloc(m, &m->w_bf[l], N * K);
    if (e == hipSuccess) e = hipMemcpy(m->w[l], w.data(), N * K * 4, hipMemcpyHostToDevice);
    for (float* p : {m->mw[l], m->vw[l]})
      if (e == hipSuccess) e = hipMemcpy(p, zero.data(), N * K * 4, hipMemcpyHostToDevice);
    for (float* p : {m->b[l], m->mb[l], m->vb[l]})
      if (e == hipSuccess) e = hipMemcpy(p, zero.data(), N * 4, hipMemcpyHostToDevice);
    if (e == hipSuccess) refresh_w_copy(m, l);
  }
  if (e == hipSuccess) e = hipStreamSynchronize(m->st);
  if (e != hipSuccess) {
    ndt_hip_mlp_destroy(m);
    return NDT_ERR_HIP;
  }
  *out = m;
  return NDT_OK;
}

void ndt_hip_mlp_destroy(ndt_mlp* m) {
  if (!m) return;
  if (m->st) (void)hipStreamSynchronize(m->st);
  for (void* p : m->allocs) (void)hipFree(p);
  for (int l = 0; l < m->layers; ++l) {
    if (m->z[l]) (void)hipFree(m->z[l]);
    if (m->act[l]) (void)hipFree(m->act[l]);
  }
  for (void* p : {(void*)m->dz, (void*)m->dy, (void*)m->dout, (void*)m->loss, (void*)m->label, (void*)m->dz_bf,
                  (void*)m->dzt_bf, (void*)m->actt_bf})
    if (p) (void)hipFree(p);
  if (m->out_host) (void)hipHostFree(m->out_host);
  if (m->splitk) (void)hipFree(m->splitk);
  if (m->st) (void)hipStreamDestroy(m->st);
  delete m;
}

const char* ndt_hip_mlp_last_error(const ndt_mlp* m) { return m ? m->err.c_str() : "null network"; }

size_t ndt_hip_mlp_num_params(const ndt_mlp* m) {
  size_t n = 0;
  for (int l = 0; m && l < m->layers; ++l) n += (size_t)m->sizes[l + 1] * (m->sizes[l] + 1);
  return n;
}

int ndt_hip_mlp_get_params(ndt_mlp* m, float* params) {
  if (!m || !params) return NDT_ERR_INVALID_ARGUMENT;
  MLP_HIP(hipStreamSynchronize(m->st), "sync");
  for (int l = 0; l < m->layers; ++l) {
    const size_t K = m->sizes[l], N = m->sizes[l + 1];
    MLP_HIP(hipMemcpy(params, m->w[l], N * K * 4, hipMemcpyDeviceToHost), "copy w");
    params += N * K;
    MLP_HIP(hipMemcpy(params, m->b[l], N * 4, hipMemcpyDeviceToHost), "copy b");
    params += N;
  }
  return NDT_OK;
}

int ndt_hip_mlp_set_params(ndt_mlp* m, const float* params) {
  if (!m || !params) return NDT_ERR_INVALID_ARGUMENT;
  for (int l = 0; l < m->layers; ++l) {
    const size_t K = m->sizes[l], N = m->sizes[l + 1];
    MLP_HIP(hipMemcpy(m->w[l], params, N * K * 4, hipMemcpyHostToDevice), "copy w");
    params += N * K;
    MLP_HIP(hipMemcpy(m->b[l], params, N * 4, hipMemcpyHostToDevice), "copy b");
    params += N;
    refresh_w_copy(m, l);
  }
  MLP_HIP(hipStreamSynchronize(m->st), "sync");
  return NDT_OK;
}

int ndt_hip_mlp_forward_device(ndt_mlp* m, const uint16_t* d_x, size_t batch, float* out) {
  if (!m || !d_x || !out || !shape_ok(m, batch)) return NDT_ERR_INVALID_ARGUMENT;
  const int B = (int)batch;
  int rc = ensure_batch(m, B);
  if (rc) return rc;
  MLP_HIP(hipMemcpyAsync(m->act[0], d_x, (size_t)B * m->sizes[0] * 2, hipMemcpyDeviceToDevice, m->st), "copy x");
  forward(m, B);
  const size_t no = (size_t)B * m->sizes.back();
  MLP_HIP(hipMemcpyAsync(m->out_host, m->z[m->layers - 1], no * 4, hipMemcpyDeviceToHost, m->st), "copy out");
  MLP_HIP(hipStreamSynchronize(m->st), "forward");
  std::memcpy(out, m->out_host, no * 4);
  return NDT_OK;
}

int ndt_hip_mlp_forward_backward(ndt_mlp* m, const float* x, size_t batch, const float* dout, float* out,
                                 float* grads) {
  if (!m || !x || !dout || !out || !grads || !shape_ok(m, batch)) return NDT_ERR_INVALID_ARGUMENT;
  const int B = (int)batch;
  int rc = ensure_batch(m, B);
  if (rc) return rc;
  const size_t nx = (size_t)B * m->sizes[0], no = (size_t)B * m->sizes.back();
  std::vector<uint16_t> xb(nx);
  for (size_t i = 0; i < nx; ++i) xb[i] = f2bf(x[i]);
  MLP_HIP(hipMemcpyAsync(m->act[0], xb.data(), nx * 2, hipMemcpyHostToDevice, m->st), "copy x");
  MLP_HIP(hipMemcpyAsync(m->dout, dout, no * 4, hipMemcpyHostToDevice, m->st), "copy dout");
  forward(m, B);
  backward(m, B);
  MLP_HIP(hipStreamSynchronize(m->st), "forward/backward");
  MLP_HIP(hipMemcpy(out, m->z[m->layers - 1], no * 4, hipMemcpyDeviceToHost), "copy out");
  for (int l = 0; l < m->layers; ++l) {
    const size_t K = m->sizes[l], N = m->sizes[l + 1];
    MLP_HIP(hipMemcpy(grads, m->gw[l], N * K * 4, hipMemcpyDeviceToHost), "copy gw");
    grads += N * K;
    MLP_HIP(hipMemcpy(grads, m->gb[l], N * 4, hipMemcpyDeviceToHost), "copy gb");
    grads += N;
  }
  return NDT_OK;
}

int ndt_hip_mlp_train_step_device(ndt_mlp* m, const uint16_t* d_x, const float* labels, size_t batch,
                                  float learning_rate, float* mean_loss) {
  if (!m || !d_x || !labels || !mean_loss || !shape_ok(m, batch) || m->sizes.back() % 4)
    return NDT_ERR_INVALID_ARGUMENT;
  const int B = (int)batch, H = m->sizes.back() / 4;
  int rc = ensure_batch(m, B);
  if (rc) return rc;
  MLP_HIP(hipMemcpyAsync(m->act[0], d_x, (size_t)B * m->sizes[0] * 2, hipMemcpyDeviceToDevice, m->st), "copy x");
  MLP_HIP(hipMemcpyAsync(m->label, labels, (size_t)B * 12, hipMemcpyHostToDevice, m->st), "copy labels");
  forward(m, B);
  hipLaunchKernelGGL(wta_loss, dim3((B + 63) / 64), dim3(64), 0, m->st, m->z[m->layers - 1], m->label, m->dout,
                     m->loss, B, H, 0.05f);
  backward(m, B);
  ++m->step;
  const float b1 = 0.9f, b2 = 0.999f;
  const float c1 = 1.0f - std::pow(b1, (float)m->step), c2 = 1.0f - std::pow(b2, (float)m->step);
  for (int l = 0; l < m->layers; ++l) {
    const size_t K = m->sizes[l], N = m->sizes[l + 1];
    hipLaunchKernelGGL(adam_step, dim3(grid1(N * K)), dim3(256), 0, m->st, m->w[l], m->gw[l], m->mw[l], m->vw[l],
                       m->wt_bf[l], N * K, learning_rate, b1, b2, c1, c2, 1e-8f);
    hipLaunchKernelGGL(adam_step, dim3(grid1(N)), dim3(256), 0, m->st, m->b[l], m->gb[l], m->mb[l], m->vb[l],
                       (uint16_t*)nullptr, N, learning_rate, b1, b2, c1, c2, 1e-8f);
    if (l > 0)
      hipLaunchKernelGGL(transpose_to_bf16<float>, dim3((K + 31) / 32, (N + 31) / 32), dim3(256), 0, m->st, m->w[l],
                         m->w_bf[l], (int)N, (int)K);
  }
  MLP_HIP(hipMemcpyAsync(m->out_host, m->loss, (size_t)B * 4, hipMemcpyDeviceToHost, m->st), "copy loss");
  MLP_HIP(hipStreamSynchronize(m->st), "train step");
  double s = 0.0;
  for (int i = 0; i < B; ++i) s += m->out_host[i];
  *mean_loss = (float)(s / B);
  return NDT_OK;
}

int ndt_hip_bev_device(const float* d_xyz, const uint32_t* d_offsets, const uint32_t* d_index, size_t rows,
                       const float* d_rot, int32_t nx, int32_t ny, int32_t nz, float cell, float zmin, float zcell,
                       uint16_t* d_bev) {
  if (!d_xyz || !d_offsets || !d_bev || rows == 0 || rows > 0x7fffffffu || nx <= 0 || ny <= 0 || nz <= 0 ||
      (size_t)nx * ny * nz > 32768 || !(cell > 0) || !(zcell > 0))
    return NDT_ERR_INVALID_ARGUMENT;
  if (d_rot && (reinterpret_cast<uintptr_t>(d_rot) & 7u)) return NDT_ERR_INVALID_ARGUMENT;
  const size_t lds = (size_t)nx * ny * nz * 4;
  if (hipFuncSetAttribute(reinterpret_cast<const void*>(bev_kernel), hipFuncAttributeMaxDynamicSharedMemorySize,
                          (int)lds) != hipSuccess)
    return NDT_ERR_HIP;
  hipLaunchKernelGGL(bev_kernel, dim3((unsigned)rows), dim3(256), lds, nullptr, d_xyz, d_offsets, d_index,
                     reinterpret_cast<const float2*>(d_rot), d_bev, nx, ny, nz, cell, zmin, zcell);
  hipError_t e = hipGetLastError();
  // The null stream only: other handles' non-blocking streams keep running.
  if (e == hipSuccess) e = hipStreamSynchronize(nullptr);
  return e == hipSuccess ? NDT_OK : NDT_ERR_HIP;
}

}  // extern "C"
