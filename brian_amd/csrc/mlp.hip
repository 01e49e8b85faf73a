// Learned proposal for model fitting: a 3-layer MLP (nin -> nh -> nh -> nout)
// with ReLU, trained on simulator-generated (spike-train features,
// parameters) pairs with a Gaussian negative log-likelihood.  No counterpart
// in the reference (which initialises the swarm uniformly at random,
// brian/library/modelfitting/modelfitting.py:232-518); plain PSO remains
// the fallback.
//
// Every product runs on gfx950's matrix cores (v_mfma_f32_32x32x16_bf16:
// BF16 operands, FP32 accumulation).  One training step is three kernels:
//   k_mlp_rows   one workgroup per 32 batch rows runs the whole forward
//                pass, the loss and the backward pass down to dZ1 with the
//                activations in LDS; weights stream from L2 as BF16 with the
//                reduction dimension contiguous.  It writes the operands of
//                the weight gradients transposed (feature-major, batch
//                contiguous) for the next kernel.
//   k_mlp_wgrad  dW = P^T Q and the bias sums over slices of the batch, one
//                wavefront per 32x32 output tile and slice, partial sums per
//                slice (no atomics: the reduction order is fixed).
//   k_mlp_adam   sums the slices, applies Adam to the FP32 master weights
//                and refreshes the BF16 copies the other kernels read.
// The step counter and the batch position live on the device, so a chunk
// of steps is captured once in a hipGraph and replayed.
#include "common.hpp"
#include "../../include/brianmi.h"

#include <cmath>
#include <memory>
#include <mutex>
#include <random>
#include <vector>

namespace bm {

typedef __bf16 bf16x8_t __attribute__((ext_vector_type(8)));
typedef float f32x16_t __attribute__((ext_vector_type(16)));
typedef uint32_t u32x4_t __attribute__((ext_vector_type(4)));

constexpr int MR = 32;          // batch rows per k_mlp_rows workgroup
constexpr int RT = 512;         // threads of a k_mlp_rows workgroup (8 wavefronts)
constexpr int RW = RT / 64;
constexpr int NO = 32;          // padded output width (2 nu <= 32)
constexpr int MAXNI = 256, MAXNH = 256;
constexpr int WG_SLICES = 8;    // batch slices of the weight-gradient kernel

__host__ __device__ inline uint16_t f2bf(float f) {
  // round to nearest even (finite inputs)
  uint32_t u = __builtin_bit_cast(uint32_t, f);
  const uint32_t lsb = (u >> 16) & 1u;
  u += 0x7FFFu + lsb;
  return (uint16_t)(u >> 16);
}
__host__ __device__ inline float bf2f(uint16_t b) {
  return __builtin_bit_cast(float, (uint32_t)b << 16);
}

__device__ inline bf16x8_t ld8(const uint16_t *p) {
  return __builtin_bit_cast(bf16x8_t, *reinterpret_cast<const u32x4_t *>(p));
}

// acc += A[32 x K] . B[K x 32]: A row-major in LDS (row stride lda), B given
// as Bt[n][k] (k contiguous, row stride ldb) starting at output column n0.
// Lane layout of v_mfma_f32_32x32x16_bf16: lane l supplies row l % 32 of A
// and column l % 32 of B at k = 8 (l / 32) .. +7 of each 16-wide K block.
__device__ inline void tile_mma(f32x16_t &acc, const uint16_t *A, int lda, const uint16_t *Bt, int ldb, int n0,
                                int K, int lane) {
  const int r = lane & 31, kh = (lane >> 5) * 8;
  const uint16_t *a = A + r * lda + kh;
  const uint16_t *b = Bt + (size_t)(n0 + r) * ldb + kh;
#pragma unroll 16
  for (int k0 = 0; k0 < K; k0 += 16)
    acc = __builtin_amdgcn_mfma_f32_32x32x16_bf16(ld8(a + k0), ld8(b + k0), acc, 0, 0, 0);
}
// row and column of accumulator element q of a 32x32 tile
__device__ inline int acc_row(int q, int lane) { return 8 * (q >> 2) + (lane >> 5) * 4 + (q & 3); }

struct MlpDev {
  int nin, ni, nh, nu, nout;          // ni: nin rounded up to 32
  int64_t off[7];                     // master parameter offsets (W1 b1 W2 b2 W3 b3 end)
  const float *p;                     // FP32 master parameters
  const uint16_t *w1t, *w2t, *w3t;    // BF16 [nh][ni], [nh][nh], [NO][nh] (forward: k = input)
  const uint16_t *w2, *w3;            // BF16 [nh][nh], [nh][NO] (backward: k = output)
};

// LDS row strides (BF16 elements): 8 elements of padding keep the 16-byte
// operand reads aligned and spread rows over the banks
__host__ __device__ inline int lds_stride(int n) { return n + 8; }

// One workgroup: rows [r0, r0 + 32) of the batch.  X rows are gathered
// through idx (FP32 [n][nin]).  TRAIN: loss per row, and the transposed
// operands of the weight gradients, feature-major with the batch (ldb
// columns) contiguous: XT [ni][B], A1T / A2T / dZ2T / dZ1T [nh][B], dOT [NO][B].
// Forward only: the nout outputs per row into out [B][nout].
template <bool TRAIN>
__global__ __launch_bounds__(RT) void k_mlp_rows(const MlpDev m, const float *__restrict__ X,
                                                  const float *__restrict__ U, const int *__restrict__ idx,
                                                  const int64_t *__restrict__ d_batch, int B, float *out,
                                                  uint16_t *XT, uint16_t *A1T, uint16_t *A2T, uint16_t *dOT,
                                                  uint16_t *dZ2T, uint16_t *dZ1T, float *loss,
                                                  int64_t *d_step) {
  extern __shared__ __align__(16) unsigned char smem[];
  const int tid = threadIdx.x, lane = tid & 63, wave = tid >> 6;
  const int ni = m.ni, nh = m.nh;
  const int sx = lds_stride(ni), sh = lds_stride(nh), so = lds_stride(NO);
  uint16_t *sX = reinterpret_cast<uint16_t *>(smem);          // [32][sx]
  uint16_t *sA1 = sX + MR * sx;                                // [32][sh]
  uint16_t *sA2 = sA1 + MR * sh;                               // [32][sh], later dZ2
  uint16_t *sZ1 = sA2 + MR * sh;                               // [32][sh] dZ1 (TRAIN)
  uint16_t *sdO = sZ1 + MR * sh;                               // [32][so]
  float *sO = reinterpret_cast<float *>(sdO + MR * so);        // [32][NO]
  const int r0 = blockIdx.x * MR;
  const int64_t boff = (TRAIN && d_batch) ? *d_batch * (int64_t)B : 0;
  if (TRAIN && d_step && blockIdx.x == 0 && tid == 0) *d_step += 1;   // Adam step of this batch

  // gather the rows (BF16), zero padding columns
  for (int e = tid; e < MR * ni; e += RT) {
    const int r = e / ni, k = e % ni;
    const int64_t src = idx ? idx[boff + r0 + r] : (int64_t)(r0 + r);
    sX[r * sx + k] = k < m.nin ? f2bf(X[src * m.nin + k]) : (uint16_t)0;
  }
  __syncthreads();
  const float *b1 = m.p + m.off[1], *b2 = m.p + m.off[3], *b3 = m.p + m.off[5];
  // layer 1: A1 = relu(X W1 + b1) (BF16; A1 > 0 is the backward mask)
  for (int t = wave; t < nh / 32; t += RW) {
    f32x16_t acc = {};
    tile_mma(acc, sX, sx, m.w1t, ni, t * 32, ni, lane);
#pragma unroll
    for (int q = 0; q < 16; ++q) {
      const int c = t * 32 + (lane & 31);
      const float z = acc[q] + b1[c];
      sA1[acc_row(q, lane) * sh + c] = f2bf(z > 0.f ? z : 0.f);
    }
  }
  __syncthreads();
  // layer 2
  for (int t = wave; t < nh / 32; t += RW) {
    f32x16_t acc = {};
    tile_mma(acc, sA1, sh, m.w2t, nh, t * 32, nh, lane);
#pragma unroll
    for (int q = 0; q < 16; ++q) {
      const int c = t * 32 + (lane & 31);
      const float z = acc[q] + b2[c];
      sA2[acc_row(q, lane) * sh + c] = f2bf(z > 0.f ? z : 0.f);
    }
  }
  __syncthreads();
  // layer 3: O = A2 W3 + b3 (one tile): the reduction split over four
  // wavefronts, partial tiles summed in a fixed order (sZ1 is free here)
  float *part3 = reinterpret_cast<float *>(sZ1);               // [4][32][32]
  if (wave < 4) {
    const int kq = nh / 4;
    f32x16_t acc = {};
    tile_mma(acc, sA2 + wave * kq, sh, m.w3t + wave * kq, nh, 0, kq, lane);
#pragma unroll
    for (int q = 0; q < 16; ++q) part3[wave * MR * NO + acc_row(q, lane) * NO + (lane & 31)] = acc[q];
  }
  __syncthreads();
  for (int e = tid; e < MR * NO; e += RT) {
    const int c = e % NO;
    sO[e] = ((part3[e] + part3[MR * NO + e]) + (part3[2 * MR * NO + e] + part3[3 * MR * NO + e])) + b3[c];
  }
  __syncthreads();
  if constexpr (!TRAIN) {
    for (int e = tid; e < MR * m.nout; e += RT) {
      const int r = e / m.nout, k = e % m.nout;
      if (r0 + r < B) out[(int64_t)(r0 + r) * m.nout + k] = sO[r * NO + k];
    }
    return;
  } else {
    // Gaussian NLL over nu outputs: O = [mu | s], sigma = exp(s) with s
    // clamped to [-7, 3]; dO (BF16) scaled by 1/B
    if (tid < MR) {
      const int r = tid;
      const int64_t src = idx ? idx[boff + r0 + r] : (int64_t)(r0 + r);
      float L = 0.f;
      const float invB = 1.f / (float)B;
      for (int k = 0; k < NO; ++k) sdO[r * so + k] = 0;
      for (int k = 0; k < m.nu; ++k) {
        const float mu = sO[r * NO + k];
        const float sr = sO[r * NO + m.nu + k];
        const float s = fminf(fmaxf(sr, -7.f), 3.f);
        const float iv = expf(-2.f * s);
        const float e = U[src * m.nu + k] - mu;
        L += 0.5f * e * e * iv + s;
        sdO[r * so + k] = f2bf(-e * iv * invB);
        const float ds = (sr > -7.f && sr < 3.f) ? (1.f - e * e * iv) : 0.f;
        sdO[r * so + m.nu + k] = f2bf(ds * invB);
      }
      // the loss is indexed by the row's position in the batch
      loss[r0 + r] = L;
    }
    // transposed operands of the weight gradients: thread = feature, 32
    // consecutive batch columns (64 bytes) per store group
    auto store_t = [&](const uint16_t *S, int stride, int nfeat, uint16_t *dst) {
      for (int f = tid; f < nfeat; f += RT) {
        u32x4_t *d = reinterpret_cast<u32x4_t *>(dst + (int64_t)f * B + r0);
#pragma unroll
        for (int c = 0; c < MR / 8; ++c) {
          u32x4_t wv;
#pragma unroll
          for (int q = 0; q < 4; ++q)
            wv[q] = (uint32_t)S[(8 * c + 2 * q) * stride + f] | ((uint32_t)S[(8 * c + 2 * q + 1) * stride + f] << 16);
          d[c] = wv;
        }
      }
    };
    store_t(sX, sx, ni, XT);
    store_t(sA1, sh, nh, A1T);
    store_t(sA2, sh, nh, A2T);
    __syncthreads();
    store_t(sdO, so, NO, dOT);
    // dZ2 = (dO W3^T) * (A2 > 0), into the A2 buffer (each element's mask
    // is read by the lane that overwrites it)
    for (int t = wave; t < nh / 32; t += RW) {
      f32x16_t acc = {};
      tile_mma(acc, sdO, so, m.w3, NO, t * 32, NO, lane);
#pragma unroll
      for (int q = 0; q < 16; ++q) {
        const int at = acc_row(q, lane) * sh + t * 32 + (lane & 31);
        sA2[at] = sA2[at] != 0 ? f2bf(acc[q]) : (uint16_t)0;
      }
    }
    __syncthreads();
    store_t(sA2, sh, nh, dZ2T);
    // dZ1 = (dZ2 W2^T) * (A1 > 0)
    for (int t = wave; t < nh / 32; t += RW) {
      f32x16_t acc = {};
      tile_mma(acc, sA2, sh, m.w2, nh, t * 32, nh, lane);
#pragma unroll
      for (int q = 0; q < 16; ++q) {
        const int at = acc_row(q, lane) * sh + t * 32 + (lane & 31);
        sZ1[at] = sA1[at] != 0 ? f2bf(acc[q]) : (uint16_t)0;
      }
    }
    __syncthreads();
    store_t(sZ1, sh, nh, dZ1T);
  }
}

// Weight-gradient tiles: task = (matrix, 32x32 tile, batch slice), one
// wavefront each, writing partial sums part[slice][param].  Bias sums are
// handled by the trailing workgroups: one thread per bias and slice.
struct WgradArgs {
  const uint16_t *XT, *A1T, *A2T, *dOT, *dZ2T, *dZ1T;
  float *part;                 // [WG_SLICES][nparams]
  int64_t nparams;
  int B, ntiles1, ntiles2, ntiles3, nbias_threads;
};

__global__ __launch_bounds__(256) void k_mlp_wgrad(const MlpDev m, const WgradArgs a) {
  const int tid = threadIdx.x, lane = tid & 63;
  const int ks = a.B / WG_SLICES;
  const int ntask = (a.ntiles1 + a.ntiles2 + a.ntiles3) * WG_SLICES;
  const int task = blockIdx.x * 4 + (tid >> 6);
  if (task < ntask) {
    const int slice = task % WG_SLICES;
    int t = task / WG_SLICES;
    const uint16_t *P, *Q;
    int np_out, rows, cols;     // matrix [rows][cols] at offset
    int64_t off;
    if (t < a.ntiles1) {
      P = a.XT; Q = a.dZ1T; rows = m.nin; cols = m.nh; off = m.off[0]; np_out = m.nh / 32;
    } else if ((t -= a.ntiles1) < a.ntiles2) {
      P = a.A1T; Q = a.dZ2T; rows = m.nh; cols = m.nh; off = m.off[2]; np_out = m.nh / 32;
    } else {
      t -= a.ntiles2;
      P = a.A2T; Q = a.dOT; rows = m.nh; cols = NO; off = m.off[4]; np_out = 1;
    }
    const int ti = t / np_out, tj = t % np_out;
    // C[i][j] = sum_b P[i][b] Q[j][b] over the slice: A = P rows (k = b), Bt = Q rows
    f32x16_t acc = {};
    const int r = lane & 31, kh = (lane >> 5) * 8;
    const uint16_t *pa = P + (int64_t)(ti * 32 + r) * a.B + slice * ks + kh;
    const uint16_t *pb = Q + (int64_t)(tj * 32 + r) * a.B + slice * ks + kh;
#pragma unroll 4
    for (int k0 = 0; k0 < ks; k0 += 16)
      acc = __builtin_amdgcn_mfma_f32_32x32x16_bf16(ld8(pa + k0), ld8(pb + k0), acc, 0, 0, 0);
    float *dst = a.part + (int64_t)slice * a.nparams + off;
#pragma unroll
    for (int q = 0; q < 16; ++q) {
      const int i = ti * 32 + acc_row(q, lane), j = tj * 32 + (lane & 31);
      if (i < rows) dst[(int64_t)i * cols + j] = acc[q];
    }
    return;
  }
  // bias sums: thread -> (slice, bias index over b1 | b2 | b3)
  const int bt = (blockIdx.x - (ntask + 3) / 4) * 256 + tid;
  if (bt < 0 || bt >= a.nbias_threads) return;
  const int nb = 2 * m.nh + NO;
  const int slice = bt / nb;
  int j = bt % nb;
  const uint16_t *Q;
  int64_t off;
  if (j < m.nh) { Q = a.dZ1T; off = m.off[1]; }
  else if ((j -= m.nh) < m.nh) { Q = a.dZ2T; off = m.off[3]; }
  else { j -= m.nh; Q = a.dOT; off = m.off[5]; }
  const uint16_t *q = Q + (int64_t)j * a.B + slice * ks;
  float s = 0.f;
  for (int b = 0; b < ks; b += 8) {
    const u32x4_t v = *reinterpret_cast<const u32x4_t *>(q + b);
#pragma unroll
    for (int c = 0; c < 4; ++c) s += bf2f((uint16_t)(v[c] & 0xFFFFu)) + bf2f((uint16_t)(v[c] >> 16));
  }
  a.part[(int64_t)slice * a.nparams + off + j] = s;
}

// Sums the slices (fixed order) into the gradient, then (adam) the Adam
// step with bias correction for the step counted on the device and the
// BF16 copies of the weights.  Without adam the gradient is written to g.
struct AdamArgs {
  float *p, *g, *m, *v;
  const float *part;
  int64_t n;
  float lr, b1, b2, eps;
  const int64_t *d_step;
  int64_t *d_batch;            // advanced by one (next batch) when non-null
  uint16_t *w1t, *w2t, *w3t, *w2, *w3;
  int adam;
};

__global__ __launch_bounds__(256) void k_mlp_adam(const MlpDev md, const AdamArgs a) {
  const int64_t i = (int64_t)blockIdx.x * 256 + threadIdx.x;
  if (a.d_batch && i == 0) *a.d_batch += 1;
  if (i >= a.n) return;
  float gi = 0.f;
#pragma unroll
  for (int s = 0; s < WG_SLICES; ++s) gi += a.part[(int64_t)s * a.n + i];
  if (!a.adam) {
    a.g[i] = gi;
    return;
  }
  const float t = (float)*a.d_step;
  const float c1 = 1.f - powf(a.b1, t), c2 = 1.f - powf(a.b2, t);
  const float mi = a.b1 * a.m[i] + (1.f - a.b1) * gi;
  const float vi = a.b2 * a.v[i] + (1.f - a.b2) * gi * gi;
  a.m[i] = mi;
  a.v[i] = vi;
  const float w = a.p[i] - a.lr * (mi / c1) / (sqrtf(vi / c2) + a.eps);
  a.p[i] = w;
  const uint16_t wb = f2bf(w);
  // BF16 copies: forward layouts are [out][in], backward [in][out]
  if (i < md.off[1]) {
    const int64_t r = i / md.nh, c = i % md.nh;       // W1[r][c]
    a.w1t[c * md.ni + r] = wb;
  } else if (i >= md.off[2] && i < md.off[3]) {
    const int64_t q = i - md.off[2], r = q / md.nh, c = q % md.nh;
    a.w2t[c * md.nh + r] = wb;
    a.w2[q] = wb;
  } else if (i >= md.off[4] && i < md.off[5]) {
    const int64_t q = i - md.off[4], r = q / NO, c = q % NO;
    a.w3t[c * md.nh + r] = wb;
    a.w3[q] = wb;
  }
}

// BF16 copies of the master weights in both layouts (W1T padding rows zero)
__global__ void k_mlp_cast(const MlpDev md, uint16_t *w1t, uint16_t *w2t, uint16_t *w3t, uint16_t *w2,
                           uint16_t *w3) {
  const int64_t i = (int64_t)blockIdx.x * 256 + threadIdx.x;
  const float *p = md.p;
  if (i < (int64_t)md.nh * md.ni) {
    const int64_t c = i / md.ni, r = i % md.ni;       // w1t[c][r] = W1[r][c]
    w1t[i] = r < md.nin ? f2bf(p[md.off[0] + r * md.nh + c]) : (uint16_t)0;
  }
  if (i < (int64_t)md.nh * md.nh) {
    w2[i] = f2bf(p[md.off[2] + i]);
    const int64_t c = i / md.nh, r = i % md.nh;       // w2t[c][r] = W2[r][c]
    w2t[i] = f2bf(p[md.off[2] + r * md.nh + c]);
  }
  if (i < (int64_t)md.nh * NO) {
    w3[i] = f2bf(p[md.off[4] + i]);
    const int64_t c = i / md.nh, r = i % md.nh;       // w3t[c][r] = W3[r][c]
    w3t[i] = f2bf(p[md.off[4] + r * NO + c]);
  }
}

// Mean of the per-row losses (one workgroup, fixed summation order).
__global__ __launch_bounds__(256) void k_mlp_loss_mean(const float *__restrict__ loss, int B, float *out) {
  __shared__ float sh[256];
  float s = 0.f;
  for (int i = threadIdx.x; i < B; i += 256) s += loss[i];
  sh[threadIdx.x] = s;
  __syncthreads();
  for (int w = 128; w > 0; w >>= 1) {
    if ((int)threadIdx.x < w) sh[threadIdx.x] += sh[threadIdx.x + w];
    __syncthreads();
  }
  if (threadIdx.x == 0) *out = sh[0] / (float)B;
}

// Scratch of bm_mlp_grad_noalloc: the per-step buffers of a batch of up to
// NA_MAXB rows for nin <= NA_MAXNI and nh <= MAXNH, one copy per GPU for the
// life of the process.
constexpr int NA_MAXB = 1024, NA_MAXNI = 128;
constexpr int64_t NA_MAXPARAMS =
    (int64_t)NA_MAXNI * MAXNH + MAXNH + (int64_t)MAXNH * MAXNH + MAXNH + (int64_t)MAXNH * NO + NO;
constexpr size_t NA_BF16 = (size_t)NA_MAXB * (NA_MAXNI + 4 * MAXNH + NO);          // BF16 elements
constexpr size_t NA_BYTES = NA_BF16 * 2 + (size_t)NA_MAXB * 4 + (size_t)WG_SLICES * NA_MAXPARAMS * 4;
static __device__ __align__(256) unsigned char scratch[NA_BYTES];

}  // namespace bm

using namespace bm;

struct bm_mlp {
  int nin = 64, ni = 64, nh = 256, nout = 8, nu = 4;
  // parameters (FP32 master) as one flat buffer: W1 b1 W2 b2 W3 b3 (W3/b3 padded to NO)
  int64_t off[7] = {};
  DevBuf p, g, m, v;
  DevBuf w1t, w2t, w3t, w2, w3;     // BF16 copies
  DevBuf data_x, data_u;
  int64_t ndata = 0;
  // per-batch operands for the batch capacity
  int cap = 0;
  DevBuf xb, ub, outb, XT, A1T, A2T, dOT, dZ2T, dZ1T, loss, part, idx;
  DevBuf d_step, d_batch;           // int64 counters
  int64_t step = 0;
  hipStream_t stream = nullptr;
  hipGraphExec_t graph = nullptr;
  int graph_batch = 0, graph_steps = 0;
  float graph_lr = 0.f;
  ~bm_mlp() {
    if (graph) (void)hipGraphExecDestroy(graph);
    if (stream) (void)hipStreamDestroy(stream);
  }
  MlpDev dev() const {
    MlpDev d{};
    d.nin = nin;
    d.ni = ni;
    d.nh = nh;
    d.nu = nu;
    d.nout = nout;
    for (int q = 0; q < 7; ++q) d.off[q] = off[q];
    d.p = p.as<float>();
    d.w1t = w1t.as<uint16_t>();
    d.w2t = w2t.as<uint16_t>();
    d.w3t = w3t.as<uint16_t>();
    d.w2 = w2.as<uint16_t>();
    d.w3 = w3.as<uint16_t>();
    return d;
  }
};

namespace {

int64_t nparams(const bm_mlp *h) {
  return (int64_t)h->nin * h->nh + h->nh + (int64_t)h->nh * h->nh + h->nh + (int64_t)h->nh * NO + NO;
}

size_t rows_shmem(const bm_mlp *h) {
  return (size_t)MR * (lds_stride(h->ni) + 3 * lds_stride(h->nh) + lds_stride(NO)) * 2 + (size_t)MR * NO * 4;
}

void drop_graph(bm_mlp *h) {
  if (h->graph) (void)hipGraphExecDestroy(h->graph);
  h->graph = nullptr;
  h->graph_steps = 0;
}

void ensure_cap(bm_mlp *h, int B) {
  if (B <= h->cap) return;
  drop_graph(h);           // a captured training chunk holds the old buffers
  const int c = (B + 255) / 256 * 256;
  h->xb.alloc((size_t)c * h->nin * 4);
  h->ub.alloc((size_t)c * h->nu * 4);
  h->outb.alloc((size_t)c * h->nout * 4);
  h->XT.alloc((size_t)c * h->ni * 2);
  h->A1T.alloc((size_t)c * h->nh * 2);
  h->A2T.alloc((size_t)c * h->nh * 2);
  h->dOT.alloc((size_t)c * NO * 2);
  h->dZ2T.alloc((size_t)c * h->nh * 2);
  h->dZ1T.alloc((size_t)c * h->nh * 2);
  h->loss.alloc((size_t)c * 4);
  h->cap = c;
}

void cast_weights(bm_mlp *h) {
  const int64_t n = std::max<int64_t>((int64_t)h->nh * h->ni, (int64_t)h->nh * h->nh);
  hipLaunchKernelGGL(k_mlp_cast, dim3((unsigned)((n + 255) / 256)), dim3(256), 0, h->stream, h->dev(),
                     h->w1t.as<uint16_t>(), h->w2t.as<uint16_t>(), h->w3t.as<uint16_t>(), h->w2.as<uint16_t>(),
                     h->w3.as<uint16_t>());
}

// forward + loss + backward of one batch of B rows (B multiple of 256 for
// the batch slices, rows gathered from X through idx or taken in order),
// then either Adam (adam = 1, lr) or the summed gradient into g.
// Device buffers one step works in: the transposed operands of the weight
// gradients, the per-row loss, the per-slice partial sums and the gradient
// (written when the step does not apply Adam).
struct StepBufs {
  uint16_t *XT, *A1T, *A2T, *dOT, *dZ2T, *dZ1T;
  float *loss, *part, *g;
};

StepBufs handle_bufs(bm_mlp *h) {
  return {h->XT.as<uint16_t>(), h->A1T.as<uint16_t>(), h->A2T.as<uint16_t>(), h->dOT.as<uint16_t>(),
          h->dZ2T.as<uint16_t>(), h->dZ1T.as<uint16_t>(), h->loss.as<float>(), h->part.as<float>(),
          h->g.as<float>()};
}

void step(bm_mlp *h, const StepBufs &sb, const float *X, const float *U, const int *idx, const int64_t *d_batch,
          int B, int adam, float lr) {
  const MlpDev md = h->dev();
  hipLaunchKernelGGL(k_mlp_rows<true>, dim3(B / MR), dim3(RT), rows_shmem(h), h->stream, md, X, U, idx, d_batch,
                     B, nullptr, sb.XT, sb.A1T, sb.A2T, sb.dOT, sb.dZ2T, sb.dZ1T, sb.loss,
                     adam ? h->d_step.as<int64_t>() : nullptr);
  WgradArgs w{};
  w.XT = sb.XT;
  w.A1T = sb.A1T;
  w.A2T = sb.A2T;
  w.dOT = sb.dOT;
  w.dZ2T = sb.dZ2T;
  w.dZ1T = sb.dZ1T;
  w.part = sb.part;
  w.nparams = h->off[6];
  w.B = B;
  w.ntiles1 = (h->ni / 32) * (h->nh / 32);
  w.ntiles2 = (h->nh / 32) * (h->nh / 32);
  w.ntiles3 = h->nh / 32;
  w.nbias_threads = WG_SLICES * (2 * h->nh + NO);
  const int ntask = (w.ntiles1 + w.ntiles2 + w.ntiles3) * WG_SLICES;
  const int blocks = (ntask + 3) / 4 + (w.nbias_threads + 255) / 256;
  hipLaunchKernelGGL(k_mlp_wgrad, dim3(blocks), dim3(256), 0, h->stream, md, w);
  AdamArgs a{};
  a.p = h->p.as<float>();
  a.g = sb.g;
  a.m = h->m.as<float>();
  a.v = h->v.as<float>();
  a.part = sb.part;
  a.n = h->off[6];
  a.lr = lr;
  a.b1 = 0.9f;
  a.b2 = 0.999f;
  a.eps = 1e-8f;
  a.d_step = h->d_step.as<int64_t>();
  a.d_batch = adam && d_batch ? const_cast<int64_t *>(d_batch) : nullptr;
  a.w1t = h->w1t.as<uint16_t>();
  a.w2t = h->w2t.as<uint16_t>();
  a.w3t = h->w3t.as<uint16_t>();
  a.w2 = h->w2.as<uint16_t>();
  a.w3 = h->w3.as<uint16_t>();
  a.adam = adam;
  hipLaunchKernelGGL(k_mlp_adam, dim3((unsigned)((a.n + 255) / 256)), dim3(256), 0, h->stream, md, a);
}

}  // namespace

extern "C" {

int bm_mlp_create(int nin, int nh, int nu, const float *params, bm_mlp **out) {
  return guard([&] {
    BM_REQUIRE(nin > 0 && nin <= MAXNI, "nin must be in [1, 256]");
    BM_REQUIRE(nh % 64 == 0 && nh > 0 && nh <= MAXNH, "nh must be a multiple of 64 in [64, 256]");
    BM_REQUIRE(nu > 0 && 2 * nu <= NO, "nu must be in [1, 16]");
    auto h = std::make_unique<bm_mlp>();
    h->nin = nin;
    h->ni = (nin + 31) / 32 * 32;
    h->nh = nh;
    h->nu = nu;
    h->nout = 2 * nu;
    BM_HIP(hipStreamCreateWithFlags(&h->stream, hipStreamNonBlocking));
    int64_t o = 0;
    h->off[0] = o; o += (int64_t)nin * nh;
    h->off[1] = o; o += nh;
    h->off[2] = o; o += (int64_t)nh * nh;
    h->off[3] = o; o += nh;
    h->off[4] = o; o += (int64_t)nh * NO;
    h->off[5] = o; o += NO;
    h->off[6] = o;
    BM_REQUIRE(rows_shmem(h.get()) <= 65536, "nin and nh too large for the row kernel's 64 KiB of LDS");
    const int64_t n = o;
    h->p.alloc(n * 4);
    h->g.alloc(n * 4);
    h->m.alloc(n * 4);
    h->v.alloc(n * 4);
    h->part.alloc((size_t)WG_SLICES * n * 4);
    BM_HIP(hipMemsetAsync(h->part.p, 0, (size_t)WG_SLICES * n * 4, h->stream));
    h->w1t.alloc((size_t)nh * h->ni * 2);
    h->w2t.alloc((size_t)nh * nh * 2);
    h->w3t.alloc((size_t)NO * nh * 2);
    h->w2.alloc((size_t)nh * nh * 2);
    h->w3.alloc((size_t)nh * NO * 2);
    h->d_step.alloc(8);
    h->d_batch.alloc(8);
    BM_HIP(hipMemcpyAsync(h->p.p, params, n * 4, hipMemcpyHostToDevice, h->stream));
    BM_HIP(hipMemsetAsync(h->m.p, 0, n * 4, h->stream));
    BM_HIP(hipMemsetAsync(h->v.p, 0, n * 4, h->stream));
    BM_HIP(hipMemsetAsync(h->g.p, 0, n * 4, h->stream));
    BM_HIP(hipMemsetAsync(h->d_step.p, 0, 8, h->stream));
    BM_HIP(hipMemsetAsync(h->d_batch.p, 0, 8, h->stream));
    cast_weights(h.get());
    BM_HIP(hipStreamSynchronize(h->stream));
    *out = h.release();
  });
}

int bm_mlp_num_params(int nin, int nh, int nu, int64_t *n) {
  return guard([&] {
    bm_mlp t;
    t.nin = nin;
    t.nh = nh;
    t.nu = nu;
    *n = nparams(&t);
  });
}

int bm_mlp_destroy(bm_mlp *h) {
  return guard([&] {
    if (!h) return;
    (void)hipStreamSynchronize(h->stream);
    delete h;
  });
}

int bm_mlp_forward(bm_mlp *h, const float *X, int B, float *out) {
  return guard([&] {
    BM_REQUIRE(B >= 0, "B must be >= 0");
    if (B == 0) return;
    const int Bp = (B + MR - 1) / MR * MR;
    ensure_cap(h, Bp);
    BM_HIP(hipMemsetAsync(h->xb.p, 0, (size_t)Bp * h->nin * 4, h->stream));
    BM_HIP(hipMemcpyAsync(h->xb.p, X, (size_t)B * h->nin * 4, hipMemcpyHostToDevice, h->stream));
    hipLaunchKernelGGL(k_mlp_rows<false>, dim3(Bp / MR), dim3(RT), rows_shmem(h), h->stream, h->dev(),
                       h->xb.as<float>(), nullptr, nullptr, nullptr, B, h->outb.as<float>(), nullptr, nullptr,
                       nullptr, nullptr, nullptr, nullptr, nullptr, nullptr);
    BM_HIP(hipGetLastError());
    BM_HIP(hipMemcpyAsync(out, h->outb.p, (size_t)B * h->nout * 4, hipMemcpyDeviceToHost, h->stream));
    BM_HIP(hipStreamSynchronize(h->stream));
  });
}

int bm_mlp_grad(bm_mlp *h, const float *X, const float *U, int B, double *loss, float *grads) {
  return guard([&] {
    BM_REQUIRE(B > 0 && B % 256 == 0, "gradient batches must be a positive multiple of 256");
    ensure_cap(h, B);
    BM_HIP(hipMemcpyAsync(h->xb.p, X, (size_t)B * h->nin * 4, hipMemcpyHostToDevice, h->stream));
    BM_HIP(hipMemcpyAsync(h->ub.p, U, (size_t)B * h->nu * 4, hipMemcpyHostToDevice, h->stream));
    step(h, handle_bufs(h), h->xb.as<float>(), h->ub.as<float>(), nullptr, nullptr, B, 0, 0.f);
    BM_HIP(hipGetLastError());
    std::vector<float> l(B);
    BM_HIP(hipMemcpyAsync(l.data(), h->loss.p, (size_t)B * 4, hipMemcpyDeviceToHost, h->stream));
    BM_HIP(hipMemcpyAsync(grads, h->g.p, (size_t)h->off[6] * 4, hipMemcpyDeviceToHost, h->stream));
    BM_HIP(hipStreamSynchronize(h->stream));
    double s = 0;
    for (float x : l) s += x;
    *loss = s / B;
  });
}

int bm_mlp_set_data(bm_mlp *h, const float *X, const float *U, int64_t n) {
  return guard([&] {
    BM_REQUIRE(n > 0, "empty data set");
    h->data_x.alloc((size_t)n * h->nin * 4);
    h->data_u.alloc((size_t)n * h->nu * 4);
    h->idx.alloc((size_t)n * 4);
    BM_HIP(hipMemcpyAsync(h->data_x.p, X, (size_t)n * h->nin * 4, hipMemcpyHostToDevice, h->stream));
    BM_HIP(hipMemcpyAsync(h->data_u.p, U, (size_t)n * h->nu * 4, hipMemcpyHostToDevice, h->stream));
    BM_HIP(hipStreamSynchronize(h->stream));
    h->ndata = n;
    drop_graph(h);
  });
}

// Adam on shuffled mini-batches of the resident data set.  The batches of
// an epoch read consecutive slices of a device permutation; the batch
// position and step count advance on the device, so chunks of steps replay
// as one captured graph.
int bm_mlp_train(bm_mlp *h, int epochs, int batch, float lr, uint64_t seed, double *last_loss) {
  return guard([&] {
    BM_REQUIRE(batch > 0 && batch % 256 == 0 && h->ndata >= batch,
               "batch must be a positive multiple of 256, at most the data size");
    ensure_cap(h, batch);
    std::mt19937_64 rng(seed);
    std::vector<int> perm(h->ndata);
    const int64_t nb = h->ndata / batch;
    const int chunk = (int)std::min<int64_t>(nb, 64);
    if (h->graph && (h->graph_batch != batch || h->graph_steps != chunk || h->graph_lr != lr)) drop_graph(h);
    if (!h->graph) {
      hipGraph_t gr;
      BM_HIP(hipStreamBeginCapture(h->stream, hipStreamCaptureModeThreadLocal));
      for (int k = 0; k < chunk; ++k)
        step(h, handle_bufs(h), h->data_x.as<float>(), h->data_u.as<float>(), h->idx.as<int>(),
             h->d_batch.as<int64_t>(), batch, 1, lr);
      hipError_t e = hipStreamEndCapture(h->stream, &gr);
      BM_HIP(e);
      e = hipGraphInstantiate(&h->graph, gr, nullptr, nullptr, 0);
      hipGraphDestroy(gr);
      BM_HIP(e);
      h->graph_batch = batch;
      h->graph_steps = chunk;
      h->graph_lr = lr;
    }
    const int64_t zero = 0;
    for (int e = 0; e < epochs; ++e) {
      for (int64_t i = 0; i < h->ndata; ++i) perm[i] = (int)i;
      std::shuffle(perm.begin(), perm.end(), rng);
      BM_HIP(hipMemcpyAsync(h->idx.p, perm.data(), (size_t)h->ndata * 4, hipMemcpyHostToDevice, h->stream));
      BM_HIP(hipMemcpyAsync(h->d_batch.p, &zero, 8, hipMemcpyHostToDevice, h->stream));
      int64_t done = 0;
      for (; done + chunk <= nb; done += chunk) BM_HIP(hipGraphLaunch(h->graph, h->stream));
      for (; done < nb; ++done)
        step(h, handle_bufs(h), h->data_x.as<float>(), h->data_u.as<float>(), h->idx.as<int>(),
             h->d_batch.as<int64_t>(), batch, 1, lr);
      // the host copies above read perm and zero: finish before reuse
      BM_HIP(hipStreamSynchronize(h->stream));
      h->step += nb;
    }
    std::vector<float> l(batch);
    BM_HIP(hipMemcpyAsync(l.data(), h->loss.p, (size_t)batch * 4, hipMemcpyDeviceToHost, h->stream));
    BM_HIP(hipStreamSynchronize(h->stream));
    double s = 0;
    for (float x : l) s += x;
    *last_loss = s / batch;
  });
}

namespace {
// Device address of this GPU's copy of the scratch array, looked up once
// per device (the array's name is only a host-side handle).
unsigned char *scratch_ptr() {
  static std::mutex mu;
  static std::vector<unsigned char *> ptrs;
  int dev = 0;
  BM_HIP(hipGetDevice(&dev));
  std::lock_guard<std::mutex> lk(mu);
  if ((int)ptrs.size() <= dev) ptrs.resize(dev + 1, nullptr);
  if (!ptrs[dev]) {
    void *p = nullptr;
    BM_HIP(hipGetSymbolAddress(&p, HIP_SYMBOL(scratch)));
    BM_REQUIRE(p != nullptr, "no device address for the gradient scratch");
    ptrs[dev] = static_cast<unsigned char *>(p);
  }
  return ptrs[dev];
}
}  // namespace

int bm_mlp_grad_noalloc(bm_mlp *h, const float *X, const float *U, int B, float *loss, float *grads) {
  return guard([&] {
    BM_REQUIRE(h && X && U && loss && grads, "null argument");
    BM_REQUIRE(B > 0 && B % 256 == 0 && B <= NA_MAXB, "B must be a multiple of 256 in [256, 1024]");
    BM_REQUIRE(h->nin <= NA_MAXNI, "bm_mlp_grad_noalloc supports nin <= 128");
    unsigned char *base = scratch_ptr();
    // buffers at their size for this B, back to back in the scratch
    uint16_t *bf = reinterpret_cast<uint16_t *>(base);
    StepBufs sb;
    sb.XT = bf;
    sb.A1T = sb.XT + (size_t)h->ni * B;
    sb.A2T = sb.A1T + (size_t)h->nh * B;
    sb.dOT = sb.A2T + (size_t)h->nh * B;
    sb.dZ2T = sb.dOT + (size_t)NO * B;
    sb.dZ1T = sb.dZ2T + (size_t)h->nh * B;
    unsigned char *fl = base + NA_BF16 * 2;
    sb.loss = reinterpret_cast<float *>(fl);
    sb.part = reinterpret_cast<float *>(fl + (size_t)NA_MAXB * 4);
    sb.g = grads;
    // every element the kernels read is written earlier in the same call
    // (all batch columns of the operands, all slices of every parameter),
    // so nothing is cleared
    step(h, sb, X, U, nullptr, nullptr, B, 0, 0.f);
    hipLaunchKernelGGL(k_mlp_loss_mean, dim3(1), dim3(256), 0, h->stream, sb.loss, B, loss);
    BM_HIP(hipGetLastError());
    BM_HIP(hipStreamSynchronize(h->stream));
  });
}

int bm_mlp_get_params(bm_mlp *h, float *out) {
  return guard([&] {
    BM_HIP(hipMemcpyAsync(out, h->p.p, (size_t)h->off[6] * 4, hipMemcpyDeviceToHost, h->stream));
    BM_HIP(hipStreamSynchronize(h->stream));
  });
}

}  // extern "C"
