// Learned proposal for model fitting: a 3-layer MLP (nin -> nh -> nh -> nout)
// with ReLU, trained on simulator-generated (spike-train features,
// parameters) pairs with a Gaussian negative log-likelihood.  No counterpart
// in the reference (which initialises the swarm uniformly at random,
// brian/library/modelfitting/modelfitting.py:232-518); plain PSO remains
// the fallback.
//
// Every product runs on the matrix cores: BF16 operands, FP32 accumulation,
// v_mfma_f32_32x32x8_bf16 on 32x32 output tiles, one wavefront per tile.
// Master weights and Adam moments are FP32; BF16 copies feed the GEMMs.
#include "common.hpp"
#include "../../include/brianmi.h"

#include <cmath>
#include <memory>
#include <random>
#include <vector>

namespace bm {

typedef __attribute__((__vector_size__(4 * sizeof(short)))) short bf16x4_t;
typedef __attribute__((__vector_size__(16 * sizeof(float)))) float f32x16_t;

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

// EPI_RELUB applies the ReLU to the BF16 copy only (FP32 keeps the
// pre-activation for the backward mask).
enum Epi { EPI_NONE = 0, EPI_BIAS = 1, EPI_RELU = 2, EPI_MASK = 4, EPI_RELUB = 8 };

// C[M x N] = op(A) . op(B); A is [M x K] (or [K x M] when ta), B is [K x N]
// (or [N x K] when tb), all BF16 row-major with leading dimensions lda/ldb.
// Epilogue: + bias[n] (EPI_BIAS), ReLU (EPI_RELU), times (mask[m][n] > 0)
// (EPI_MASK); writes FP32 C (if non-null) and a BF16 copy Cb (if non-null).
// M, N multiples of 32; K multiple of 8.
__global__ __launch_bounds__(64) void k_gemm_bf16(const uint16_t *__restrict__ A, int lda, int ta,
                                                  const uint16_t *__restrict__ B, int ldb, int tb,
                                                  float *C, uint16_t *Cb, int ldc, int M, int N, int K,
                                                  const float *__restrict__ bias,
                                                  const float *__restrict__ mask, int epi) {
  const int lane = threadIdx.x;
  const int m0 = blockIdx.y * 32, n0 = blockIdx.x * 32;
  const int r = lane & 31, kg = (lane >> 5) * 4;
  f32x16_t acc = {};
  for (int k0 = 0; k0 < K; k0 += 8) {
    bf16x4_t a, b;
#pragma unroll
    for (int v = 0; v < 4; ++v) {
      const int k = k0 + kg + v;
      const int m = m0 + r, n = n0 + r;
      a[v] = (short)(ta ? A[(size_t)k * lda + m] : A[(size_t)m * lda + k]);
      b[v] = (short)(tb ? B[(size_t)n * ldb + k] : B[(size_t)k * ldb + n]);
    }
    acc = __builtin_amdgcn_mfma_f32_32x32x8bf16_1k(a, b, acc, 0, 0, 0);
  }
#pragma unroll
  for (int q = 0; q < 16; ++q) {
    const int m = m0 + 8 * (q / 4) + (lane >> 5) * 4 + (q % 4);
    const int n = n0 + r;
    float x = acc[q];
    if (epi & EPI_BIAS) x += bias[n];
    if (epi & EPI_RELU) x = x > 0.f ? x : 0.f;
    if (epi & EPI_MASK) x = mask[(size_t)m * ldc + n] > 0.f ? x : 0.f;
    if (C) C[(size_t)m * ldc + n] = x;
    if (Cb) Cb[(size_t)m * ldc + n] = f2bf((epi & EPI_RELUB) ? (x > 0.f ? x : 0.f) : x);
  }
}

__global__ void k_cast_bf16(const float *__restrict__ x, uint16_t *y, int64_t n) {
  const int64_t i = (int64_t)blockIdx.x * 256 + threadIdx.x;
  if (i < n) y[i] = f2bf(x[i]);
}

// Gaussian NLL over nu outputs: O = [mu (nu) | s (nu) | padding], sigma = exp(s)
// with s clamped to [-7, 3].  Writes dO (BF16, scaled by 1/B) and per-sample loss.
__global__ void k_nll(const float *__restrict__ O, int ldo, const float *__restrict__ U, int nu, int B,
                      uint16_t *dO, float *loss) {
  const int i = blockIdx.x * 256 + threadIdx.x;
  if (i >= B) return;
  float L = 0.f;
  const float invB = 1.f / (float)B;
  for (int k = 0; k < ldo; ++k) dO[(size_t)i * ldo + k] = 0;
  for (int k = 0; k < nu; ++k) {
    const float mu = O[(size_t)i * ldo + k];
    const float sr = O[(size_t)i * ldo + nu + k];
    const float s = fminf(fmaxf(sr, -7.f), 3.f);
    const float iv = expf(-2.f * s);
    const float e = U[(size_t)i * nu + k] - mu;
    L += 0.5f * e * e * iv + s;
    dO[(size_t)i * ldo + k] = f2bf(-e * iv * invB);
    const float ds = (sr > -7.f && sr < 3.f) ? (1.f - e * e * iv) : 0.f;
    dO[(size_t)i * ldo + nu + k] = f2bf(ds * invB);
  }
  loss[i] = L;
}

// column sums of a BF16 [B x N] matrix into FP32 g[N]
__global__ void k_colsum_bf16(const uint16_t *__restrict__ X, int B, int N, float *g) {
  const int n = blockIdx.x * 256 + threadIdx.x;
  if (n >= N) return;
  float s = 0.f;
  for (int i = 0; i < B; ++i) s += bf2f(X[(size_t)i * N + n]);
  g[n] = s;
}

__global__ void k_adam(float *w, const float *__restrict__ g, float *m, float *v, int64_t n, float lr,
                       float b1, float b2, float c1, float c2, float eps) {
  const int64_t i = (int64_t)blockIdx.x * 256 + threadIdx.x;
  if (i >= n) return;
  const float gi = g[i];
  const float mi = b1 * m[i] + (1.f - b1) * gi;
  const float vi = b2 * v[i] + (1.f - b2) * gi * gi;
  m[i] = mi;
  v[i] = vi;
  w[i] -= lr * (mi / c1) / (sqrtf(vi / c2) + eps);
}

// gathers batch rows idx[0..B) of X (FP32 [n x nin]) into BF16 and of U
__global__ void k_gather(const float *__restrict__ X, const float *__restrict__ U, const int *__restrict__ idx,
                         int B, int nin, int nu, uint16_t *Xb, float *Ub) {
  const int i = blockIdx.x;
  if (i >= B) return;
  const int64_t src = idx[i];
  for (int k = threadIdx.x; k < nin; k += blockDim.x) Xb[(size_t)i * nin + k] = f2bf(X[src * nin + k]);
  for (int k = threadIdx.x; k < nu; k += blockDim.x) Ub[(size_t)i * nu + k] = U[src * nu + k];
}

}  // namespace bm

using namespace bm;

struct bm_mlp {
  int nin = 64, nh = 256, nout = 8, nu = 4, npad = 32;
  // parameters (FP32 master) as one flat buffer: W1 b1 W2 b2 W3 b3 (W3/b3 padded to npad)
  int64_t off[7] = {};
  DevBuf p, g, m, v, wb;           // params, grads, adam m, adam v, bf16 weights (W1, W2, W3)
  DevBuf data_x, data_u;
  int64_t ndata = 0;
  // activations for the current batch capacity
  int cap = 0;
  DevBuf Xb, Ub, Z1, A1b, Z2, A2b, O, dOb, dZ2b, dZ1b, loss, idx;
  int64_t step = 0;
  hipStream_t stream = nullptr;
  ~bm_mlp() {
    if (stream) (void)hipStreamDestroy(stream);
  }
  float *P() { return p.as<float>(); }
  float *G() { return g.as<float>(); }
  uint16_t *W1b() { return wb.as<uint16_t>(); }
  uint16_t *W2b() { return wb.as<uint16_t>() + (size_t)nin * nh; }
  uint16_t *W3b() { return wb.as<uint16_t>() + (size_t)nin * nh + (size_t)nh * nh; }
};

namespace {

int64_t nparams(const bm_mlp *h) {
  return (int64_t)h->nin * h->nh + h->nh + (int64_t)h->nh * h->nh + h->nh + (int64_t)h->nh * h->npad + h->npad;
}

void ensure_cap(bm_mlp *h, int B) {
  if (B <= h->cap) return;
  const int c = (B + 31) / 32 * 32;
  h->Xb.alloc((size_t)c * h->nin * 2);
  h->Ub.alloc((size_t)c * h->nu * 4);
  h->Z1.alloc((size_t)c * h->nh * 4);
  h->A1b.alloc((size_t)c * h->nh * 2);
  h->Z2.alloc((size_t)c * h->nh * 4);
  h->A2b.alloc((size_t)c * h->nh * 2);
  h->O.alloc((size_t)c * h->npad * 4);
  h->dOb.alloc((size_t)c * h->npad * 2);
  h->dZ2b.alloc((size_t)c * h->nh * 2);
  h->dZ1b.alloc((size_t)c * h->nh * 2);
  h->loss.alloc((size_t)c * 4);
  h->idx.alloc((size_t)c * 4);
  h->cap = c;
}

void gemm(bm_mlp *h, const uint16_t *A, int lda, int ta, const uint16_t *B, int ldb, int tb, float *C,
          uint16_t *Cb, int ldc, int M, int N, int K, const float *bias, const float *mask, int epi) {
  BM_REQUIRE(M % 32 == 0 && N % 32 == 0 && K % 8 == 0, "GEMM shape must be 32x32x8-aligned");
  hipLaunchKernelGGL(k_gemm_bf16, dim3(N / 32, M / 32), dim3(64), 0, h->stream, A, lda, ta, B, ldb, tb, C, Cb,
                     ldc, M, N, K, bias, mask, epi);
}

void refresh_bf16(bm_mlp *h) {
  const float *P = h->P();
  auto cast = [&](const float *src, uint16_t *dst, int64_t n) {
    hipLaunchKernelGGL(k_cast_bf16, dim3((unsigned)((n + 255) / 256)), dim3(256), 0, h->stream, src, dst, n);
  };
  cast(P + h->off[0], h->W1b(), (int64_t)h->nin * h->nh);
  cast(P + h->off[2], h->W2b(), (int64_t)h->nh * h->nh);
  cast(P + h->off[4], h->W3b(), (int64_t)h->nh * h->npad);
}

// forward on the batch already in Xb (B rows, B multiple of 32)
void forward(bm_mlp *h, int B) {
  const float *P = h->P();
  // Z1 (FP32, pre-activation) and A1 = bf16(relu(Z1)) from one GEMM
  gemm(h, h->Xb.as<uint16_t>(), h->nin, 0, h->W1b(), h->nh, 0, h->Z1.as<float>(), h->A1b.as<uint16_t>(), h->nh,
       B, h->nh, h->nin, P + h->off[1], nullptr, EPI_BIAS | EPI_RELUB);
  gemm(h, h->A1b.as<uint16_t>(), h->nh, 0, h->W2b(), h->nh, 0, h->Z2.as<float>(), h->A2b.as<uint16_t>(), h->nh,
       B, h->nh, h->nh, P + h->off[3], nullptr, EPI_BIAS | EPI_RELUB);
  gemm(h, h->A2b.as<uint16_t>(), h->nh, 0, h->W3b(), h->npad, 0, h->O.as<float>(), nullptr, h->npad, B, h->npad,
       h->nh, P + h->off[5], nullptr, EPI_BIAS);
}

// loss + gradients into G for the batch in Xb/Ub
void backward(bm_mlp *h, int B) {
  float *G = h->G();
  hipLaunchKernelGGL(k_nll, dim3((B + 255) / 256), dim3(256), 0, h->stream, h->O.as<float>(), h->npad,
                     h->Ub.as<float>(), h->nu, B, h->dOb.as<uint16_t>(), h->loss.as<float>());
  // layer 3: dW3 = A2^T dO ; db3 = colsum(dO) ; dZ2 = (dO W3^T) * (Z2 > 0)
  gemm(h, h->A2b.as<uint16_t>(), h->nh, 1, h->dOb.as<uint16_t>(), h->npad, 0, G + h->off[4], nullptr, h->npad,
       h->nh, h->npad, B, nullptr, nullptr, EPI_NONE);
  hipLaunchKernelGGL(k_colsum_bf16, dim3((h->npad + 255) / 256), dim3(256), 0, h->stream, h->dOb.as<uint16_t>(),
                     B, h->npad, G + h->off[5]);
  gemm(h, h->dOb.as<uint16_t>(), h->npad, 0, h->W3b(), h->npad, 1, nullptr, h->dZ2b.as<uint16_t>(), h->nh, B,
       h->nh, h->npad, nullptr, h->Z2.as<float>(), EPI_MASK);
  // layer 2
  gemm(h, h->A1b.as<uint16_t>(), h->nh, 1, h->dZ2b.as<uint16_t>(), h->nh, 0, G + h->off[2], nullptr, h->nh, h->nh,
       h->nh, B, nullptr, nullptr, EPI_NONE);
  hipLaunchKernelGGL(k_colsum_bf16, dim3((h->nh + 255) / 256), dim3(256), 0, h->stream, h->dZ2b.as<uint16_t>(),
                     B, h->nh, G + h->off[3]);
  gemm(h, h->dZ2b.as<uint16_t>(), h->nh, 0, h->W2b(), h->nh, 1, nullptr, h->dZ1b.as<uint16_t>(), h->nh, B, h->nh,
       h->nh, nullptr, h->Z1.as<float>(), EPI_MASK);
  // layer 1
  gemm(h, h->Xb.as<uint16_t>(), h->nin, 1, h->dZ1b.as<uint16_t>(), h->nh, 0, G + h->off[0], nullptr, h->nh, h->nin,
       h->nh, B, nullptr, nullptr, EPI_NONE);
  hipLaunchKernelGGL(k_colsum_bf16, dim3((h->nh + 255) / 256), dim3(256), 0, h->stream, h->dZ1b.as<uint16_t>(),
                     B, h->nh, G + h->off[1]);
}

void upload_batch(bm_mlp *h, const float *X, const float *U, int B, int Bp) {
  std::vector<uint16_t> xb((size_t)Bp * h->nin, 0);
  for (size_t i = 0; i < (size_t)B * h->nin; ++i) xb[i] = f2bf(X[i]);
  BM_HIP(hipMemcpyAsync(h->Xb.p, xb.data(), xb.size() * 2, hipMemcpyHostToDevice, h->stream));
  if (U) {
    std::vector<float> ub((size_t)Bp * h->nu, 0.f);
    std::copy(U, U + (size_t)B * h->nu, ub.begin());
    BM_HIP(hipMemcpyAsync(h->Ub.p, ub.data(), ub.size() * 4, hipMemcpyHostToDevice, h->stream));
  }
}

}  // namespace

extern "C" {

int bm_mlp_create(int nin, int nh, int nu, const float *params, bm_mlp **out) {
  return guard([&] {
    BM_REQUIRE(nin % 8 == 0 && nin > 0, "nin must be a positive multiple of 8");
    BM_REQUIRE(nh % 32 == 0 && nh > 0, "nh must be a positive multiple of 32");
    BM_REQUIRE(nu > 0 && 2 * nu <= 32, "nu must be in [1, 16]");
    auto h = std::make_unique<bm_mlp>();
    h->nin = nin;
    h->nh = nh;
    h->nu = nu;
    h->nout = 2 * nu;
    h->npad = 32;
    BM_HIP(hipStreamCreateWithFlags(&h->stream, hipStreamNonBlocking));
    int64_t o = 0;
    h->off[0] = o; o += (int64_t)nin * nh;
    h->off[1] = o; o += nh;
    h->off[2] = o; o += (int64_t)nh * nh;
    h->off[3] = o; o += nh;
    h->off[4] = o; o += (int64_t)nh * h->npad;
    h->off[5] = o; o += h->npad;
    h->off[6] = o;
    const int64_t n = o;
    h->p.alloc(n * 4);
    h->g.alloc(n * 4);
    h->m.alloc(n * 4);
    h->v.alloc(n * 4);
    h->wb.alloc(((size_t)nin * nh + (size_t)nh * nh + (size_t)nh * h->npad) * 2);
    BM_HIP(hipMemcpyAsync(h->p.p, params, n * 4, hipMemcpyHostToDevice, h->stream));
    BM_HIP(hipMemsetAsync(h->m.p, 0, n * 4, h->stream));
    BM_HIP(hipMemsetAsync(h->v.p, 0, n * 4, h->stream));
    BM_HIP(hipMemsetAsync(h->g.p, 0, n * 4, h->stream));
    refresh_bf16(h.get());
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
    const int Bp = (B + 31) / 32 * 32;
    ensure_cap(h, Bp);
    upload_batch(h, X, nullptr, B, Bp);
    forward(h, Bp);
    std::vector<float> o((size_t)Bp * h->npad);
    BM_HIP(hipMemcpyAsync(o.data(), h->O.p, o.size() * 4, hipMemcpyDeviceToHost, h->stream));
    BM_HIP(hipStreamSynchronize(h->stream));
    for (int i = 0; i < B; ++i)
      for (int k = 0; k < h->nout; ++k) out[(size_t)i * h->nout + k] = o[(size_t)i * h->npad + k];
  });
}

int bm_mlp_grad(bm_mlp *h, const float *X, const float *U, int B, double *loss, float *grads) {
  return guard([&] {
    BM_REQUIRE(B % 32 == 0, "gradient batches must be a multiple of 32");
    ensure_cap(h, B);
    upload_batch(h, X, U, B, B);
    forward(h, B);
    backward(h, B);
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
    h->data_x.alloc((size_t)n * h->nin * 4);
    h->data_u.alloc((size_t)n * h->nu * 4);
    BM_HIP(hipMemcpyAsync(h->data_x.p, X, (size_t)n * h->nin * 4, hipMemcpyHostToDevice, h->stream));
    BM_HIP(hipMemcpyAsync(h->data_u.p, U, (size_t)n * h->nu * 4, hipMemcpyHostToDevice, h->stream));
    BM_HIP(hipStreamSynchronize(h->stream));
    h->ndata = n;
  });
}

// Adam on shuffled mini-batches of the resident data set.
int bm_mlp_train(bm_mlp *h, int epochs, int batch, float lr, uint64_t seed, double *last_loss) {
  return guard([&] {
    BM_REQUIRE(h->ndata >= batch && batch % 32 == 0 && batch > 0, "batch must be a multiple of 32 <= data size");
    ensure_cap(h, batch);
    std::mt19937_64 rng(seed);
    std::vector<int> perm(h->ndata);
    const float b1 = 0.9f, b2 = 0.999f, eps = 1e-8f;
    const int64_t n = h->off[6];
    double acc = 0;
    int64_t nb_total = 0;
    for (int e = 0; e < epochs; ++e) {
      for (int64_t i = 0; i < h->ndata; ++i) perm[i] = (int)i;
      std::shuffle(perm.begin(), perm.end(), rng);
      const int64_t nb = h->ndata / batch;
      acc = 0;
      for (int64_t bi = 0; bi < nb; ++bi) {
        BM_HIP(hipMemcpyAsync(h->idx.p, perm.data() + bi * batch, (size_t)batch * 4, hipMemcpyHostToDevice,
                              h->stream));
        hipLaunchKernelGGL(k_gather, dim3(batch), dim3(64), 0, h->stream, h->data_x.as<float>(),
                           h->data_u.as<float>(), h->idx.as<int>(), batch, h->nin, h->nu, h->Xb.as<uint16_t>(),
                           h->Ub.as<float>());
        forward(h, batch);
        backward(h, batch);
        ++h->step;
        const float c1 = 1.f - std::pow(b1, (float)h->step), c2 = 1.f - std::pow(b2, (float)h->step);
        hipLaunchKernelGGL(k_adam, dim3((unsigned)((n + 255) / 256)), dim3(256), 0, h->stream, h->P(), h->G(),
                           h->m.as<float>(), h->v.as<float>(), n, lr, b1, b2, c1, c2, eps);
        refresh_bf16(h);
        if (bi == nb - 1 || bi % 64 == 0) {
          std::vector<float> l(batch);
          BM_HIP(hipMemcpyAsync(l.data(), h->loss.p, (size_t)batch * 4, hipMemcpyDeviceToHost, h->stream));
          BM_HIP(hipStreamSynchronize(h->stream));
          double s = 0;
          for (float x : l) s += x;
          acc = s / batch;
        }
        ++nb_total;
      }
    }
    BM_HIP(hipStreamSynchronize(h->stream));
    *last_loss = acc;
  });
}

int bm_mlp_get_params(bm_mlp *h, float *out) {
  return guard([&] {
    BM_HIP(hipMemcpyAsync(out, h->p.p, (size_t)h->off[6] * 4, hipMemcpyDeviceToHost, h->stream));
    BM_HIP(hipStreamSynchronize(h->stream));
  });
}

int bm_mlp_gemm_test(int M, int N, int K, int ta, int tb, const float *A, const float *B, float *C) {
  return guard([&] {
    bm_mlp h;
    BM_HIP(hipStreamCreateWithFlags(&h.stream, hipStreamNonBlocking));
    std::vector<uint16_t> a((size_t)M * K), b((size_t)K * N);
    for (size_t i = 0; i < a.size(); ++i) a[i] = f2bf(A[i]);
    for (size_t i = 0; i < b.size(); ++i) b[i] = f2bf(B[i]);
    DevBuf da, db, dc;
    da.alloc(a.size() * 2);
    db.alloc(b.size() * 2);
    dc.alloc((size_t)M * N * 4);
    BM_HIP(hipMemcpyAsync(da.p, a.data(), a.size() * 2, hipMemcpyHostToDevice, h.stream));
    BM_HIP(hipMemcpyAsync(db.p, b.data(), b.size() * 2, hipMemcpyHostToDevice, h.stream));
    gemm(&h, da.as<uint16_t>(), ta ? M : K, ta, db.as<uint16_t>(), tb ? K : N, tb, dc.as<float>(), nullptr, N, M,
         N, K, nullptr, nullptr, EPI_NONE);
    BM_HIP(hipMemcpyAsync(C, dc.p, (size_t)M * N * 4, hipMemcpyDeviceToHost, h.stream));
    BM_HIP(hipStreamSynchronize(h.stream));
  });
}

}  // extern "C"
