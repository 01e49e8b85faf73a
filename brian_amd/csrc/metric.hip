// Pairwise squared distances of sampled traces on the device: the van Rossum
// metric's distance matrix (brian/monitor.py:1427-1486).  Row i of X holds
// neuron i's filtered spike train sampled every step (T samples);
//   D[i][j] = scale * sum_t (X[i][t] - X[j][t])^2,   scale = dt / tau,
// symmetric with a zero diagonal.  The differences are formed as the
// reference does (no |x|^2 + |y|^2 - 2 x.y expansion, which cancels), so the
// work is n^2 T / 2 subtract/multiply-adds on the FP64 vector units: one
// 64x64 output tile per workgroup of the lower triangle, both row panels
// staged through LDS 32 samples at a time, 4x4 outputs per thread.
#include "common.hpp"

namespace bm {
namespace {

constexpr int TILE = 64;      // output tile edge
constexpr int TK = 32;        // samples per LDS stage
constexpr int MT = 256;       // threads per workgroup: 16 x 16, 4 x 4 outputs each

__global__ __launch_bounds__(MT) void k_pair_sqdist(const double *X, int64_t n, int64_t T, double scale,
                                                    double *D) {
  // lower-triangle tile (bi >= bj) from the linear block index
  const int64_t b = blockIdx.x;
  int64_t bi = (int64_t)((__builtin_sqrt(8.0 * (double)b + 1.0) - 1.0) * 0.5);
  while ((bi + 1) * (bi + 2) / 2 <= b) ++bi;
  while (bi * (bi + 1) / 2 > b) --bi;
  const int64_t bj = b - bi * (bi + 1) / 2;
  const int64_t i0 = bi * TILE, j0 = bj * TILE;

  __shared__ double As[TK][TILE + 1];
  __shared__ double Bs[TK][TILE + 1];
  const int tid = threadIdx.x;
  const int ty = tid >> 4, tx = tid & 15;
  double acc[4][4];
#pragma unroll
  for (int a = 0; a < 4; ++a)
#pragma unroll
    for (int c = 0; c < 4; ++c) acc[a][c] = 0.0;

  for (int64_t t0 = 0; t0 < T; t0 += TK) {
    // stage X[i0 .. i0+64)[t0 .. t0+32) and the j panel; consecutive threads
    // read consecutive samples of one row
    for (int e = tid; e < TILE * TK; e += MT) {
      const int r = e / TK, k = e % TK;
      const int64_t t = t0 + k;
      const int64_t ri = i0 + r, rj = j0 + r;
      As[k][r] = (ri < n && t < T) ? X[ri * T + t] : 0.0;
      Bs[k][r] = (rj < n && t < T) ? X[rj * T + t] : 0.0;
    }
    __syncthreads();
#pragma unroll 4
    for (int k = 0; k < TK; ++k) {
      double a[4], c[4];
#pragma unroll
      for (int q = 0; q < 4; ++q) {
        a[q] = As[k][ty + 16 * q];
        c[q] = Bs[k][tx + 16 * q];
      }
#pragma unroll
      for (int p = 0; p < 4; ++p)
#pragma unroll
        for (int q = 0; q < 4; ++q) {
          const double d = a[p] - c[q];
          acc[p][q] = __builtin_fma(d, d, acc[p][q]);
        }
    }
    __syncthreads();
  }
#pragma unroll
  for (int p = 0; p < 4; ++p)
#pragma unroll
    for (int q = 0; q < 4; ++q) {
      const int64_t i = i0 + ty + 16 * p, j = j0 + tx + 16 * q;
      if (i >= n || j >= n || j > i) continue;
      const double v = i == j ? 0.0 : scale * acc[p][q];
      D[i * n + j] = v;
      D[j * n + i] = v;
    }
}

}  // namespace
}  // namespace bm

using namespace bm;

extern "C" {

int bm_pairwise_sq_distance(const double *X, int64_t n, int64_t T, double scale, double *D) {
  return guard([&] {
    BM_REQUIRE(n >= 0 && T >= 0 && (n == 0 || (X && D)), "bad arguments");
    BM_REQUIRE(n <= (int64_t(1) << 20), "at most 2^20 traces");
    if (n == 0) return;
    hipStream_t st;
    BM_HIP(hipStreamCreate(&st));
    struct StreamGuard {
      hipStream_t s;
      ~StreamGuard() { hipStreamDestroy(s); }
    } sg{st};
    DevBuf x, d;
    x.alloc((size_t)(n * T) * 8);
    d.alloc((size_t)(n * n) * 8);
    if (T > 0) BM_HIP(hipMemcpyAsync(x.p, X, (size_t)(n * T) * 8, hipMemcpyHostToDevice, st));
    const int64_t nb = (n + TILE - 1) / TILE;
    const int64_t blocks = nb * (nb + 1) / 2;
    BM_REQUIRE(blocks < (int64_t(1) << 31), "too many traces");
    hipLaunchKernelGGL(k_pair_sqdist, dim3((unsigned)blocks), dim3(MT), 0, st, x.as<double>(), n, T, scale,
                       d.as<double>());
    BM_HIP(hipGetLastError());
    BM_HIP(hipMemcpyAsync(D, d.p, (size_t)(n * n) * 8, hipMemcpyDeviceToHost, st));
    BM_HIP(hipStreamSynchronize(st));
  });
}

}  // extern "C"
