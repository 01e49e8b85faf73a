// On-device construction of block-diagonal random connectivity (the role of
// connect_random / random_matrix, brian/connections/construction.py:42-177,
// for many independent networks at once).  Pair (i, j) is connected iff
// pair_u32(seed, i, j) < floor(p * 2^32); its delay, when requested, is
// dmin + (pair_u32(seed', i, j) + 0.5) * 2^-32 * (dmax - dmin) seconds,
// quantised as int(invdt * d) (delayconnection.py:123, :152).  The hash is
// integer-only so a host restatement reproduces the matrix exactly.
// Two passes over all candidate pairs: row counts, exclusive scan, fill.
#include "net_launch.hpp"

#include <hipcub/hipcub.hpp>

namespace bm {

__host__ __device__ inline uint32_t fmix32(uint32_t h) {
  h ^= h >> 16;
  h *= 0x85EBCA6Bu;
  h ^= h >> 13;
  h *= 0xC2B2AE35u;
  h ^= h >> 16;
  return h;
}

__host__ __device__ inline uint32_t pair_u32(uint64_t seed, uint32_t i, uint32_t j) {
  const uint32_t key = fmix32((i * 0x9E3779B1u) ^ (uint32_t)seed);
  return fmix32((key + j * 0x85EBCA77u) ^ (uint32_t)(seed >> 32));
}

struct BlockDesc {
  const int64_t *src_off, *src_n, *tgt_off, *tgt_n;  // [nblocks], sorted by src_off
  int nblocks;
  int64_t src_lo;        // first source row of the connection (row r = src_lo + r)
  int64_t tgt_lo;        // target range origin (columns are relative to it)
  int64_t nrows;
  uint32_t thr;
  uint64_t seed, dseed;
  double dmin, dmax, invdt;
  int with_delay, dly_bytes;
};

__device__ inline int find_block(const BlockDesc &b, int64_t i) {
  int lo = 0, hi = b.nblocks - 1, res = -1;
  while (lo <= hi) {
    const int mid = (lo + hi) >> 1;
    if (b.src_off[mid] <= i) {
      res = mid;
      lo = mid + 1;
    } else {
      hi = mid - 1;
    }
  }
  if (res >= 0 && i < b.src_off[res] + b.src_n[res]) return res;
  return -1;
}

// one wavefront per row
__global__ __launch_bounds__(256) void k_gen_count(const BlockDesc b, int64_t *counts) {
  const int64_t r = ((int64_t)blockIdx.x * 256 + threadIdx.x) >> 6;
  const int lane = threadIdx.x & 63;
  if (r >= b.nrows) return;
  const int64_t i = b.src_lo + r;
  const int blk = find_block(b, i);
  int cnt = 0;
  if (blk >= 0) {
    const int64_t t0 = b.tgt_off[blk], tn = b.tgt_n[blk];
    for (int64_t q = lane; q < tn; q += 64) cnt += pair_u32(b.seed, (uint32_t)i, (uint32_t)(t0 + q)) < b.thr;
  }
#pragma unroll
  for (int off = 32; off > 0; off >>= 1) cnt += __shfl_down(cnt, off);
  if (lane == 0) counts[r] = cnt;
}

__global__ __launch_bounds__(256) void k_gen_fill(const BlockDesc b, const int64_t *row_ptr, int32_t *col,
                                                  void *dly) {
  const int64_t r = ((int64_t)blockIdx.x * 256 + threadIdx.x) >> 6;
  const int lane = threadIdx.x & 63;
  if (r >= b.nrows) return;
  const int64_t i = b.src_lo + r;
  const int blk = find_block(b, i);
  if (blk < 0) return;
  const int64_t t0 = b.tgt_off[blk], tn = b.tgt_n[blk];
  int64_t pos = row_ptr[r];
  for (int64_t q0 = 0; q0 < tn; q0 += 64) {
    const int64_t q = q0 + lane;
    const bool hit = q < tn && pair_u32(b.seed, (uint32_t)i, (uint32_t)(t0 + q)) < b.thr;
    const unsigned long long mask = __ballot(hit);
    if (hit) {
      const int below = __popcll(mask & ((1ull << lane) - 1ull));
      const int64_t p = pos + below;
      col[p] = (int32_t)(t0 + q - b.tgt_lo);
      if (b.with_delay) {
        // no fused multiply-add: the host restatement rounds each operation
#pragma clang fp contract(off)
        const uint32_t u = pair_u32(b.dseed, (uint32_t)i, (uint32_t)(t0 + q));
        const double d = b.dmin + ((double)u + 0.5) * 2.3283064365386963e-10 * (b.dmax - b.dmin);
        const int st = (int)(b.invdt * d);
        if (b.dly_bytes == 1) static_cast<uint8_t *>(dly)[p] = (uint8_t)st;
        else static_cast<uint16_t *>(dly)[p] = (uint16_t)st;
      }
    }
    pos += __popcll(mask);
  }
}

// Builds the CSR of a block-diagonal random connection on the device.
// Returns nnz; row_ptr (nrows+1, int64), col (int32) and optional per-synapse
// delays (u8/u16) are allocated into the given buffers.
int64_t generate_block_csr(const GenSpec &s, DevBuf &row_ptr, DevBuf &col, DevBuf &dly, hipStream_t st) {
  DevBuf d_so, d_sn, d_to, d_tn, counts, temp;
  const size_t nb = s.src_off.size();
  d_so.alloc(nb * 8);
  d_sn.alloc(nb * 8);
  d_to.alloc(nb * 8);
  d_tn.alloc(nb * 8);
  BM_HIP(hipMemcpyAsync(d_so.p, s.src_off.data(), nb * 8, hipMemcpyHostToDevice, st));
  BM_HIP(hipMemcpyAsync(d_sn.p, s.src_n.data(), nb * 8, hipMemcpyHostToDevice, st));
  BM_HIP(hipMemcpyAsync(d_to.p, s.tgt_off.data(), nb * 8, hipMemcpyHostToDevice, st));
  BM_HIP(hipMemcpyAsync(d_tn.p, s.tgt_n.data(), nb * 8, hipMemcpyHostToDevice, st));
  BlockDesc b{};
  b.src_off = d_so.as<int64_t>();
  b.src_n = d_sn.as<int64_t>();
  b.tgt_off = d_to.as<int64_t>();
  b.tgt_n = d_tn.as<int64_t>();
  b.nblocks = (int)nb;
  b.src_lo = s.src_lo;
  b.tgt_lo = s.tgt_lo;
  b.nrows = s.nrows;
  double t = std::floor(s.p * 4294967296.0);
  t = std::min(std::max(t, 0.0), 4294967295.0);
  b.thr = (uint32_t)t;
  b.seed = s.seed;
  b.dseed = s.seed ^ 0xA5A5A5A55A5A5A5Aull;
  b.dmin = s.dmin;
  b.dmax = s.dmax;
  b.invdt = s.invdt;
  b.with_delay = s.with_delay;
  b.dly_bytes = s.dly_bytes;
  // int64 counts so the scan accumulates in 64 bits (nnz may exceed 2^31)
  counts.alloc((size_t)(s.nrows + 1) * 8);
  BM_HIP(hipMemsetAsync(counts.p, 0, (size_t)(s.nrows + 1) * 8, st));
  row_ptr.alloc((size_t)(s.nrows + 1) * 8);
  const unsigned grid = (unsigned)((s.nrows * 64 + 255) / 256);
  if (s.nrows > 0) hipLaunchKernelGGL(k_gen_count, dim3(grid), dim3(256), 0, st, b, counts.as<int64_t>());
  BM_HIP(hipGetLastError());
  size_t tb = 0;
  BM_HIP(hipcub::DeviceScan::ExclusiveSum(nullptr, tb, counts.as<int64_t>(), row_ptr.as<int64_t>(),
                                          (int)(s.nrows + 1), st));
  temp.alloc(tb);
  BM_HIP(hipcub::DeviceScan::ExclusiveSum(temp.p, tb, counts.as<int64_t>(), row_ptr.as<int64_t>(),
                                          (int)(s.nrows + 1), st));
  int64_t nnz = 0;
  BM_HIP(hipMemcpyAsync(&nnz, row_ptr.as<int64_t>() + s.nrows, 8, hipMemcpyDeviceToHost, st));
  BM_HIP(hipStreamSynchronize(st));
  col.alloc((size_t)std::max<int64_t>(nnz, 1) * 4);
  if (s.with_delay) dly.alloc((size_t)std::max<int64_t>(nnz, 1) * s.dly_bytes);
  if (s.nrows > 0)
    hipLaunchKernelGGL(k_gen_fill, dim3(grid), dim3(256), 0, st, b, row_ptr.as<int64_t>(), col.as<int32_t>(),
                       s.with_delay ? dly.p : nullptr);
  BM_HIP(hipGetLastError());
  BM_HIP(hipStreamSynchronize(st));
  return nnz;
}

// ------------------------------------------------------------ fixed in-degree
// random_matrix_fixed_column (construction.py:149-177): every target column
// j gets k distinct sources drawn uniformly from [0, n).  Floyd's algorithm
// per column: for t = 0..k-1 with m = n-k+t, draw r uniform in [0, m]
// (r = (pair_u32(seed, j, t) * (m+1)) >> 32); take r, or m if r was already
// taken.  One thread per column; the column's set lives in pre[j*k ..].
__global__ __launch_bounds__(256) void k_fixed_sample(int64_t ncols, int k, uint32_t n, uint64_t seed,
                                                      int32_t *pre, int64_t *counts) {
  const int64_t j = (int64_t)blockIdx.x * 256 + threadIdx.x;
  if (j >= ncols) return;
  int32_t *S = pre + j * k;
  for (int t = 0; t < k; ++t) {
    const uint32_t m = n - (uint32_t)k + (uint32_t)t;
    const uint32_t u = pair_u32(seed, (uint32_t)j, (uint32_t)t);
    uint32_t r = (uint32_t)(((uint64_t)u * ((uint64_t)m + 1ull)) >> 32);
    bool seen = false;
    for (int q = 0; q < t; ++q) seen = seen || (S[q] == (int32_t)r);
    if (seen) r = m;
    S[t] = (int32_t)r;
    atomicAdd(reinterpret_cast<unsigned long long *>(counts) + r, 1ull);
  }
}

// Scatter (source, column) pairs into the source rows; the order inside a
// row is fixed afterwards by a segmented sort.
__global__ __launch_bounds__(256) void k_fixed_scatter(int64_t ncols, int k, const int32_t *pre,
                                                       int64_t *cursor, int32_t *col) {
  const int64_t e = (int64_t)blockIdx.x * 256 + threadIdx.x;
  if (e >= ncols * (int64_t)k) return;
  const int64_t j = e / k;
  const int32_t src = pre[e];
  const unsigned long long pos = atomicAdd(reinterpret_cast<unsigned long long *>(cursor) + src, 1ull);
  col[pos] = (int32_t)j;
}

int64_t generate_fixed_indegree_all(const GenSpec &s, DevBuf &row_ptr, DevBuf &col, hipStream_t st);

__global__ __launch_bounds__(256) void k_rebase_rows(const int64_t *in, int64_t n, int64_t *out) {
  const int64_t r = (int64_t)blockIdx.x * 256 + threadIdx.x;
  if (r <= n) out[r] = in[r] - in[0];
}

// Source-row CSR (columns ascending within each row) of a fixed in-degree
// connection: nrows sampled sources, ncols targets, k sources per target;
// rows [keep_off, keep_off + keep_n) of the sampled range are returned (two
// connections with the same seed split one sampling by source range).
// Returns nnz of the kept rows.
int64_t generate_fixed_indegree_csr(const GenSpec &s, DevBuf &row_ptr, DevBuf &col, hipStream_t st) {
  DevBuf full_rp, full_col;
  const int64_t nnz_all = generate_fixed_indegree_all(s, full_rp, full_col, st);
  (void)nnz_all;
  int64_t lo = 0, hi = 0;
  BM_HIP(hipMemcpyAsync(&lo, full_rp.as<int64_t>() + s.keep_off, 8, hipMemcpyDeviceToHost, st));
  BM_HIP(hipMemcpyAsync(&hi, full_rp.as<int64_t>() + s.keep_off + s.keep_n, 8, hipMemcpyDeviceToHost, st));
  BM_HIP(hipStreamSynchronize(st));
  row_ptr.alloc((size_t)(s.keep_n + 1) * 8);
  hipLaunchKernelGGL(k_rebase_rows, dim3((unsigned)((s.keep_n + 1 + 255) / 256)), dim3(256), 0, st,
                     full_rp.as<int64_t>() + s.keep_off, s.keep_n, row_ptr.as<int64_t>());
  BM_HIP(hipGetLastError());
  col.alloc((size_t)std::max<int64_t>(hi - lo, 1) * 4);
  if (hi > lo)
    BM_HIP(hipMemcpyAsync(col.p, full_col.as<int32_t>() + lo, (size_t)(hi - lo) * 4, hipMemcpyDeviceToDevice, st));
  BM_HIP(hipStreamSynchronize(st));
  return hi - lo;
}

int64_t generate_fixed_indegree_all(const GenSpec &s, DevBuf &row_ptr, DevBuf &col, hipStream_t st) {
  const int64_t nnz = s.ncols * (int64_t)s.k;
  BM_REQUIRE(nnz < (int64_t)INT32_MAX, "fixed in-degree connection too large for one sort");
  DevBuf pre, counts, cursor, tmp_col, temp;
  pre.alloc((size_t)std::max<int64_t>(nnz, 1) * 4);
  counts.alloc((size_t)(s.nrows + 1) * 8);
  BM_HIP(hipMemsetAsync(counts.p, 0, (size_t)(s.nrows + 1) * 8, st));
  if (s.ncols > 0)
    hipLaunchKernelGGL(k_fixed_sample, dim3((unsigned)((s.ncols + 255) / 256)), dim3(256), 0, st, s.ncols, s.k,
                       (uint32_t)s.nrows, s.seed, pre.as<int32_t>(), counts.as<int64_t>());
  BM_HIP(hipGetLastError());
  row_ptr.alloc((size_t)(s.nrows + 1) * 8);
  size_t tb = 0;
  BM_HIP(hipcub::DeviceScan::ExclusiveSum(nullptr, tb, counts.as<int64_t>(), row_ptr.as<int64_t>(),
                                          (int)(s.nrows + 1), st));
  temp.alloc(tb);
  BM_HIP(hipcub::DeviceScan::ExclusiveSum(temp.p, tb, counts.as<int64_t>(), row_ptr.as<int64_t>(),
                                          (int)(s.nrows + 1), st));
  cursor.alloc((size_t)(s.nrows + 1) * 8);
  BM_HIP(hipMemcpyAsync(cursor.p, row_ptr.p, (size_t)(s.nrows + 1) * 8, hipMemcpyDeviceToDevice, st));
  tmp_col.alloc((size_t)std::max<int64_t>(nnz, 1) * 4);
  if (nnz > 0)
    hipLaunchKernelGGL(k_fixed_scatter, dim3((unsigned)((nnz + 255) / 256)), dim3(256), 0, st, s.ncols, s.k,
                       pre.as<int32_t>(), cursor.as<int64_t>(), tmp_col.as<int32_t>());
  BM_HIP(hipGetLastError());
  col.alloc((size_t)std::max<int64_t>(nnz, 1) * 4);
  const int64_t *rp = row_ptr.as<int64_t>();
  size_t tb2 = 0;
  BM_HIP(hipcub::DeviceSegmentedRadixSort::SortKeys(nullptr, tb2, tmp_col.as<int32_t>(), col.as<int32_t>(),
                                                    (int)nnz, (int)s.nrows, rp, rp + 1, 0, 32, st));
  DevBuf temp2;
  temp2.alloc(tb2);
  BM_HIP(hipcub::DeviceSegmentedRadixSort::SortKeys(temp2.p, tb2, tmp_col.as<int32_t>(), col.as<int32_t>(),
                                                    (int)nnz, (int)s.nrows, rp, rp + 1, 0, 32, st));
  BM_HIP(hipStreamSynchronize(st));
  return nnz;
}

}  // namespace bm
