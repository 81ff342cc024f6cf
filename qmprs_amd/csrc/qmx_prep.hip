// Input side of prepare_state (qmprs/synthesis/mps_encoding/base.py:96-104):
// quick Ket re-indexing ("snake") and compression of the statevector, and
// the dense state of an MPS given as site tensors (MPS(mps=...),
// qmprs/primitives/mps.py:183-197).
//
// Compression zeroes the c smallest-magnitude amplitudes of each state
// exactly: a most-significant-byte-first radix select over the bit patterns
// of |a|^2 (IEEE doubles >= 0 order like their bit patterns) finds the c-th
// smallest key K, keys below K are zeroed, and of the keys equal to K the
// first ones in index order (a two-level count / scan over chunks of 4096).
#include "qmx_device.h"
#include "qmx_internal.h"

namespace qmx {

constexpr int kPrepChunk = 4096;   // amplitudes per block in the tie pass

QMX_DEV uint64_t mag2_key(float2 a) {
  const double x = a.x, y = a.y;
  return (uint64_t)__double_as_longlong(__dadd_rn(__dmul_rn(x, x), __dmul_rn(y, y)));
}
QMX_DEV uint64_t mag2_key(double2 a) {
  return (uint64_t)__double_as_longlong(__dadd_rn(__dmul_rn(a.x, a.x), __dmul_rn(a.y, a.y)));
}

// out = in re-indexed: identity, or "snake" (n >= 3): the second half reversed
template <typename S>
__global__ __launch_bounds__(256) void reindex_kernel(const S* in, S* out, int n, int snake) {
  const size_t len = (size_t)1 << n, half = len >> 1;
  const S* src = in + (size_t)blockIdx.y * len;
  S* dst = out + (size_t)blockIdx.y * len;
  const bool rev = snake && n >= 3;
  for (size_t i = (size_t)blockIdx.x * blockDim.x + threadIdx.x; i < len; i += (size_t)gridDim.x * blockDim.x)
    dst[i] = src[(rev && i >= half) ? 3 * half - 1 - i : i];
}

// histogram of key byte `shift` over the keys whose higher bytes equal the
// state's prefix so far
template <typename S>
__global__ __launch_bounds__(256) void radix_hist_kernel(const S* data, int n, int shift, const uint64_t* prefix,
                                                         uint32_t* hist) {
  __shared__ uint32_t h[256];
  h[threadIdx.x] = 0;
  __syncthreads();
  const int b = blockIdx.y;
  const size_t len = (size_t)1 << n;
  const S* d = data + (size_t)b * len;
  const uint64_t hi = shift >= 56 ? 0ull : (~0ull << (shift + 8));
  const uint64_t want = prefix[b] & hi;
  for (size_t i = (size_t)blockIdx.x * blockDim.x + threadIdx.x; i < len; i += (size_t)gridDim.x * blockDim.x) {
    const uint64_t k = mag2_key(d[i]);
    if ((k & hi) == want) atomicAdd(&h[(k >> shift) & 255u], 1u);
  }
  __syncthreads();
  const uint32_t c = h[threadIdx.x];
  if (c) atomicAdd(&hist[(size_t)b * 256 + threadIdx.x], c);
}

// pick the byte of the rank-`rank` key, narrow the prefix, clear the histogram
__global__ __launch_bounds__(256) void radix_pick_kernel(uint32_t* hist, uint64_t* prefix, uint32_t* rank,
                                                         uint32_t* below, int shift) {
  const int b = blockIdx.x;
  uint32_t* hb = hist + (size_t)b * 256;
  if (threadIdx.x == 0) {
    uint32_t cum = 0;
    for (int bin = 0; bin < 256; ++bin) {
      const uint32_t c = hb[bin];
      if (rank[b] < cum + c) {
        prefix[b] |= (uint64_t)bin << shift;
        rank[b] -= cum;
        below[b] += cum;
        break;
      }
      cum += c;
    }
  }
  __syncthreads();
  hb[threadIdx.x] = 0;
}

// number of keys equal to K per chunk
template <typename S>
__global__ __launch_bounds__(256) void tie_count_kernel(const S* data, int n, const uint64_t* kth, uint32_t* cnt,
                                                        int nchunks) {
  __shared__ uint32_t s;
  if (threadIdx.x == 0) s = 0;
  __syncthreads();
  const int b = blockIdx.y;
  const size_t len = (size_t)1 << n;
  const S* d = data + (size_t)b * len;
  const size_t c0 = (size_t)blockIdx.x * kPrepChunk;
  uint32_t mine = 0;
  for (size_t i = c0 + threadIdx.x; i < c0 + kPrepChunk && i < len; i += blockDim.x)
    mine += mag2_key(d[i]) == kth[b] ? 1u : 0u;
  if (mine) atomicAdd(&s, mine);
  __syncthreads();
  if (threadIdx.x == 0) cnt[(size_t)b * nchunks + blockIdx.x] = s;
}

// exclusive scan of the chunk counts (one block per state)
__global__ __launch_bounds__(1024) void tie_scan_kernel(uint32_t* cnt, int nchunks) {
  __shared__ uint32_t part[1024];
  uint32_t* c = cnt + (size_t)blockIdx.x * nchunks;
  const int per = (nchunks + 1023) / 1024;
  const int lo = threadIdx.x * per, hi = min(nchunks, lo + per);
  uint32_t sum = 0;
  for (int i = lo; i < hi; ++i) sum += c[i];
  part[threadIdx.x] = sum;
  __syncthreads();
  for (int off = 1; off < 1024; off <<= 1) {
    const uint32_t v = threadIdx.x >= (unsigned)off ? part[threadIdx.x - off] : 0u;
    __syncthreads();
    part[threadIdx.x] += v;
    __syncthreads();
  }
  uint32_t run = part[threadIdx.x] - sum;
  for (int i = lo; i < hi; ++i) {
    const uint32_t v = c[i];
    c[i] = run;
    run += v;
  }
}

// zero keys below K, and keys equal to K whose index-order rank is < take
// (each thread owns 16 consecutive amplitudes of the chunk)
template <typename S>
__global__ __launch_bounds__(256) void compress_zero_kernel(S* data, int n, const uint64_t* kth,
                                                            const uint32_t* rank, const uint32_t* offs,
                                                            int nchunks) {
  __shared__ uint32_t part[256];
  const int b = blockIdx.y;
  const size_t len = (size_t)1 << n;
  S* d = data + (size_t)b * len;
  const uint64_t K = kth[b];
  const uint32_t take = rank[b] + 1u;   // equal keys to zero
  const size_t i0 = (size_t)blockIdx.x * kPrepChunk + (size_t)threadIdx.x * 16;
  uint64_t k[16];
  uint32_t eq = 0;
#pragma unroll
  for (int j = 0; j < 16; ++j) {
    k[j] = (i0 + j < len) ? mag2_key(d[i0 + j]) : ~0ull;
    eq += k[j] == K ? 1u : 0u;
  }
  part[threadIdx.x] = eq;
  __syncthreads();
  for (int off = 1; off < 256; off <<= 1) {
    const uint32_t v = threadIdx.x >= (unsigned)off ? part[threadIdx.x - off] : 0u;
    __syncthreads();
    part[threadIdx.x] += v;
    __syncthreads();
  }
  uint32_t r = offs[(size_t)b * nchunks + blockIdx.x] + part[threadIdx.x] - eq;
  S zero;
  zero.x = 0;
  zero.y = 0;
#pragma unroll
  for (int j = 0; j < 16; ++j) {
    if (i0 + j >= len) break;
    if (k[j] < K) {
      d[i0 + j] = zero;
    } else if (k[j] == K) {
      if (r < take) d[i0 + j] = zero;
      ++r;
    }
  }
}

size_t prep_input_bytes(int n, int batch) {
  const size_t nchunks = std::max<size_t>(1, ((size_t)1 << n) / kPrepChunk);
  return (size_t)batch * (256 * sizeof(uint32_t) + 3 * sizeof(uint64_t) + nchunks * sizeof(uint32_t)) + 1024;
}

template <typename S>
void launch_prep_input(const S* in, S* out, int n, int batch, int snake, size_t cutoff, void* ws,
                       hipStream_t st) {
  const size_t len = (size_t)1 << n;
  const unsigned nb = (unsigned)std::min<size_t>(4096, std::max<size_t>(1, len / 256 / 4));
  hipLaunchKernelGGL(reindex_kernel<S>, dim3(nb, batch), dim3(256), 0, st, in, out, n, snake);
  if (cutoff == 0) return;
  const int nchunks = (int)std::max<size_t>(1, len / kPrepChunk);
  char* p = static_cast<char*>(ws);
  uint32_t* hist = reinterpret_cast<uint32_t*>(p);
  p += (size_t)batch * 256 * sizeof(uint32_t);
  uint64_t* prefix = reinterpret_cast<uint64_t*>(p);
  p += (size_t)batch * sizeof(uint64_t);
  uint32_t* rank = reinterpret_cast<uint32_t*>(p);
  p += (size_t)batch * sizeof(uint64_t);
  uint32_t* below = reinterpret_cast<uint32_t*>(p);
  p += (size_t)batch * sizeof(uint64_t);
  uint32_t* cnt = reinterpret_cast<uint32_t*>(p);
  // prefix = 0, rank = cutoff - 1, below = 0, histograms cleared
  launch_fill_int(reinterpret_cast<int*>(hist), 0, batch * 256, st);
  launch_fill_int(reinterpret_cast<int*>(prefix), 0, batch * 2, st);
  launch_fill_int(reinterpret_cast<int*>(rank), (int)(cutoff - 1), batch, st);
  launch_fill_int(reinterpret_cast<int*>(below), 0, batch, st);
  for (int shift = 56; shift >= 0; shift -= 8) {
    hipLaunchKernelGGL(radix_hist_kernel<S>, dim3(nb, batch), dim3(256), 0, st, out, n, shift, prefix, hist);
    hipLaunchKernelGGL(radix_pick_kernel, dim3(batch), dim3(256), 0, st, hist, prefix, rank, below, shift);
  }
  hipLaunchKernelGGL(tie_count_kernel<S>, dim3(nchunks, batch), dim3(256), 0, st, out, n, prefix, cnt, nchunks);
  hipLaunchKernelGGL(tie_scan_kernel, dim3(batch), dim3(1024), 0, st, cnt, nchunks);
  hipLaunchKernelGGL(compress_zero_kernel<S>, dim3(nchunks, batch), dim3(256), 0, st, out, n, prefix, rank, cnt,
                     nchunks);
}
template void launch_prep_input<float2>(const float2*, float2*, int, int, int, size_t, void*, hipStream_t);
template void launch_prep_input<double2>(const double2*, double2*, int, int, int, size_t, void*, hipStream_t);

}  // namespace qmx
