// FP64 linear algebra of the chi-truncation chain (statevector -> chi-MPS).
//
// Step i of the right-to-left truncated projection chain works on
// Y_i = X_{i+1} viewed as a (2^i x n_i) row-major complex128 matrix:
//   rho_i = Y_i^H Y_i             gram + gram_reduce   (n_i x n_i, ld = nmax)
//   rho_i = V diag(lam) V^H       jacobi               (one block per state)
//   W_i   = V[:, top k_i]         select               (rsum2 cutoff, max bond chi)
//   X_i   = Y_i W_i               gemm
// and the state is contracted back with Z_{i+1} = Z_i W_i^H (gemm, conj_b).
// Shapes that depend on the cutoff live on the device (k_i, n_i per state);
// launches are sized by host upper bounds and idle blocks exit early.
#include <algorithm>

#include "qmx_device.h"
#include "qmx_internal.h"

namespace qmx {

// ------------------------------------------------------------------ gram ----
// part[b][chunk] (nmax x nmax) = sum over the chunk's rows of conj(Y)[:, p] Y[:, q]
__global__ __launch_bounds__(256) void gram_kernel(const double2* __restrict__ y, size_t y_stride,
                                                   const int* ncols_dev, int rows_log2, int nmax,
                                                   double2* part, int nchunks) {
  __shared__ double2 ya[32][33];
  __shared__ double2 yb[32][33];
  const int b = blockIdx.z;
  const int nc = ncols_dev[b];
  const int ntile = (nmax + 31) / 32;
  const int ti = blockIdx.x / ntile, tj = blockIdx.x % ntile;
  if (ti * 32 >= nc || tj * 32 >= nc) return;
  const size_t rows = (size_t)1 << rows_log2;
  const size_t per = (rows + nchunks - 1) / nchunks;
  const size_t r0 = per * blockIdx.y;
  const size_t r1 = (r0 + per < rows) ? r0 + per : rows;
  const double2* yb0 = y + (size_t)b * y_stride;
  const int tx = threadIdx.x & 31, ty = threadIdx.x >> 5;
  double2 acc[4] = {cz(), cz(), cz(), cz()};
  for (size_t rb = r0; rb < r1; rb += 32) {
    for (int e = threadIdx.x; e < 32 * 32; e += 256) {
      int rr = e >> 5, cc = e & 31;
      size_t row = rb + rr;
      int ca = ti * 32 + cc, cb = tj * 32 + cc;
      ya[rr][cc] = (row < r1 && ca < nc) ? yb0[row * nc + ca] : cz();
      yb[rr][cc] = (row < r1 && cb < nc) ? yb0[row * nc + cb] : cz();
    }
    __syncthreads();
#pragma unroll 4
    for (int rr = 0; rr < 32; ++rr) {
      double2 q = yb[rr][tx];
#pragma unroll
      for (int j = 0; j < 4; ++j) acc[j] = cfmac(ya[rr][ty * 4 + j], q, acc[j]);
    }
    __syncthreads();
  }
  double2* out = part + ((size_t)b * nchunks + blockIdx.y) * nmax * nmax;
  for (int j = 0; j < 4; ++j) {
    int p = ti * 32 + ty * 4 + j, q = tj * 32 + tx;
    if (p < nc && q < nc) out[(size_t)p * nmax + q] = acc[j];
  }
}

__global__ void gram_reduce_kernel(const double2* part, int nchunks, const int* ncols_dev, int nmax,
                                   double2* a, size_t a_stride) {
  const int b = blockIdx.y;
  const int nc = ncols_dev[b];
  int e = blockIdx.x * blockDim.x + threadIdx.x;
  if (e >= nc * nc) return;
  int p = e / nc, q = e % nc;
  double2 s = cz();
  for (int c = 0; c < nchunks; ++c)
    s = cadd(s, part[((size_t)b * nchunks + c) * nmax * nmax + (size_t)p * nmax + q]);
  a[(size_t)b * a_stride + (size_t)p * nmax + q] = s;
}

// ---------------------------------------------------------------- jacobi ----
// Parallel cyclic two-sided Jacobi for a Hermitian n x n matrix (n even,
// n <= 2 * 256), one block per state.  Round r pairs the indices by the circle
// method; all n/2 rotations of a round are disjoint, so the update
// A <- J^H A J is done blockwise: each 2x2 block (P, Q) of A becomes
// J_P^H A_PQ J_Q, independently of every other block.  V accumulates J.
constexpr int kJacobiMaxPairs = 512;

// A pair is rotated only if its off-diagonal entry exceeds kJacobiAbsTol times
// the trace (and 1e-15 of the geometric mean of its diagonal entries).  The
// Gram matrix carries rounding errors of order eps * trace from its own
// formation, and every rotation that involves a large eigenvalue leaves
// errors of that order in the entries it touches; with a floor well below
// that level the rotations inside a cluster of numerically zero eigenvalues
// (rank-deficient Gram matrices: fewer rows than columns) regenerate each
// other's entries and a sweep may never come out clean.  A quarter of eps
// ends those sweeps and leaves the kept subspace as accurate as the Gram
// matrix it comes from.
constexpr double kJacobiAbsTol = 0.25 * 2.220446049250313e-16;

__global__ __launch_bounds__(512) void jacobi_kernel(double2* a_all, double2* v_all, size_t stride,
                                                     const int* ncols_dev, int ld, int max_sweeps,
                                                     int* error) {
  __shared__ int pp[kJacobiMaxPairs], qq[kJacobiMaxPairs];
  __shared__ double jc[kJacobiMaxPairs], js[kJacobiMaxPairs];
  __shared__ double2 je[kJacobiMaxPairs];
  __shared__ int any_rot;
  __shared__ double s_scale;
  const int b = blockIdx.x;
  const int n = ncols_dev[b];
  double2* a = a_all + (size_t)b * stride;
  double2* v = v_all + (size_t)b * stride;
  const int tid = threadIdx.x, nt = blockDim.x;
  for (int e = tid; e < n * n; e += nt) {
    int i = e / n, j = e % n;
    v[(size_t)i * ld + j] = make_double2(i == j ? 1.0 : 0.0, 0.0);
  }
  if (tid == 0) {
    double tr = 0.0;
    for (int i = 0; i < n; ++i) tr += fabs(a[(size_t)i * ld + i].x);
    s_scale = tr;
  }
  __syncthreads();
  if (n < 2) return;
  const int np = n / 2;
  const double abs_tol = kJacobiAbsTol * s_scale + 1e-300;
  bool converged = false;
  for (int sweep = 0; sweep < max_sweeps; ++sweep) {
    __syncthreads();
    if (tid == 0) any_rot = 0;
    __syncthreads();
    for (int r = 0; r < n - 1; ++r) {
      if (tid < np) {
        int m = tid;
        int p = (m == 0) ? n - 1 : (r + m) % (n - 1);
        int q = (r + n - 1 - m) % (n - 1);
        if (m == 0) q = r % (n - 1);
        if (p > q) { int t = p; p = q; q = t; }
        double al = a[(size_t)p * ld + p].x, be = a[(size_t)q * ld + q].x;
        double2 g = a[(size_t)p * ld + q];
        double ag = cabs_(g);
        double c = 1.0, s = 0.0;
        double2 em = make_double2(1.0, 0.0);
        if (ag > abs_tol && ag > 1e-15 * sqrt(fabs(al * be))) {
          double tau = (be - al) / (2.0 * ag);
          double t = (tau >= 0 ? 1.0 : -1.0) / (fabs(tau) + sqrt(1.0 + tau * tau));
          c = 1.0 / sqrt(1.0 + t * t);
          s = t * c;
          em = make_double2(g.x / ag, -g.y / ag);   // e^{-i phi}
          any_rot = 1;
        }
        pp[m] = p; qq[m] = q; jc[m] = c; js[m] = s; je[m] = em;
      }
      __syncthreads();
      for (int e = tid; e < np * np; e += nt) {
        int P = e / np, Q = e % np;
        int p0 = pp[P], p1 = qq[P], q0 = pp[Q], q1 = qq[Q];
        // J = [[c, s], [-s em, c em]]
        double2 J00 = make_double2(jc[Q], 0.0), J01 = make_double2(js[Q], 0.0);
        double2 J10 = cscale(je[Q], -js[Q]), J11 = cscale(je[Q], jc[Q]);
        double2 K00 = make_double2(jc[P], 0.0), K01 = make_double2(js[P], 0.0);
        double2 K10 = cscale(je[P], -js[P]), K11 = cscale(je[P], jc[P]);
        double2 a00 = a[(size_t)p0 * ld + q0], a01 = a[(size_t)p0 * ld + q1];
        double2 a10 = a[(size_t)p1 * ld + q0], a11 = a[(size_t)p1 * ld + q1];
        double2 b00 = cadd(cmul(a00, J00), cmul(a01, J10));
        double2 b01 = cadd(cmul(a00, J01), cmul(a01, J11));
        double2 b10 = cadd(cmul(a10, J00), cmul(a11, J10));
        double2 b11 = cadd(cmul(a10, J01), cmul(a11, J11));
        double2 c00 = cadd(cmulc(K00, b00), cmulc(K10, b10));
        double2 c01 = cadd(cmulc(K00, b01), cmulc(K10, b11));
        double2 c10 = cadd(cmulc(K01, b00), cmulc(K11, b10));
        double2 c11 = cadd(cmulc(K01, b01), cmulc(K11, b11));
        if (P == Q) {
          c00.y = 0.0; c11.y = 0.0;
          c01 = cz(); c10 = cz();
        }
        a[(size_t)p0 * ld + q0] = c00; a[(size_t)p0 * ld + q1] = c01;
        a[(size_t)p1 * ld + q0] = c10; a[(size_t)p1 * ld + q1] = c11;
      }
      for (int e = tid; e < n * np; e += nt) {
        int i = e / np, Q = e % np;
        int q0 = pp[Q], q1 = qq[Q];
        double2 J00 = make_double2(jc[Q], 0.0), J01 = make_double2(js[Q], 0.0);
        double2 J10 = cscale(je[Q], -js[Q]), J11 = cscale(je[Q], jc[Q]);
        double2 x = v[(size_t)i * ld + q0], y = v[(size_t)i * ld + q1];
        v[(size_t)i * ld + q0] = cadd(cmul(x, J00), cmul(y, J10));
        v[(size_t)i * ld + q1] = cadd(cmul(x, J01), cmul(y, J11));
      }
      __syncthreads();
    }
    __syncthreads();
    if (!any_rot) { converged = true; break; }
  }
  if (!converged && tid == 0 && error) atomicOr(error, kErrEigNoConv);
}

// ------------------------------------------------ LDS-resident Jacobi -------
// For n <= kJacobiLdsMax the upper triangle of A (n (n + 1) / 2 complex128,
// 132 KiB at n = 128) lives in LDS for a whole sweep: one block per state,
// the same parallel (circle) ordering and rotations as jacobi_kernel.  Each
// sweep logs its rotations (c, s, e^{-i phi} per pair and round); a second
// kernel applies the log to the rows of V, 16 rows per block held in LDS.
// The host launches sweep / update pairs; a state none of whose off-diagonal
// entries passes the rotation test at the start of a sweep is converged (that
// sweep would rotate nothing and is not run), as is one whose sweep rotated
// nothing, and every later launch returns at once for it.
constexpr int kJacobiLdsMax = 128;
constexpr int kJacobiRowsPerBlock = 16;

QMX_DEV int tri_index(int i, int j, int n) { return i * (2 * n - i + 1) / 2 + (j - i); }

__global__ __launch_bounds__(1024) void jacobi_lds_sweep_kernel(double2* a_all, size_t stride,
                                                                const int* ncols_dev, int ld,
                                                                double4* log_all, size_t log_stride,
                                                                int* state, int sweep) {
  extern __shared__ double2 tri[];
  __shared__ int pp[kJacobiLdsMax / 2], qq[kJacobiLdsMax / 2];
  __shared__ double jc[kJacobiLdsMax / 2], js[kJacobiLdsMax / 2];
  __shared__ double2 je[kJacobiLdsMax / 2];
  __shared__ int any_rot;
  __shared__ double s_scale;
  const int b = blockIdx.x;
  // state[b]: 0 = running, 1 = converged; state[batch + b] = sweeps run
  if (state[b]) return;
  const int n = ncols_dev[b];
  if (n < 2) {
    if (threadIdx.x == 0) state[b] = 1;
    return;
  }
  double2* a = a_all + (size_t)b * stride;
  double4* lg = log_all + (size_t)b * log_stride;
  const int tid = threadIdx.x, nt = blockDim.x;
  const int ntri = n * (n + 1) / 2;
  for (int e = tid; e < ntri; e += nt) {
    // row i: the largest i with tri_index(i, i) <= e
    int i = (int)((2.0 * n + 1.0 - sqrt((2.0 * n + 1.0) * (2.0 * n + 1.0) - 8.0 * e)) * 0.5);
    while (i > 0 && tri_index(i, i, n) > e) --i;
    while (i + 1 < n && tri_index(i + 1, i + 1, n) <= e) ++i;
    const int j = i + (e - tri_index(i, i, n));
    tri[e] = a[(size_t)i * ld + j];
  }
  if (tid == 0) {
    any_rot = 0;
    double tr = 0.0;
    for (int i = 0; i < n; ++i) tr += fabs(a[(size_t)i * ld + i].x);
    s_scale = tr;
  }
  __syncthreads();
  const int np = n / 2;
  const double abs_tol = kJacobiAbsTol * s_scale + 1e-300;
  // a sweep would rotate nothing iff no off-diagonal entry passes the rotation
  // test now: then the state is converged and the sweep is not run
  for (int e = tid; e < ntri; e += nt) {
    int i = (int)((2.0 * n + 1.0 - sqrt((2.0 * n + 1.0) * (2.0 * n + 1.0) - 8.0 * e)) * 0.5);
    while (i > 0 && tri_index(i, i, n) > e) --i;
    while (i + 1 < n && tri_index(i + 1, i + 1, n) <= e) ++i;
    const int j = i + (e - tri_index(i, i, n));
    if (j == i) continue;
    const double ag = cabs_(tri[e]);
    if (ag > abs_tol && ag > 1e-15 * sqrt(fabs(tri[tri_index(i, i, n)].x * tri[tri_index(j, j, n)].x)))
      any_rot = 1;
  }
  __syncthreads();
  const bool go = any_rot;
  __syncthreads();
  if (!go) {
    if (tid == 0) state[b] = 1;
    return;
  }
  if (tid == 0) any_rot = 0;
  __syncthreads();
  // this thread's pair blocks (P, Q), P <= Q, row-major over the upper
  // triangle of pair blocks; the same in every round (the kernel is
  // launched with 1024 threads: kBlkPer * 1024 covers every block)
  constexpr int kBlkPer = (kJacobiLdsMax / 2 * (kJacobiLdsMax / 2 + 1) / 2 + 1023) / 1024;
  const int nblocks = np * (np + 1) / 2;
  int bP[kBlkPer], bQ[kBlkPer];
#pragma unroll
  for (int k = 0; k < kBlkPer; ++k) {
    const int e = tid + k * nt;
    int P = 0, Q = 0;
    if (e < nblocks) {
      P = (int)((2.0 * np + 1.0 - sqrt((2.0 * np + 1.0) * (2.0 * np + 1.0) - 8.0 * e)) * 0.5);
      while (P > 0 && tri_index(P, P, np) > e) --P;
      while (P + 1 < np && tri_index(P + 1, P + 1, np) <= e) ++P;
      Q = P + (e - tri_index(P, P, np));
    }
    bP[k] = P;
    bQ[k] = Q;
  }
  for (int r = 0; r < n - 1; ++r) {
    if (tid < np) {
      const int m = tid;
      int p = (m == 0) ? n - 1 : (r + m) % (n - 1);
      int q = (r + n - 1 - m) % (n - 1);
      if (m == 0) q = r % (n - 1);
      if (p > q) { int t = p; p = q; q = t; }
      const double al = tri[tri_index(p, p, n)].x, be = tri[tri_index(q, q, n)].x;
      const double2 g = tri[tri_index(p, q, n)];
      const double ag = cabs_(g);
      double c = 1.0, sn = 0.0;
      double2 em = make_double2(1.0, 0.0);
      // one comparison against the larger threshold: the diagonal reads are
      // issued together with the off-diagonal one
      if (ag > fmax(abs_tol, 1e-15 * sqrt(fabs(al * be)))) {
        const double tau = (be - al) / (2.0 * ag);
        const double t = (tau >= 0 ? 1.0 : -1.0) / (fabs(tau) + sqrt(1.0 + tau * tau));
        c = 1.0 / sqrt(1.0 + t * t);
        sn = t * c;
        em = make_double2(g.x / ag, -g.y / ag);   // e^{-i phi}
        any_rot = 1;
      }
      pp[m] = p; qq[m] = q; jc[m] = c; js[m] = sn; je[m] = em;
      lg[(size_t)r * np + m] = make_double4(c, sn, em.x, em.y);
    }
    __syncthreads();
    // blocks (P, Q), P <= Q: B' = J_P^H B J_Q with B = A[{p0, p1}][{q0, q1}];
    // (Q, P) is its conjugate transpose and is never stored separately
#pragma unroll
    for (int k = 0; k < kBlkPer; ++k) {
      if (tid + k * nt >= nblocks) break;
      const int P = bP[k], Q = bQ[k];
      const int p0 = pp[P], p1 = qq[P], q0 = pp[Q], q1 = qq[Q];
      // one LDS access per entry: (i, j) with i > j is the conjugate of (j, i)
      auto get = [&](int i, int j) {
        const bool lo = i > j;
        const double2 v = tri[tri_index(lo ? j : i, lo ? i : j, n)];
        return make_double2(v.x, lo ? -v.y : v.y);
      };
      auto put = [&](int i, int j, double2 v) {
        const bool lo = i > j;
        tri[tri_index(lo ? j : i, lo ? i : j, n)] = make_double2(v.x, lo ? -v.y : v.y);
      };
      const double2 J00 = make_double2(jc[Q], 0.0), J01 = make_double2(js[Q], 0.0);
      const double2 J10 = cscale(je[Q], -js[Q]), J11 = cscale(je[Q], jc[Q]);
      const double2 K00 = make_double2(jc[P], 0.0), K01 = make_double2(js[P], 0.0);
      const double2 K10 = cscale(je[P], -js[P]), K11 = cscale(je[P], jc[P]);
      const double2 a00 = get(p0, q0), a01 = get(p0, q1), a10 = get(p1, q0), a11 = get(p1, q1);
      const double2 b00 = cadd(cmul(a00, J00), cmul(a01, J10));
      const double2 b01 = cadd(cmul(a00, J01), cmul(a01, J11));
      const double2 b10 = cadd(cmul(a10, J00), cmul(a11, J10));
      const double2 b11 = cadd(cmul(a10, J01), cmul(a11, J11));
      double2 c00 = cadd(cmulc(K00, b00), cmulc(K10, b10));
      double2 c01 = cadd(cmulc(K00, b01), cmulc(K10, b11));
      double2 c10 = cadd(cmulc(K01, b00), cmulc(K11, b10));
      double2 c11 = cadd(cmulc(K01, b01), cmulc(K11, b11));
      if (P == Q) {
        c00.y = 0.0; c11.y = 0.0;
        c01 = cz();
        put(p0, q0, c00); put(p0, q1, c01); put(p1, q1, c11);
      } else {
        put(p0, q0, c00); put(p0, q1, c01); put(p1, q0, c10); put(p1, q1, c11);
      }
    }
    __syncthreads();
  }
  // write the triangle back (the lower part stays stale: readers use i <= j)
  for (int e = tid; e < ntri; e += nt) {
    int i = (int)((2.0 * n + 1.0 - sqrt((2.0 * n + 1.0) * (2.0 * n + 1.0) - 8.0 * e)) * 0.5);
    while (i > 0 && tri_index(i, i, n) > e) --i;
    while (i + 1 < n && tri_index(i + 1, i + 1, n) <= e) ++i;
    const int j = i + (e - tri_index(i, i, n));
    a[(size_t)i * ld + j] = tri[e];
  }
  if (tid == 0) {
    state[gridDim.x + b] = sweep + 1;
    // a sweep that rotated nothing leaves the log at identity: converged
    if (!any_rot) state[b] = 1;
  }
}

// V <- V J over one sweep's log; rows [16 blockIdx.x, +16) of state blockIdx.y
__global__ __launch_bounds__(1024) void jacobi_lds_rows_kernel(double2* v_all, size_t stride,
                                                               const int* ncols_dev, int ld,
                                                               const double4* log_all, size_t log_stride,
                                                               const int* state, int sweep) {
  __shared__ double2 rows[kJacobiRowsPerBlock][kJacobiLdsMax];
  const int b = blockIdx.y;
  // the sweep ran for this state iff its sweep counter reached sweep + 1
  if (state[gridDim.y + b] != sweep + 1) return;
  const int n = ncols_dev[b];
  const int r0 = blockIdx.x * kJacobiRowsPerBlock;
  if (n < 2 || r0 >= n) return;
  double2* v = v_all + (size_t)b * stride;
  const double4* lg = log_all + (size_t)b * log_stride;
  const int nr = min(kJacobiRowsPerBlock, n - r0);
  const int np = n / 2;
  for (int e = threadIdx.x; e < nr * n; e += blockDim.x) rows[e / n][e % n] = v[(size_t)(r0 + e / n) * ld + e % n];
  __syncthreads();
  for (int r = 0; r < n - 1; ++r) {
    for (int e = threadIdx.x; e < nr * np; e += blockDim.x) {
      const int i = e / np, m = e % np;
      int p = (m == 0) ? n - 1 : (r + m) % (n - 1);
      int q = (r + n - 1 - m) % (n - 1);
      if (m == 0) q = r % (n - 1);
      if (p > q) { int t = p; p = q; q = t; }
      const double4 R = lg[(size_t)r * np + m];
      const double2 em = make_double2(R.z, R.w);
      const double2 J00 = make_double2(R.x, 0.0), J01 = make_double2(R.y, 0.0);
      const double2 J10 = cscale(em, -R.y), J11 = cscale(em, R.x);
      const double2 x = rows[i][p], y = rows[i][q];
      rows[i][p] = cadd(cmul(x, J00), cmul(y, J10));
      rows[i][q] = cadd(cmul(x, J01), cmul(y, J11));
    }
    __syncthreads();
  }
  for (int e = threadIdx.x; e < nr * n; e += blockDim.x) v[(size_t)(r0 + e / n) * ld + e % n] = rows[e / n][e % n];
}

__global__ void jacobi_lds_init_kernel(double2* v_all, size_t stride, const int* ncols_dev, int ld, int* state,
                                       int batch) {
  const int b = blockIdx.y;
  const int n = ncols_dev[b];
  double2* v = v_all + (size_t)b * stride;
  for (int e = blockIdx.x * blockDim.x + threadIdx.x; e < n * n; e += gridDim.x * blockDim.x) {
    const int i = e / n, j = e % n;
    v[(size_t)i * ld + j] = make_double2(i == j ? 1.0 : 0.0, 0.0);
  }
  if (blockIdx.x == 0 && threadIdx.x == 0) {
    state[b] = 0;
    state[batch + b] = 0;
  }
}

__global__ void jacobi_lds_check_kernel(const int* state, int batch, int* error) {
  const int b = blockIdx.x * blockDim.x + threadIdx.x;
  if (b < batch && !state[b] && error) atomicOr(error, kErrEigNoConv);
}

// ------------------------------------------------ circle ordering --------
// Round r of the parallel (circle) ordering of n indices (n even): pair m of
// the n / 2 disjoint pairs, p < q; the n - 1 rounds of a sweep pair every
// two indices once.
QMX_DEV void circle_pair(int n, int r, int m, int& p, int& q) {
  p = (m == 0) ? n - 1 : (r + m) % (n - 1);
  q = (r + n - 1 - m) % (n - 1);
  if (m == 0) q = r % (n - 1);
  if (p > q) { const int t = p; p = q; q = t; }
}

// ----------------------------------------------------- block Jacobi -------
// For Gram matrices too large for one workgroup (n > 512, chi > 256): a
// two-sided block Jacobi over blocks of kBjBs columns.  The matrix is copied
// into an npad x npad buffer (npad = n rounded up to 2 kBjBs, zero padding:
// padded indices are never rotated, as every off-diagonal entry they have is
// zero), and the nb = npad / kBjBs blocks are paired by the same circle
// method as the scalar solvers: nb - 1 rounds per sweep, nb / 2 disjoint block
// pairs per round.  Per round:
//   solve: each pair's 2 kBjBs x 2 kBjBs Hermitian submatrix gets one cyclic
//          Jacobi sweep in LDS (the scalar solvers' rotation test), the
//          product U of its rotations is logged;
//   A <- A J (in place), A' = (A J)^H J into the other buffer, V <- V J (in
//   place), J the block-diagonal unitary of the round's U (columns permuted).
// After each sweep a state none of whose off-diagonal entries passes the
// rotation test is converged; the diagonal and V go back to the solver's
// own layout for select_kernel.
constexpr int kBjBs = 16;
constexpr int kBjSub = 2 * kBjBs;   // subproblem size (32)
constexpr int kBjRows = 32;         // rows per block of the apply kernel (ldp is a multiple of 64)
// The update A' = (A J)^H J is Hermitian only up to rounding, and a unitary
// conjugation carries an anti-Hermitian part forward unchanged (of order eps
// times the largest eigenvalue).  The solve rotates the Hermitian part of
// each submatrix, so the convergence test measures the Hermitian part of the
// off-diagonal entries too (testing the raw entries, a sweep never comes out
// clean once they are at that level); the rotation floor is the scalar
// solvers' (kJacobiAbsTol times the trace).

// Rotation (c, s, e^{-i phi}) of a pair from its diagonal entries al, be and
// its upper off-diagonal entry g: the scalar solvers' test (|g| above tol
// and above 1e-15 sqrt|al be|) and formulas, with a short chain of dependent
// FP64 operations for the latency-bound block solve: the test compares
// squares, and one reciprocal square root of |g|^2 gives both e^{-i phi} and
// tau = (be - al) / (2 |g|); t = sign(tau) / (|tau| + sqrt(1 + tau^2)),
// c = rsqrt(1 + t^2), s = t c.
QMX_DEV double4 bj_rotation(double al, double be, double2 g, double tol) {
  const double ag2 = g.x * g.x + g.y * g.y;
  const double thr = fmax(tol * tol, 1e-30 * fabs(al * be));
  if (!(ag2 > thr)) return make_double4(1.0, 0.0, 1.0, 0.0);
  const double ri = rsqrt(ag2);   // 1 / |g|
  const double tau = 0.5 * (be - al) * ri;
  const double t = (tau >= 0 ? 1.0 : -1.0) / (fabs(tau) + sqrt(1.0 + tau * tau));
  const double c = rsqrt(1.0 + t * t);
  return make_double4(c, t * c, g.x * ri, -g.y * ri);
}

QMX_DEV int bj_col(int nb, int r, int m, int l) {
  int P, Q;
  circle_pair(nb, r, m, P, Q);
  return l < kBjBs ? P * kBjBs + l : Q * kBjBs + (l - kBjBs);
}

// padded copies: ap = A (n x n of ld) with zeros up to npad, vp = I; scale =
// trace, conv = (n < 2), rot = 0
__global__ __launch_bounds__(256) void bj_init_kernel(const double2* a_all, size_t stride, const int* ncols_dev,
                                                      int ld, double2* ap_all, double2* vp_all, size_t pstride,
                                                      int ldp, double* scale, int* conv, int* rot) {
  const int b = blockIdx.y;
  const int n = ncols_dev[b];
  const double2* a = a_all + (size_t)b * stride;
  double2* ap = ap_all + (size_t)b * pstride;
  double2* vp = vp_all + (size_t)b * pstride;
  for (size_t e = blockIdx.x * (size_t)blockDim.x + threadIdx.x; e < (size_t)ldp * ldp;
       e += (size_t)gridDim.x * blockDim.x) {
    const int i = (int)(e / ldp), j = (int)(e % ldp);
    ap[e] = (i < n && j < n) ? a[(size_t)i * ld + j] : cz();
    vp[e] = make_double2(i == j ? 1.0 : 0.0, 0.0);
  }
  if (blockIdx.x == 0 && threadIdx.x == 0) {
    double tr = 0.0;
    for (int i = 0; i < n; ++i) tr += fabs(a[(size_t)i * ld + i].x);
    scale[b] = tr;
    conv[b] = n < 2 ? 1 : 0;
    rot[b] = 0;
  }
}

// one block per (block pair m, state b): one Jacobi sweep of the pair's
// symmetrised submatrix S in LDS; U (product of the rotations) -> log
__global__ __launch_bounds__(256) void bj_solve_kernel(const double2* ap_all, size_t pstride, int ldp, int nb,
                                                       int r, double2* log_all, const double* scale,
                                                       const int* conv) {
  __shared__ double2 S[kBjSub][kBjSub + 1];
  __shared__ double2 U[kBjSub][kBjSub + 1];
  __shared__ double4 R[kBjSub / 2];
  __shared__ int col[kBjSub];
  const int b = blockIdx.y;
  if (conv[b]) return;
  const int m = blockIdx.x;
  const double2* ap = ap_all + (size_t)b * pstride;
  if (threadIdx.x < kBjSub) col[threadIdx.x] = bj_col(nb, r, m, threadIdx.x);
  __syncthreads();
  for (int e = threadIdx.x; e < kBjSub * kBjSub; e += blockDim.x) {
    const int i = e / kBjSub, j = e % kBjSub;
    const double2 x = ap[(size_t)col[i] * ldp + col[j]], y = ap[(size_t)col[j] * ldp + col[i]];
    S[i][j] = i == j ? make_double2(x.x, 0.0) : make_double2(0.5 * (x.x + y.x), 0.5 * (x.y - y.y));
    U[i][j] = make_double2(i == j ? 1.0 : 0.0, 0.0);
  }
  __syncthreads();
  const double tol = kJacobiAbsTol * scale[b] + 1e-300;
  constexpr int np = kBjSub / 2;
  for (int ir = 0; ir < kBjSub - 1; ++ir) {
    if (threadIdx.x < np) {
      int p, q;
      circle_pair(kBjSub, ir, threadIdx.x, p, q);
      R[threadIdx.x] = bj_rotation(S[p][p].x, S[q][q].x, S[p][q], tol);
    }
    __syncthreads();
    {
      // 2x2 block (P', Q') of S: K^H S J (one block per thread, in place)
      const int Pb = threadIdx.x / np, Qb = threadIdx.x % np;
      int p0, p1, q0, q1;
      circle_pair(kBjSub, ir, Pb, p0, p1);
      circle_pair(kBjSub, ir, Qb, q0, q1);
      const double4 RP = R[Pb], RQ = R[Qb];
      const double2 K00 = make_double2(RP.x, 0.0), K01 = make_double2(RP.y, 0.0);
      const double2 K10 = cscale(make_double2(RP.z, RP.w), -RP.y);
      const double2 K11 = cscale(make_double2(RP.z, RP.w), RP.x);
      const double2 J00 = make_double2(RQ.x, 0.0), J01 = make_double2(RQ.y, 0.0);
      const double2 J10 = cscale(make_double2(RQ.z, RQ.w), -RQ.y);
      const double2 J11 = cscale(make_double2(RQ.z, RQ.w), RQ.x);
      const double2 a00 = S[p0][q0], a01 = S[p0][q1], a10 = S[p1][q0], a11 = S[p1][q1];
      const double2 b00 = cadd(cmul(a00, J00), cmul(a01, J10));
      const double2 b01 = cadd(cmul(a00, J01), cmul(a01, J11));
      const double2 b10 = cadd(cmul(a10, J00), cmul(a11, J10));
      const double2 b11 = cadd(cmul(a10, J01), cmul(a11, J11));
      double2 c00 = cadd(cmulc(K00, b00), cmulc(K10, b10));
      double2 c01 = cadd(cmulc(K00, b01), cmulc(K10, b11));
      double2 c10 = cadd(cmulc(K01, b00), cmulc(K11, b10));
      double2 c11 = cadd(cmulc(K01, b01), cmulc(K11, b11));
      if (Pb == Qb) {
        c00.y = 0.0;
        c11.y = 0.0;
        c01 = cz();
        c10 = cz();
      }
      S[p0][q0] = c00;
      S[p0][q1] = c01;
      S[p1][q0] = c10;
      S[p1][q1] = c11;
      // U <- U J: items (row i, pair k), two per thread
      for (int it = threadIdx.x; it < kBjSub * np; it += blockDim.x) {
        const int i = it / np, k = it % np;
        int p, q;
        circle_pair(kBjSub, ir, k, p, q);
        const double4 Rk = R[k];
        const double2 em = make_double2(Rk.z, Rk.w);
        const double2 x = U[i][p], y = U[i][q];
        U[i][p] = cadd(cmul(x, make_double2(Rk.x, 0.0)), cmul(y, cscale(em, -Rk.y)));
        U[i][q] = cadd(cmul(x, make_double2(Rk.y, 0.0)), cmul(y, cscale(em, Rk.x)));
      }
    }
    __syncthreads();
  }
  double2* lg = log_all + ((size_t)b * (nb / 2) + m) * kBjSub * kBjSub;
  for (int e = threadIdx.x; e < kBjSub * kBjSub; e += blockDim.x) lg[e] = U[e / kBjSub][e % kBjSub];
}

// out[i][col_m[l]] = sum_k in'[i][col_m[k]] U_m[k][l] for rows [32 blockIdx.x, +32)
// and the round's pair m = blockIdx.y; in' = in (conjt = 0) or in^H (conjt = 1).
// blockIdx.z = mat * batch + b: matrix mat (0: in0 -> out0, 1: in1 -> out1)
// of state b, so that two independent updates share one launch.  With
// conjt = 0, in == out is allowed (a block reads its whole tile first).
__global__ __launch_bounds__(256) void bj_apply_kernel(const double2* in0, double2* out0, const double2* in1,
                                                       double2* out1, size_t pstride, int ldp, int nb, int r,
                                                       const double2* log_all, int conjt, const int* conv,
                                                       int batch) {
  __shared__ double2 X[kBjRows][kBjSub + 1];
  __shared__ double2 Um[kBjSub][kBjSub + 1];
  __shared__ int col[kBjSub];
  const int mat = blockIdx.z / batch, b = blockIdx.z % batch;
  if (conv[b]) return;
  const int m = blockIdx.y, i0 = blockIdx.x * kBjRows;
  const double2* in = (mat ? in1 : in0) + (size_t)b * pstride;
  double2* out = (mat ? out1 : out0) + (size_t)b * pstride;
  const double2* lg = log_all + ((size_t)b * (nb / 2) + m) * kBjSub * kBjSub;
  if (threadIdx.x < kBjSub) col[threadIdx.x] = bj_col(nb, r, m, threadIdx.x);
  for (int e = threadIdx.x; e < kBjSub * kBjSub; e += blockDim.x) Um[e / kBjSub][e % kBjSub] = lg[e];
  __syncthreads();
  if (!conjt) {
    for (int e = threadIdx.x; e < kBjRows * kBjSub; e += blockDim.x) {
      const int ii = e / kBjSub, k = e % kBjSub;
      X[ii][k] = in[(size_t)(i0 + ii) * ldp + col[k]];
    }
  } else {
    for (int e = threadIdx.x; e < kBjRows * kBjSub; e += blockDim.x) {
      const int ii = e % kBjRows, k = e / kBjRows;
      X[ii][k] = cconj(in[(size_t)col[k] * ldp + i0 + ii]);
    }
  }
  __syncthreads();
  // thread: column l = tid % 32, rows ii = tid / 32 + 8 t (t < kBjRows / 8)
  constexpr int kPer = kBjRows / 8;
  const int l = threadIdx.x % kBjSub, ib = threadIdx.x / kBjSub;
  double2 acc[kPer];
#pragma unroll
  for (int t = 0; t < kPer; ++t) acc[t] = cz();
  for (int k = 0; k < kBjSub; ++k) {
    const double2 u = Um[k][l];
#pragma unroll
    for (int t = 0; t < kPer; ++t) acc[t] = cfma(X[ib + 8 * t][k], u, acc[t]);
  }
#pragma unroll
  for (int t = 0; t < kPer; ++t) out[(size_t)(i0 + ib + 8 * t) * ldp + col[l]] = acc[t];
}

// rot[b] = 1 if the Hermitian part of an off-diagonal entry of the current A
// passes the rotation test (rot is clear on entry)
__global__ __launch_bounds__(256) void bj_check_kernel(const double2* acur_all, size_t pstride, const int* ncols_dev,
                                                       int ldp, const double* scale, const int* conv, int* rot) {
  const int b = blockIdx.y;
  if (conv[b]) return;
  const int n = ncols_dev[b];
  const double2* a = acur_all + (size_t)b * pstride;
  const double tol = kJacobiAbsTol * scale[b] + 1e-300;
  bool pass = false;
  for (size_t e = blockIdx.x * (size_t)blockDim.x + threadIdx.x; e < (size_t)n * n;
       e += (size_t)gridDim.x * blockDim.x) {
    const int i = (int)(e / n), j = (int)(e % n);
    if (j <= i) continue;
    const double2 x = a[(size_t)i * ldp + j], y = a[(size_t)j * ldp + i];
    const double ag = 0.5 * sqrt((x.x + y.x) * (x.x + y.x) + (x.y - y.y) * (x.y - y.y));
    if (ag > tol && ag > 1e-15 * sqrt(fabs(a[(size_t)i * ldp + i].x * a[(size_t)j * ldp + j].x))) pass = true;
  }
  if (pass) rot[b] = 1;
}

// end of a sweep: a running state whose A passed no rotation test converges
__global__ void bj_sweep_end_kernel(int* conv, int* rot, int batch) {
  const int b = blockIdx.x * blockDim.x + threadIdx.x;
  if (b >= batch) return;
  if (!conv[b] && !rot[b]) conv[b] = 1;
  rot[b] = 0;
}

// eigenvalues (diagonal of the final A) and V back into the solver's layout;
// a state still running after the last sweep did not converge
__global__ __launch_bounds__(256) void bj_store_kernel(const double2* cur_all, const double2* vp_all, size_t pstride,
                                                       int ldp, double2* a_all, double2* v_all, size_t stride,
                                                       int ld, const int* ncols_dev, const int* conv, int* error) {
  const int b = blockIdx.y;
  const int n = ncols_dev[b];
  const double2* cur = cur_all + (size_t)b * pstride;
  const double2* vp = vp_all + (size_t)b * pstride;
  double2* a = a_all + (size_t)b * stride;
  double2* v = v_all + (size_t)b * stride;
  for (size_t e = blockIdx.x * (size_t)blockDim.x + threadIdx.x; e < (size_t)n * n;
       e += (size_t)gridDim.x * blockDim.x) {
    const int i = (int)(e / n), j = (int)(e % n);
    v[(size_t)i * ld + j] = vp[(size_t)i * ldp + j];
    if (i == j) a[(size_t)i * ld + i] = make_double2(cur[(size_t)i * ldp + i].x, 0.0);
  }
  if (blockIdx.x == 0 && threadIdx.x == 0 && !conv[b]) atomicOr(error, kErrEigNoConv);
}

// ---------------------------------------------------------------- select ----
// Sort eigenvalues (diag of A) in decreasing order, pick k (rsum2 cutoff and
// max bond chi; k = n when n <= chi), and write the sorted eigenvectors into
// the columns of w (same layout as v).  k_out[b] = k; spectra (optional).
__global__ __launch_bounds__(512) void select_kernel(const double2* a_all, const double2* v_all,
                                                     double2* w_all, size_t stride,
                                                     const int* ncols_dev, int ld, int chi,
                                                     double cutoff, int* k_out, int kstride,
                                                     double* spectra, size_t spec_stride) {
  __shared__ double lam[2 * kJacobiMaxPairs];
  __shared__ int perm[2 * kJacobiMaxPairs];
  const int b = blockIdx.x;
  const int n = ncols_dev[b];
  const double2* a = a_all + (size_t)b * stride;
  const double2* v = v_all + (size_t)b * stride;
  double2* w = w_all + (size_t)b * stride;
  for (int j = threadIdx.x; j < n; j += blockDim.x) {
    double lj = a[(size_t)j * ld + j].x;
    int rank = 0;
    for (int i = 0; i < n; ++i) {
      double li = a[(size_t)i * ld + i].x;
      if (li > lj || (li == lj && i < j)) ++rank;
    }
    lam[rank] = lj;
    perm[rank] = j;
  }
  __syncthreads();
  if (threadIdx.x == 0) {
    int k;
    if (n <= chi) {
      k = n;
    } else {
      double total = 0.0;
      for (int i = 0; i < n; ++i) total += fmax(lam[i], 0.0);
      k = n;
      if (total > 0.0 && cutoff > 0.0) {
        double tail = 0.0;
        for (int j = n - 1; j > 0; --j) {
          tail += fmax(lam[j], 0.0);
          if (tail <= cutoff * total) k = j; else break;
        }
      }
      if (k > chi) k = chi;
      if (k < 1) k = 1;
    }
    k_out[(size_t)b * kstride] = k;
  }
  __syncthreads();
  for (int e = threadIdx.x; e < n * n; e += blockDim.x) {
    int i = e / n, c = e % n;
    w[(size_t)i * ld + c] = v[(size_t)i * ld + perm[c]];
  }
  if (spectra)
    for (int j = threadIdx.x; j < n; j += blockDim.x) spectra[(size_t)b * spec_stride + j] = lam[j];
}

// ------------------------------------------------------------------ gemm ----
// C (m x nc, ld nc) = A (m x kd, ld kd) * op(B); B has leading dimension ldb.
// 64 x 64 output tile per block, 4 x 4 per thread, K in chunks of 16.
__global__ __launch_bounds__(256) void gemm_kernel(const double2* __restrict__ a_all, size_t a_stride,
                                                   const double2* __restrict__ b_all, size_t b_stride,
                                                   double2* c_all, size_t c_stride, int rows_log2,
                                                   const int* kdim, const int* ncols, int dim_stride,
                                                   int kmax, int nmax, int ldb, int conj_b) {
  __shared__ double2 as[64][17];
  __shared__ double2 bs[16][65];
  const int b = blockIdx.z;
  const int kd = kdim ? kdim[(size_t)b * dim_stride] : kmax;
  const int nc = ncols ? ncols[(size_t)b * dim_stride] : nmax;
  const size_t m = (size_t)1 << rows_log2;
  const size_t row0 = (size_t)blockIdx.y * 64;
  const int col0 = blockIdx.x * 64;
  if (col0 >= nc || row0 >= m) return;
  const double2* A = a_all + (size_t)b * a_stride;
  const double2* B = b_all + (size_t)b * b_stride;
  double2* C = c_all + (size_t)b * c_stride;
  const int tx = threadIdx.x & 15, ty = threadIdx.x >> 4;
  double2 acc[4][4];
#pragma unroll
  for (int i = 0; i < 4; ++i)
#pragma unroll
    for (int j = 0; j < 4; ++j) acc[i][j] = cz();
  for (int k0 = 0; k0 < kd; k0 += 16) {
    for (int e = threadIdx.x; e < 64 * 16; e += 256) {
      int rr = e >> 4, kk = e & 15;
      size_t row = row0 + rr;
      int kc = k0 + kk;
      as[rr][kk] = (row < m && kc < kd) ? A[row * kd + kc] : cz();
    }
    for (int e = threadIdx.x; e < 16 * 64; e += 256) {
      int kk = e >> 6, cc = e & 63;
      int kc = k0 + kk, col = col0 + cc;
      double2 val = cz();
      if (kc < kd && col < nc) val = conj_b ? cconj(B[(size_t)col * ldb + kc]) : B[(size_t)kc * ldb + col];
      bs[kk][cc] = val;
    }
    __syncthreads();
#pragma unroll 4
    for (int kk = 0; kk < 16; ++kk) {
      double2 av[4], bv[4];
#pragma unroll
      for (int i = 0; i < 4; ++i) av[i] = as[ty * 4 + i][kk];
#pragma unroll
      for (int j = 0; j < 4; ++j) bv[j] = bs[kk][tx * 4 + j];
#pragma unroll
      for (int i = 0; i < 4; ++i)
#pragma unroll
        for (int j = 0; j < 4; ++j) acc[i][j] = cfma(av[i], bv[j], acc[i][j]);
    }
    __syncthreads();
  }
#pragma unroll
  for (int i = 0; i < 4; ++i) {
    size_t row = row0 + ty * 4 + i;
    if (row >= m) continue;
#pragma unroll
    for (int j = 0; j < 4; ++j) {
      int col = col0 + tx * 4 + j;
      if (col < nc) C[row * nc + col] = acc[i][j];
    }
  }
}

// ------------------------------------------------------------- launchers ----
void launch_gram(const double2* y, size_t y_stride, const int* ncols_dev, int rows_log2,
                 int ncols_max, double2* part, int nchunks, int batch, hipStream_t st) {
  int nt = (ncols_max + 31) / 32;
  hipLaunchKernelGGL(gram_kernel, dim3(nt * nt, nchunks, batch), dim3(256), 0, st, y, y_stride,
                     ncols_dev, rows_log2, ncols_max, part, nchunks);
}

void launch_gram_reduce(const double2* part, int nchunks, const int* ncols_dev, int ncols_max,
                        double2* a, size_t a_stride, int batch, hipStream_t st) {
  int tot = ncols_max * ncols_max;
  hipLaunchKernelGGL(gram_reduce_kernel, dim3((tot + 255) / 256, batch), dim3(256), 0, st, part,
                     nchunks, ncols_dev, ncols_max, a, a_stride);
}

void launch_jacobi(double2* a, double2* v, size_t stride, const int* ncols_dev, int ncols_max,
                   int batch, int max_sweeps, int* error, hipStream_t st) {
  hipLaunchKernelGGL(jacobi_kernel, dim3(batch), dim3(512), 0, st, a, v, stride, ncols_dev,
                     ncols_max, max_sweeps, error);
}

// block Jacobi workspace: padded A (two buffers) and V, the round's U log,
// scale, conv, rot
int bj_ldp(int ncols_max) { return (ncols_max + 63) / 64 * 64; }

size_t jacobi_block_bytes(int ncols_max, int batch) {
  const size_t ldp = (size_t)bj_ldp(ncols_max), nb = ldp / kBjBs;
  return 3 * (size_t)batch * ldp * ldp * sizeof(double2) + 256 +
         (size_t)batch * (nb / 2) * kBjSub * kBjSub * sizeof(double2) + 256 +
         (size_t)batch * (sizeof(double) + 2 * sizeof(int)) + 256;
}

int launch_jacobi_block(double2* a, double2* v, size_t stride, const int* ncols_dev, int ncols_max, int batch,
                        void* ws, int* conv_host_buf, int max_sweeps, int* error, hipStream_t st) {
  if (stride != (size_t)ncols_max * ncols_max) return -1;
  const int ldp = bj_ldp(ncols_max), nb = ldp / kBjBs;
  const size_t pstride = (size_t)ldp * ldp;
  auto up = [](size_t x) { return (x + 255) & ~static_cast<size_t>(255); };
  char* p = static_cast<char*>(ws);
  double2* ap = reinterpret_cast<double2*>(p);
  double2* a2 = ap + (size_t)batch * pstride;
  double2* vp = a2 + (size_t)batch * pstride;
  p += up(3 * (size_t)batch * pstride * sizeof(double2));
  double2* lg = reinterpret_cast<double2*>(p);
  p += up((size_t)batch * (nb / 2) * kBjSub * kBjSub * sizeof(double2));
  double* scale = reinterpret_cast<double*>(p);
  int* conv = reinterpret_cast<int*>(scale + batch);
  int* rot = conv + batch;
  hipLaunchKernelGGL(bj_init_kernel, dim3(std::max(1, ldp * ldp / 256 / 16), batch), dim3(256), 0, st, a, stride,
                     ncols_dev, ncols_max, ap, vp, pstride, ldp, scale, conv, rot);
  double2* cur = ap;
  double2* oth = a2;
  const dim3 sgrid(nb / 2, batch), agrid2(ldp / kBjRows, nb / 2, 2 * batch), agrid(ldp / kBjRows, nb / 2, batch);
  for (int s = 0; s < max_sweeps; ++s) {
    for (int r = 0; r < nb - 1; ++r) {
      hipLaunchKernelGGL(bj_solve_kernel, sgrid, dim3(256), 0, st, cur, pstride, ldp, nb, r, lg, scale, conv);
      // A <- A J and V <- V J (in place) in one launch, then A' = (A J)^H J
      hipLaunchKernelGGL(bj_apply_kernel, agrid2, dim3(256), 0, st, cur, cur, vp, vp, pstride, ldp, nb, r, lg, 0,
                         conv, batch);
      hipLaunchKernelGGL(bj_apply_kernel, agrid, dim3(256), 0, st, cur, oth, cur, oth, pstride, ldp, nb, r, lg, 1,
                         conv, batch);
      std::swap(cur, oth);
    }
    hipLaunchKernelGGL(bj_check_kernel, dim3(256, batch), dim3(256), 0, st, cur, pstride, ncols_dev, ldp, scale,
                       conv, rot);
    hipLaunchKernelGGL(bj_sweep_end_kernel, dim3((batch + 255) / 256), dim3(256), 0, st, conv, rot, batch);
    // stop once every state has converged (one host read per sweep)
    if (hipMemcpyAsync(conv_host_buf, conv, sizeof(int) * batch, hipMemcpyDeviceToHost, st) != hipSuccess ||
        hipStreamSynchronize(st) != hipSuccess)
      return -1;
    bool all = true;
    for (int b = 0; b < batch; ++b) all = all && conv_host_buf[b];
    if (all) break;
  }
  hipLaunchKernelGGL(bj_store_kernel, dim3(std::max(1, ncols_max * ncols_max / 256 / 16), batch), dim3(256), 0, st,
                     cur, vp, pstride, ldp, a, v, stride, ncols_max, ncols_dev, conv, error);
  return 0;
}

size_t jacobi_lds_log_doubles(int ncols_max) {
  return ncols_max <= kJacobiLdsMax ? (size_t)(ncols_max - 1) * (ncols_max / 2) * 4 : 0;
}

bool jacobi_lds_fits(int ncols_max) { return ncols_max >= 2 && ncols_max <= kJacobiLdsMax; }

void launch_jacobi_lds(double2* a, double2* v, size_t stride, const int* ncols_dev, int ncols_max,
                       int batch, double* log, int* state, int max_sweeps, int* error, hipStream_t st) {
  const size_t tri_bytes = sizeof(double2) * (size_t)ncols_max * (ncols_max + 1) / 2;
  const size_t log_stride = (size_t)(ncols_max - 1) * (ncols_max / 2);
  double4* lg = reinterpret_cast<double4*>(log);
  hipLaunchKernelGGL(jacobi_lds_init_kernel, dim3((ncols_max * ncols_max + 255) / 256, batch), dim3(256), 0,
                     st, v, stride, ncols_dev, ncols_max, state, batch);
  const int row_blocks = (ncols_max + kJacobiRowsPerBlock - 1) / kJacobiRowsPerBlock;
  for (int s = 0; s < max_sweeps; ++s) {
    hipLaunchKernelGGL(jacobi_lds_sweep_kernel, dim3(batch), dim3(1024), tri_bytes, st, a, stride,
                       ncols_dev, ncols_max, lg, log_stride, state, s);
    hipLaunchKernelGGL(jacobi_lds_rows_kernel, dim3(row_blocks, batch), dim3(1024), 0, st, v, stride,
                       ncols_dev, ncols_max, lg, log_stride, state, s);
  }
  hipLaunchKernelGGL(jacobi_lds_check_kernel, dim3((batch + 255) / 256), dim3(256), 0, st, state, batch,
                     error);
}

// The dynamic-LDS limit is set to the largest triangle the kernel takes
// (kJacobiLdsMax columns), whatever ncols_max this call needs: calls from
// several host threads (one per ragged bucket) then never lower it under a
// launch that needs more.
hipError_t jacobi_lds_setup(int ncols_max) {
  if (ncols_max > kJacobiLdsMax) return hipErrorInvalidValue;
  const size_t tri_bytes = sizeof(double2) * (size_t)kJacobiLdsMax * (kJacobiLdsMax + 1) / 2;
  return hipFuncSetAttribute(reinterpret_cast<const void*>(jacobi_lds_sweep_kernel),
                             hipFuncAttributeMaxDynamicSharedMemorySize, (int)tri_bytes);
}

void launch_select(const double2* a, const double2* v, double2* w, size_t stride,
                   const int* ncols_dev, int ncols_max, int chi, double cutoff, int* k_out,
                   int kstride, double* spectra, size_t spec_stride, int batch, hipStream_t st) {
  hipLaunchKernelGGL(select_kernel, dim3(batch), dim3(512), 0, st, a, v, w, stride, ncols_dev,
                     ncols_max, chi, cutoff, k_out, kstride, spectra, spec_stride);
}

// nn[b][step] = ncur[b] = first_cols (> 0) or 2 * kk[b][step + 1]
__global__ void chain_dims_kernel(int* nn, int* ncur, const int* kk, int n, int step,
                                  int first_cols, int batch) {
  int b = blockIdx.x * blockDim.x + threadIdx.x;
  if (b >= batch) return;
  int v = first_cols > 0 ? first_cols : 2 * kk[(size_t)b * (n + 1) + step + 1];
  nn[(size_t)b * (n + 1) + step] = v;
  ncur[b] = v;
}

void launch_chain_dims(int* nn, int* ncur, const int* kk, int n, int step, int first_cols,
                       int batch, hipStream_t st) {
  hipLaunchKernelGGL(chain_dims_kernel, dim3((batch + 63) / 64), dim3(64), 0, st, nn, ncur, kk, n,
                     step, first_cols, batch);
}

// Sweep early stop: skip[b] = 1 once |<psi|C|0>| >= goal[b]; active[0] counts
// the states still running, whose sweeps_done[b] is incremented.
__global__ void goal_check_kernel(const double2* dot, int dot_stride, const double* goal, int* skip,
                                  int* active, int* sweeps_done, int batch) {
  int b = blockIdx.x * blockDim.x + threadIdx.x;
  if (b >= batch) return;
  if (!skip[b]) {
    double2 d = dot[(size_t)b * dot_stride];
    double f = sqrt(d.x * d.x + d.y * d.y);
    if (f >= goal[b]) {
      skip[b] = 1;
    } else {
      atomicAdd(active, 1);
      if (sweeps_done) sweeps_done[b] += 1;
    }
  }
}

void launch_goal_check(const double2* dot, int dot_stride, const double* goal, int* skip,
                       int* active, int* sweeps_done, int batch, hipStream_t st) {
  hipLaunchKernelGGL(goal_check_kernel, dim3((batch + 63) / 64), dim3(64), 0, st, dot, dot_stride,
                     goal, skip, active, sweeps_done, batch);
}

// one block: chunks of 1024 states, each compacted by a block-wide exclusive
// scan of the live flags
__global__ __launch_bounds__(1024) void live_list_kernel(const int* skip, int batch, int* list) {
  __shared__ int sc[1024];
  int base = 0;
  for (int b0 = 0; b0 < batch; b0 += 1024) {
    const int b = b0 + (int)threadIdx.x;
    const int live = (b < batch && !skip[b]) ? 1 : 0;
    sc[threadIdx.x] = live;
    __syncthreads();
    for (int off = 1; off < 1024; off <<= 1) {
      const int v = threadIdx.x >= (unsigned)off ? sc[threadIdx.x - off] : 0;
      __syncthreads();
      sc[threadIdx.x] += v;
      __syncthreads();
    }
    if (live) list[base + sc[threadIdx.x] - 1] = b;
    const int total = sc[1023];
    __syncthreads();
    base += total;
  }
}

void launch_live_list(const int* skip, int batch, int* list, hipStream_t st) {
  hipLaunchKernelGGL(live_list_kernel, dim3(1), dim3(1024), 0, st, skip, batch, list);
}

void launch_gemm(const double2* a, size_t a_stride, const double2* bm, size_t b_stride, double2* c,
                 size_t c_stride, const GemmDims& d, int conj_b, int batch, hipStream_t st) {
  size_t m = (size_t)1 << d.rows_log2;
  dim3 grid((d.nmax + 63) / 64, (unsigned)((m + 63) / 64), batch);
  hipLaunchKernelGGL(gemm_kernel, grid, dim3(256), 0, st, a, a_stride, bm, b_stride, c, c_stride,
                     d.rows_log2, d.kdim, d.ncols, d.dim_stride, d.kmax, d.nmax, d.ldb, conj_b);
}

}  // namespace qmx
