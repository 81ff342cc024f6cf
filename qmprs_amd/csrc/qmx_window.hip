// Windowed environment sweep (sequential.py:452-505), the dominant kernel.
//
// A window is a run of consecutive sweep steps (gates in reverse circuit
// order) whose qubits fit in a 4-bit set U.  For the window starting at gate k
// the pass kernel streams (T', phi) once:
//   T'  <- G'^dagger T' for the previous window's updated gates (in order)
//   phi <- G_k^dagger phi
//   R[x][y] = sum_{r not in U} conj(T'[x, r]) phi[y, r]     (16 x 16, MFMA f64)
//   phi <- G_j^dagger phi for the window's other gates
// and the finalize kernel replays the window on the 16-dim space: for each
// gate j the environment is the partial trace of R over U's other bits,
// G'_j = conj(polar(E_j)), then R <- G'_j^T R (target side) and
// R <- R conj(G_{j+1}) (state side).  This equals the gate-by-gate sweep
// exactly (every step of it is local to U), with one pass over HBM per window
// instead of one per gate.
//
// MFMA v_mfma_f64_16x16x4f64 operand layout (checked on gfx950):
//   A[i][k] lane i + 16k,  B[k][j] lane j + 16k,  D[i][j] lane j + 16(i%4), item i/4.
#include "qmx_device.h"
#include "qmx_internal.h"

namespace qmx {

typedef double d4v __attribute__((ext_vector_type(4)));

struct WGeom {
  int t, c0, nh;
  uint64_t qmask;
  int hoffs[1 << kMaxHighBits];
  int ux[4];            // local positions of U's bits (ascending)
  int xoff[16];         // local offset of x (bit m of x -> ux[m])
  int restpos[kWinTileBits];  // local positions outside U (ascending)
};

QMX_DEV uint64_t wdeposit(uint64_t x, uint64_t mask) {
  uint64_t r = 0;
  while (mask) {
    uint64_t low = mask & (~mask + 1ull);
    if (x & 1ull) r |= low;
    x >>= 1;
    mask &= mask - 1ull;
  }
  return r;
}

QMX_DEV int wlocal(uint64_t pmask, int gb) { return __popcll(pmask & ((1ull << gb) - 1ull)); }

template <typename S>
QMX_DEV double2 lds_get(const S* s, int l) { return to_c(s[l]); }

// gate (adjoint) on an LDS tile of storage type S, local bits m1 (site) / m0 (site+1)
template <typename S>
QMX_DEV void wtile_gate_adj(S* s, int t, int width, int m1, int m0, const double2* gm) {
  if (width == 2) {
    const int ng = 1 << (t - 2);
    const int s0 = 1 << m0, s1 = 1 << m1;
    for (int q = threadIdx.x; q < ng; q += blockDim.x) {
      int l0 = (int)insert_zero(insert_zero((uint64_t)q, m0), m1);
      double2 x0 = to_c(s[l0]), x1 = to_c(s[l0 + s0]), x2 = to_c(s[l0 + s1]), x3 = to_c(s[l0 + s0 + s1]);
#pragma unroll
      for (int o = 0; o < 4; ++o) {
        double2 acc = cmulc(gm[0 * 4 + o], x0);
        acc = cfmac(gm[1 * 4 + o], x1, acc);
        acc = cfmac(gm[2 * 4 + o], x2, acc);
        acc = cfmac(gm[3 * 4 + o], x3, acc);
        int idx = o == 0 ? l0 : (o == 1 ? l0 + s0 : (o == 2 ? l0 + s1 : l0 + s0 + s1));
        s[idx] = from_c<S>(acc);
      }
    }
  } else {
    const int ng = 1 << (t - 1);
    const int s1 = 1 << m1;
    for (int q = threadIdx.x; q < ng; q += blockDim.x) {
      int l0 = (int)insert_zero((uint64_t)q, m1);
      double2 x0 = to_c(s[l0]), x1 = to_c(s[l0 + s1]);
      double2 y0 = cfmac(gm[1 * 4 + 0], x1, cmulc(gm[0 * 4 + 0], x0));
      double2 y1 = cfmac(gm[1 * 4 + 1], x1, cmulc(gm[0 * 4 + 1], x0));
      s[l0] = from_c<S>(y0);
      s[l0 + s1] = from_c<S>(y1);
    }
  }
}

QMX_DEV int wsane(int n, int site, int w) { return (w == 2 && site < n - 1) ? 2 : 1; }

// step -> (layer, site) of the sweep order (reverse circuit order)
QMX_DEV void step_gate(int n, int s, int& layer, int& site) {
  layer = s / n;
  site = n - 1 - s % n;
}

template <typename S>
__global__ __launch_bounds__(256) void window_pass_kernel(WinPassArgs<S> a) {
  extern __shared__ char wsm[];
  __shared__ WGeom g;
  __shared__ double2 gmat[2 * kMaxWindow][16];
  __shared__ int ginfo[2 * kMaxWindow][3];   // width, m1, m0
  __shared__ int s_np, s_nc;
  const int b = blockIdx.y;
  const int n = a.n;
  if (a.skip && a.skip[b]) return;
  const int kb = a.layers[b] * n;
  if (a.s0 >= kb) return;
  const int nc = min(a.s1, kb) - a.s0;
  const int np = a.p1 - a.p0;   // previous window (all valid: p1 == s0 < kb)
  if (threadIdx.x == 0) {
    uint64_t pm = a.umask;
    int sites[2 * kMaxWindow], widths[2 * kMaxWindow];
    for (int j = 0; j < np + nc; ++j) {
      int s = j < np ? a.p0 + j : a.s0 + (j - np);
      int layer, site;
      step_gate(n, s, layer, site);
      size_t gi = ((size_t)b * a.lmax + layer) * n + site;
      int w = wsane(n, site, a.widths[gi]);
      sites[j] = site;
      widths[j] = w;
      for (int e = 0; e < 16; ++e) gmat[j][e] = a.gates[gi * 16 + e];
      pm |= 1ull << (n - 1 - site);
      if (w == 2) pm |= 1ull << (n - 2 - site);
    }
    // tile = gate bits + lowest bits
    int t = a.tbits < n ? a.tbits : n;
    uint64_t full = (1ull << n) - 1ull;
    for (int bit = 0; bit < n && __popcll(pm) < t; ++bit) pm |= 1ull << bit;
    g.t = t;
    int c0 = 0;
    while (c0 < n && ((pm >> c0) & 1ull)) ++c0;
    g.c0 = c0;
    uint64_t high = pm & ~((1ull << c0) - 1ull);
    int nh = __popcll(high);
    if (nh > kMaxHighBits) nh = kMaxHighBits;
    g.nh = nh;
    g.qmask = full & ~pm;
    for (int h = 0; h < (1 << nh); ++h) g.hoffs[h] = (int)wdeposit((uint64_t)h, high);
    int m = 0;
    for (int bit = 0; bit < n; ++bit)
      if ((a.umask >> bit) & 1ull) g.ux[m++] = wlocal(pm, bit);
    for (int x = 0; x < 16; ++x) {
      int off = 0;
      for (int q = 0; q < 4; ++q)
        if ((x >> q) & 1) off |= 1 << g.ux[q];
      g.xoff[x] = off;
    }
    int rp = 0;
    for (int l = 0; l < t; ++l) {
      bool inu = false;
      for (int q = 0; q < 4; ++q) inu |= (g.ux[q] == l);
      if (!inu) g.restpos[rp++] = l;
    }
    for (int j = 0; j < np + nc; ++j) {
      ginfo[j][0] = widths[j];
      ginfo[j][1] = wlocal(pm, n - 1 - sites[j]);
      ginfo[j][2] = widths[j] == 2 ? wlocal(pm, n - 2 - sites[j]) : 0;
    }
    s_np = np;
    s_nc = nc;
  }
  __syncthreads();
  const int t = g.t, tl = 1 << t;
  constexpr int kPer = (1 << kWinTileBits) / 256;
  const size_t len = (size_t)1 << n;
  S* sT = reinterpret_cast<S*>(wsm);
  S* sP = sT + tl;
  const bool first = (np == 0);
  // T' is first written by window 1: windows 0 and 1 read the target
  const S* tsrc = (a.p0 == 0 ? a.target : a.tprime) + (size_t)b * len;
  S* tdst = a.tprime + (size_t)b * len;
  S* phi = a.phi + (size_t)b * len;
  const long ntiles = 1l << (n - t);
  const int lane = threadIdx.x & 63, wave = threadIdx.x >> 6;
  const int nrest = t - 4;
  const int kper_wave = (1 << nrest) / 4;   // rest indices per wave per tile
  d4v acc_rr = {0, 0, 0, 0}, acc_ii = {0, 0, 0, 0}, acc_ri = {0, 0, 0, 0}, acc_ir = {0, 0, 0, 0};
  S rT[kPer], rP[kPer];
  auto load_tile = [&](long tile) {
    uint64_t base = wdeposit((uint64_t)tile, g.qmask);
#pragma unroll
    for (int j = 0; j < kPer; ++j) {
      int l = threadIdx.x + 256 * j;
      if (l < tl) {
        size_t ad = (size_t)(base + (uint64_t)(l & ((1 << g.c0) - 1)) + (uint64_t)g.hoffs[l >> g.c0]);
        rT[j] = tsrc[ad];
        rP[j] = phi[ad];
      }
    }
  };
  long tile = blockIdx.x;
  if (tile < ntiles) load_tile(tile);
  for (; tile < ntiles; tile += gridDim.x) {
#pragma unroll
    for (int j = 0; j < kPer; ++j) {
      int l = threadIdx.x + 256 * j;
      if (l < tl) { sT[l] = rT[j]; sP[l] = rP[j]; }
    }
    __syncthreads();
    long next = tile + gridDim.x;
    if (next < ntiles) load_tile(next);
    for (int j = 0; j < s_np; ++j) {
      wtile_gate_adj<S>(sT, t, ginfo[j][0], ginfo[j][1], ginfo[j][2], gmat[j]);
      __syncthreads();
    }
    wtile_gate_adj<S>(sP, t, ginfo[s_np][0], ginfo[s_np][1], ginfo[s_np][2], gmat[s_np]);
    __syncthreads();
    // R += conj(T') phi^T over this tile's rest indices: wave w takes a quarter
    {
      const int i = lane & 15, kk = lane >> 4;
      for (int ks = 0; ks < kper_wave; ks += 4) {
        int r = wave * kper_wave + ks + kk;
        int roff = 0;
        for (int q = 0; q < nrest; ++q)
          if ((r >> q) & 1) roff |= 1 << g.restpos[q];
        double2 tv = to_c(sT[g.xoff[i] + roff]);
        double2 pv = to_c(sP[g.xoff[i] + roff]);
        double ar = tv.x, ai = -tv.y;      // A = conj(T')  (row x = i, k = r)
        double br = pv.x, bi = pv.y;       // B = phi       (k = r, column y = i)
        acc_rr = __builtin_amdgcn_mfma_f64_16x16x4f64(ar, br, acc_rr, 0, 0, 0);
        acc_ii = __builtin_amdgcn_mfma_f64_16x16x4f64(ai, bi, acc_ii, 0, 0, 0);
        acc_ri = __builtin_amdgcn_mfma_f64_16x16x4f64(ar, bi, acc_ri, 0, 0, 0);
        acc_ir = __builtin_amdgcn_mfma_f64_16x16x4f64(ai, br, acc_ir, 0, 0, 0);
      }
    }
    __syncthreads();
    for (int j = 1; j < s_nc; ++j) {
      wtile_gate_adj<S>(sP, t, ginfo[s_np + j][0], ginfo[s_np + j][1], ginfo[s_np + j][2], gmat[s_np + j]);
      __syncthreads();
    }
    uint64_t base = wdeposit((uint64_t)tile, g.qmask);
    for (int l = threadIdx.x; l < tl; l += blockDim.x) {
      size_t ad = (size_t)(base + (uint64_t)(l & ((1 << g.c0) - 1)) + (uint64_t)g.hoffs[l >> g.c0]);
      if (!first) tdst[ad] = sT[l];
      phi[ad] = sP[l];
    }
    __syncthreads();
  }
  // combine waves: D[i][j] at lane j + 16 (i % 4), item i / 4
  double* red = reinterpret_cast<double*>(wsm);   // 4 waves x 512
  {
    const int j = lane & 15, ib = lane >> 4;
#pragma unroll
    for (int it = 0; it < 4; ++it) {
      int i = it * 4 + ib;
      red[wave * 512 + (i * 16 + j) * 2] = acc_rr[it] - acc_ii[it];
      red[wave * 512 + (i * 16 + j) * 2 + 1] = acc_ri[it] + acc_ir[it];
    }
  }
  __syncthreads();
  double* pb = a.part + ((size_t)b * gridDim.x + blockIdx.x) * 512;
  for (int e = threadIdx.x; e < 512; e += blockDim.x)
    pb[e] = ((red[e] + red[512 + e]) + (red[1024 + e] + red[1536 + e]));
}

// Replay the window on the 16-dim space of U (one block per state).
__global__ __launch_bounds__(256) void window_finalize_kernel(WinFinArgs a) {
  __shared__ double2 q[256];
  __shared__ double2 qn[256];
  __shared__ double2 gm[kMaxWindow][16];
  __shared__ int gw[kMaxWindow], gh[kMaxWindow], gl[kMaxWindow];
  __shared__ double2 gnew[16];
  __shared__ double part_sum[512];
  const int b = blockIdx.x;
  const int n = a.n;
  if (a.skip && a.skip[b]) return;
  const int kb = a.layers[b] * n;
  if (a.s0 >= kb) return;
  const int nc = min(a.s1, kb) - a.s0;
  // R = sum of the block partials in a fixed order
  const double* pb = a.part + (size_t)b * a.nblk * 512;
  for (int e = threadIdx.x; e < 512; e += blockDim.x) {
    double s = 0.0;
    int qq = 0;
    for (; qq + 3 < a.nblk; qq += 4)
      s += ((pb[(size_t)qq * 512 + e] + pb[(size_t)(qq + 1) * 512 + e]) +
            (pb[(size_t)(qq + 2) * 512 + e] + pb[(size_t)(qq + 3) * 512 + e]));
    for (; qq < a.nblk; ++qq) s += pb[(size_t)qq * 512 + e];
    part_sum[e] = s;
  }
  if (threadIdx.x == 0) {
    // positions of each gate's bits inside U (rank among U's bits)
    for (int j = 0; j < nc; ++j) {
      int layer, site;
      step_gate(n, a.s0 + j, layer, site);
      size_t gi = ((size_t)b * a.lmax + layer) * n + site;
      int w = wsane(n, site, a.widths[gi]);
      gw[j] = w;
      int hb = n - 1 - site;
      gh[j] = __popcll(a.umask & ((1ull << hb) - 1ull));
      gl[j] = w == 2 ? __popcll(a.umask & ((1ull << (hb - 1)) - 1ull)) : 0;
      for (int e = 0; e < 16; ++e) gm[j][e] = a.gates[gi * 16 + e];
    }
  }
  __syncthreads();
  const int x = threadIdx.x >> 4, y = threadIdx.x & 15;
  q[threadIdx.x] = make_double2(part_sum[2 * threadIdx.x], part_sum[2 * threadIdx.x + 1]);
  __syncthreads();
  for (int j = 0; j < nc; ++j) {
    const int w = gw[j], hb = gh[j], lb = gl[j];
    if (threadIdx.x == 0) {
      // E[o][i] = sum_z Q[(o,z)][(i,z)] over the U bits outside gate j
      double2 e[16];
      for (int k = 0; k < 16; ++k) e[k] = cz();
      const int d = w == 2 ? 4 : 2;
      for (int xx = 0; xx < 16; ++xx) {
        int o = w == 2 ? (((xx >> hb) & 1) << 1) | ((xx >> lb) & 1) : ((xx >> hb) & 1);
        for (int i = 0; i < d; ++i) {
          // y = x with gate bits replaced by i
          int yy = xx;
          if (w == 2) {
            yy = (yy & ~((1 << hb) | (1 << lb))) | (((i >> 1) & 1) << hb) | ((i & 1) << lb);
          } else {
            yy = (yy & ~(1 << hb)) | ((i & 1) << hb);
          }
          e[o * 4 + i] = cadd(e[o * 4 + i], q[xx * 16 + yy]);
        }
      }
      double2 u[16];
      double nuc = 0.0;
      polar4(d, e, u, &nuc);
      int layer, site;
      step_gate(n, a.s0 + j, layer, site);
      size_t gi = ((size_t)b * a.lmax + layer) * n + site;
      for (int k = 0; k < 16; ++k) {
        int o = k / 4, i = k % 4;
        double2 v = (o < d && i < d) ? cconj(u[k]) : cz();
        gnew[k] = v;
        a.gates[gi * 16 + k] = v;
      }
      if (a.s0 + j == kb - 1 && a.overlap) a.overlap[(size_t)b * a.overlap_stride] = nuc;
    }
    __syncthreads();
    // target side: Q[x][y] <- sum_{x'} G'[x'_g][x_g] Q[x'][y]
    {
      double2 s = cz();
      const int xg = w == 2 ? (((x >> hb) & 1) << 1) | ((x >> lb) & 1) : ((x >> hb) & 1);
      const int d = w == 2 ? 4 : 2;
      for (int pg = 0; pg < d; ++pg) {
        int xp = x;
        if (w == 2) xp = (xp & ~((1 << hb) | (1 << lb))) | (((pg >> 1) & 1) << hb) | ((pg & 1) << lb);
        else xp = (xp & ~(1 << hb)) | ((pg & 1) << hb);
        s = cfma(gnew[pg * 4 + xg], q[xp * 16 + y], s);
      }
      qn[threadIdx.x] = s;
    }
    __syncthreads();
    // state side: Q[x][y] <- sum_{y'} Q[x][y'] conj(G_{j+1}[y'_g][y_g])
    if (j + 1 < nc) {
      const int w2 = gw[j + 1], hb2 = gh[j + 1], lb2 = gl[j + 1];
      const int yg = w2 == 2 ? (((y >> hb2) & 1) << 1) | ((y >> lb2) & 1) : ((y >> hb2) & 1);
      const int d2 = w2 == 2 ? 4 : 2;
      double2 s = cz();
      for (int pg = 0; pg < d2; ++pg) {
        int yp = y;
        if (w2 == 2) yp = (yp & ~((1 << hb2) | (1 << lb2))) | (((pg >> 1) & 1) << hb2) | ((pg & 1) << lb2);
        else yp = (yp & ~(1 << hb2)) | ((pg & 1) << hb2);
        s = cfmac(gm[j + 1][pg * 4 + yg], qn[x * 16 + yp], s);
      }
      q[threadIdx.x] = s;
    } else {
      q[threadIdx.x] = qn[threadIdx.x];
    }
    __syncthreads();
  }
}

template <typename S>
void launch_window_pass(const WinPassArgs<S>& a, int nblk, int batch, hipStream_t st) {
  int t = a.tbits < a.n ? a.tbits : a.n;
  size_t shm = 2 * (sizeof(S) << t);
  if (shm < 4 * 512 * sizeof(double)) shm = 4 * 512 * sizeof(double);
  hipLaunchKernelGGL(window_pass_kernel<S>, dim3(nblk, batch), dim3(256), shm, st, a);
}
template void launch_window_pass<float2>(const WinPassArgs<float2>&, int, int, hipStream_t);
template void launch_window_pass<double2>(const WinPassArgs<double2>&, int, int, hipStream_t);

void launch_window_finalize(const WinFinArgs& a, int batch, hipStream_t st) {
  hipLaunchKernelGGL(window_finalize_kernel, dim3(batch), dim3(256), 0, st, a);
}

}  // namespace qmx
