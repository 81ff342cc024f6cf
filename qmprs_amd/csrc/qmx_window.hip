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

// complex64 tiles: gate arithmetic in FP32 (the storage precision)
QMX_DEV float2 fcmulc(float2 a, float2 b) {
  return make_float2(fmaf(a.x, b.x, a.y * b.y), fmaf(a.x, b.y, -a.y * b.x));
}
QMX_DEV float2 fcfmac(float2 a, float2 b, float2 acc) {
  acc.x = fmaf(a.x, b.x, acc.x);
  acc.x = fmaf(a.y, b.y, acc.x);
  acc.y = fmaf(a.x, b.y, acc.y);
  acc.y = fmaf(-a.y, b.x, acc.y);
  return acc;
}

QMX_DEV void wtile_gate_adj_f32(float2* s, int t, int width, int m1, int m0, const float2* gm) {
  if (width == 2) {
    const int ng = 1 << (t - 2);
    const int s0 = 1 << m0, s1 = 1 << m1;
    for (int q = threadIdx.x; q < ng; q += blockDim.x) {
      int l0 = (int)insert_zero(insert_zero((uint64_t)q, m0), m1);
      float2 x0 = s[l0], x1 = s[l0 + s0], x2 = s[l0 + s1], x3 = s[l0 + s0 + s1];
      float2 y[4];
#pragma unroll
      for (int o = 0; o < 4; ++o) {
        float2 acc = fcmulc(gm[0 * 4 + o], x0);
        acc = fcfmac(gm[1 * 4 + o], x1, acc);
        acc = fcfmac(gm[2 * 4 + o], x2, acc);
        acc = fcfmac(gm[3 * 4 + o], x3, acc);
        y[o] = acc;
      }
      s[l0] = y[0]; s[l0 + s0] = y[1]; s[l0 + s1] = y[2]; s[l0 + s0 + s1] = y[3];
    }
  } else {
    const int ng = 1 << (t - 1);
    const int s1 = 1 << m1;
    for (int q = threadIdx.x; q < ng; q += blockDim.x) {
      int l0 = (int)insert_zero((uint64_t)q, m1);
      float2 x0 = s[l0], x1 = s[l0 + s1];
      s[l0] = fcfmac(gm[1 * 4 + 0], x1, fcmulc(gm[0 * 4 + 0], x0));
      s[l0 + s1] = fcfmac(gm[1 * 4 + 1], x1, fcmulc(gm[0 * 4 + 1], x0));
    }
  }
}

template <typename S>
QMX_DEV void wgate(S* s, int t, int width, int m1, int m0, const double2* gm, const float2* gmf);
template <>
QMX_DEV void wgate<float2>(float2* s, int t, int width, int m1, int m0, const double2*, const float2* gmf) {
  wtile_gate_adj_f32(s, t, width, m1, m0, gmf);
}
template <>
QMX_DEV void wgate<double2>(double2* s, int t, int width, int m1, int m0, const double2* gm, const float2*) {
  wtile_gate_adj<double2>(s, t, width, m1, m0, gm);
}

QMX_DEV int wsane(int n, int site, int w) { return (w == 2 && site < n - 1) ? 2 : 1; }

// step -> (layer, site) of the sweep order (reverse circuit order)
QMX_DEV void step_gate(int n, int s, int& layer, int& site) {
  layer = s / n;
  site = n - 1 - s % n;
}

// ---- register-group gate application ------------------------------------
// A lane holds the 16 amplitudes of one group over a 4-bit set (index x, bit
// m of x = m-th lowest bit of the set).  Gate adjoints act on bits (HB, LB)
// of x (LB < 0: one-qubit gate on HB); indices are compile-time.
template <typename C> struct CT;
template <> struct CT<float> {
  using c2 = float2;
  static QMX_DEV c2 mk(double2 v) { return make_float2((float)v.x, (float)v.y); }
  static QMX_DEV c2 mulc(c2 a, c2 b) { return make_float2(fmaf(a.x, b.x, a.y * b.y), fmaf(a.x, b.y, -a.y * b.x)); }
  static QMX_DEV c2 fmac(c2 a, c2 b, c2 acc) {
    acc.x = fmaf(a.x, b.x, acc.x); acc.x = fmaf(a.y, b.y, acc.x);
    acc.y = fmaf(a.x, b.y, acc.y); acc.y = fmaf(-a.y, b.x, acc.y);
    return acc;
  }
};
template <> struct CT<double> {
  using c2 = double2;
  static QMX_DEV c2 mk(double2 v) { return v; }
  static QMX_DEV c2 mulc(c2 a, c2 b) { return cmulc(a, b); }
  static QMX_DEV c2 fmac(c2 a, c2 b, c2 acc) { return cfmac(a, b, acc); }
};

template <typename C, int HB, int LB>
QMX_DEV void grp_adj(typename CT<C>::c2 (&a)[16], const typename CT<C>::c2* g) {
  using T = CT<C>;
  if (LB >= 0) {
#pragma unroll
    for (int x = 0; x < 16; ++x) {
      if (((x >> HB) & 1) || ((x >> (LB < 0 ? 0 : LB)) & 1)) continue;
      const int i0 = x, i1 = x | (1 << (LB < 0 ? 0 : LB)), i2 = x | (1 << HB), i3 = i1 | (1 << HB);
      typename T::c2 x0 = a[i0], x1 = a[i1], x2 = a[i2], x3 = a[i3];
      typename T::c2 y[4];
#pragma unroll
      for (int o = 0; o < 4; ++o) {
        typename T::c2 acc = T::mulc(g[0 * 4 + o], x0);
        acc = T::fmac(g[1 * 4 + o], x1, acc);
        acc = T::fmac(g[2 * 4 + o], x2, acc);
        acc = T::fmac(g[3 * 4 + o], x3, acc);
        y[o] = acc;
      }
      a[i0] = y[0]; a[i1] = y[1]; a[i2] = y[2]; a[i3] = y[3];
    }
  } else {
#pragma unroll
    for (int x = 0; x < 16; ++x) {
      if ((x >> HB) & 1) continue;
      const int i0 = x, i1 = x | (1 << HB);
      typename T::c2 x0 = a[i0], x1 = a[i1];
      a[i0] = T::fmac(g[1 * 4 + 0], x1, T::mulc(g[0 * 4 + 0], x0));
      a[i1] = T::fmac(g[1 * 4 + 1], x1, T::mulc(g[0 * 4 + 1], x0));
    }
  }
}

// code = hb * 2 + (width == 2): width-2 gates have lb = hb - 1
template <typename C>
QMX_DEV void grp_apply(int code, typename CT<C>::c2 (&a)[16], const typename CT<C>::c2* g) {
  switch (code) {
    case 3: grp_adj<C, 1, 0>(a, g); break;
    case 5: grp_adj<C, 2, 1>(a, g); break;
    case 7: grp_adj<C, 3, 2>(a, g); break;
    case 0: grp_adj<C, 0, -1>(a, g); break;
    case 2: grp_adj<C, 1, -1>(a, g); break;
    case 4: grp_adj<C, 2, -1>(a, g); break;
    case 6: grp_adj<C, 3, -1>(a, g); break;
    default: break;
  }
}

// LDS element swizzle: low 5 bits XOR a fold of the higher bits (bijective)
QMX_DEV int swz(int l) { return l ^ (((l >> 5) ^ (l >> 10)) & 31); }

template <typename S> struct StoreC;
template <> struct StoreC<float2> { using C = float; };
template <> struct StoreC<double2> { using C = double; };

// deposit of the low bits of x into the set bits of mask (32-bit)
QMX_DEV uint32_t dep32(uint32_t x, uint32_t mask) {
  uint32_t r = 0;
  while (mask) {
    uint32_t low = mask & (~mask + 1u);
    if (x & 1u) r |= low;
    x >>= 1;
    mask &= mask - 1u;
  }
  return r;
}

// offset of group element x from the 4 positions p[] of its set
QMX_DEV uint32_t xoffs(const int* p, int x) {
  return (((uint32_t)x & 1u) << p[0]) | ((((uint32_t)x >> 1) & 1u) << p[1]) |
         ((((uint32_t)x >> 2) & 1u) << p[2]) | ((((uint32_t)x >> 3) & 1u) << p[3]);
}

QMX_DEV uint32_t group_roffg(const GroupMap& m, int g) { return (uint32_t)wdeposit((uint64_t)g, m.restg); }
QMX_DEV int group_roffl(const GroupMap& m, int g) { return (int)dep32((uint32_t)g, m.restl); }

// gate code of (site, width) inside a 4-bit group set: hb * 2 + (width == 2)
QMX_DEV int gate_code(uint64_t mask, int n, int site, int w) {
  int hb = __popcll(mask & ((1ull << (n - 1 - site)) - 1ull));
  return hb * 2 + (w == 2 ? 1 : 0);
}

// complex64 fits 4 waves per SIMD without spills; complex128 runs at 2 (a
// 4-wave budget would spill its FP64 groups)
// STAGED: the tiles of T' and phi are copied between HBM and LDS in tile
// order by all threads (coalesced runs of 2^c0 amplitudes, c0 = number of
// lowest state bits inside the tile) and the register groups are read from
// and written to LDS.  The host selects it for windows whose groups stride
// the state by 4 or more amplitudes per lane (gate bits at the lowest state
// bits), where direct group loads would scatter every wave-wide access.
template <typename S, int WPE, bool STAGED>
__global__ __launch_bounds__(256) __attribute__((amdgpu_waves_per_eu(WPE))) void window_pass_kernel(
    WinPassArgs<S> a) {
  using C = typename StoreC<S>::C;
  using T = CT<C>;
  using c2 = typename T::c2;
  extern __shared__ char wsm[];
  __shared__ c2 gmat[2 * kMaxWindow][16];
  __shared__ int gcode[2 * kMaxWindow];
  // group element offsets, [side][x]: side 0 = T' over the previous set,
  // side 1 = phi over the current set; global (state) and tile-local
  __shared__ uint32_t xog[2][16];
  __shared__ int xol[2][16];
  __shared__ int rdep[8];
  __shared__ uint32_t hoff[STAGED ? 256 : 1];
  const int b = a.list ? a.list[blockIdx.y] : blockIdx.y;
  const int n = a.n;
  if (a.skip && a.skip[b]) return;
  const int kb = a.layers[b] * n;
  if (a.s0 >= kb) return;
  const int nc = min(a.s1, kb) - a.s0;
  const int np = a.p1 - a.p0;   // previous window (all valid: p1 == s0 < kb)
  if (threadIdx.x >= 64 && threadIdx.x < 96) {
    const int x = threadIdx.x & 15, side = (threadIdx.x >> 4) & 1;
    xog[side][x] = side ? xoffs(a.gu.ubit, x) : xoffs(a.gp.ubit, x);
    xol[side][x] = (int)(side ? xoffs(a.gu.ux, x) : xoffs(a.gp.ux, x));
    if (threadIdx.x < 72) rdep[x] = group_roffl(a.gu, 4 * x);
  }
  // one gate entry per thread ((np + nc) * 16 <= 256)
  if ((int)threadIdx.x < (np + nc) * 16) {
    const int j = threadIdx.x >> 4, e = threadIdx.x & 15;
    const int s = j < np ? a.p0 + j : a.s0 + (j - np);
    int layer, site;
    step_gate(n, s, layer, site);
    const size_t gi = ((size_t)b * a.lmax + layer) * n + site;
    gmat[j][e] = T::mk(a.gates[gi * 16 + e]);
    if (e == 0) gcode[j] = gate_code(j < np ? a.gp.mask : a.gu.mask, n, site, wsane(n, site, a.widths[gi]));
  }
  // staged copies: tile-local index l -> state offset (l & lomask) + hoff[l >> c0]
  const uint64_t pbits = ((1ull << n) - 1ull) & ~a.qmask;
  const int c0 = STAGED ? __builtin_ctzll(~pbits) : 0;
  if (STAGED) {
    for (int h = threadIdx.x; h < (1 << (a.t - c0)); h += blockDim.x)
      hoff[h] = (uint32_t)wdeposit((uint64_t)h << c0, pbits);
  }
  __syncthreads();
  const int t = a.t;
  const size_t len = (size_t)1 << n;
  S* sT = reinterpret_cast<S*>(wsm);
  S* sP = sT + (1 << t);
  const bool first = (np == 0);
  // T' is first written by window 1: windows 0 and 1 read the target
  const S* tsrc = (a.p0 == 0 ? a.target : a.tprime) + (size_t)b * len;
  S* tdst = a.tprime + (size_t)b * len;
  S* phi = a.phi + (size_t)b * len;
  const long ntiles = 1l << (n - t);
  const int lane = threadIdx.x & 63, wave = threadIdx.x >> 6;
  const int ngroups = 1 << (t - 4);
  // lanes [0, ngroups): T' groups over the previous set; [ngroups, 2 ngroups): phi groups
  const int task = threadIdx.x;
  const bool tside = task < ngroups;
  const bool active = task < 2 * ngroups;
  const int grp = tside ? task : task - ngroups;
  // One code path for both sides: this lane's set (bit positions in the state
  // and in the tile), source/destination, gate range and the point at which
  // the group is deposited in LDS (after all previous-window gates for T',
  // after the first current-window gate for phi).
  const uint64_t restg = tside ? a.gp.restg : a.gu.restg;
  const uint32_t restl = tside ? a.gp.restl : a.gu.restl;
  const uint32_t roffg = active ? (uint32_t)wdeposit((uint64_t)grp, restg) : 0u;
  const int roffl = active ? (int)dep32((uint32_t)grp, restl) : 0;
  const uint32_t* xg = xog[tside ? 0 : 1];
  const int* xl = xol[tside ? 0 : 1];
  const S* src = tside ? tsrc : phi;
  S* dst = tside ? tdst : phi;
  S* lds = tside ? sT : sP;
  const bool store_dst = a.store && (tside ? !first : true);
  const int g0 = tside ? 0 : np;
  const int ngate = tside ? np : nc;
  const int lds_after = tside ? np : 1;
  // R phase: wave w takes rest indices [w kper, (w + 1) kper), 4 per k-step
  const int kper = ngroups / 4;
  const int xl_i = (int)xoffs(a.gu.ux, lane & 15);
  // rest index of this lane at k-step k: r = wave kper + 4k + lane / 16, three
  // disjoint bit fields (kper = 2^(t-6) >= 4), so its tile offset is the OR of
  // a per-lane part and the per-step part rdep[k]
  const int rlane = xl_i + group_roffl(a.gu, wave * kper + (lane >> 4));
  d4v acc_r = {0, 0, 0, 0}, acc_i = {0, 0, 0, 0};
  const uint32_t lomask = (1u << c0) - 1u;
  // Direct loads: a group is dead once stored, so the next tile's group is
  // loaded into the same registers right then and its loads stay in flight
  // through the barriers and the R phase.  The next tile is this block's
  // alone (tiles go round-robin over the blocks), so no other block writes it.
  c2 v[16];
  auto load_group = [&](uint32_t g) {
#pragma unroll
    for (int x = 0; x < 16; ++x) v[x] = T::mk(to_c(src[g + xg[x]]));
  };
  if (!STAGED && active && (long)blockIdx.x < ntiles)
    load_group((uint32_t)wdeposit((uint64_t)blockIdx.x, a.qmask) + roffg);
  // Staged tiles of 2^kWinTileBits amplitudes over 256 threads: the 128
  // T'-side threads copy the T' tile (16 amplitudes each, tile order), the
  // 128 phi-side threads the phi tile, through the group registers v.  The
  // next tile's loads are issued into v once v is dead (T' side: after the
  // group phase; phi side: once its final values are back in LDS), and v is
  // written to LDS at the next tile's start.
  // With a contiguous low run of c0 >= 7 bits, tile index sl0 + 128 i
  // (sl0 < 128) sits at state offset sl0 + ((128 i) & lomask) + hoff[(128 i)
  // >> c0]: a per-lane part and a block-uniform part.
  const bool fast_stage = STAGED && t == kWinTileBits && blockDim.x == 256 && c0 >= 7;
  const int sl0 = threadIdx.x & 127;
  uint32_t ust[16];
#pragma unroll
  for (int i = 0; i < 16; ++i)
    ust[i] = fast_stage ? __builtin_amdgcn_readfirstlane(((128u * i) & lomask) + hoff[(128 * i) >> c0]) : 0u;
  auto load_stage = [&](uint32_t tb) {
#pragma unroll
    for (int i = 0; i < 16; ++i) v[i] = T::mk(to_c(src[tb + ust[i] + sl0]));
  };
  if (fast_stage && (long)blockIdx.x < ntiles) load_stage((uint32_t)wdeposit((uint64_t)blockIdx.x, a.qmask));
  for (long tile = blockIdx.x; tile < ntiles; tile += gridDim.x) {
    const uint32_t base = (uint32_t)wdeposit((uint64_t)tile, a.qmask);
    const uint32_t gb = base + roffg;
    const long nxt = tile + gridDim.x;
    if (STAGED) {
      if (fast_stage) {
#pragma unroll
        for (int i = 0; i < 16; ++i) lds[swz(sl0 + 128 * i)] = from_c<S>(make_double2(v[i].x, v[i].y));
      } else {
        for (int l = threadIdx.x; l < (1 << t); l += blockDim.x) {
          const uint32_t off = base + ((uint32_t)l & lomask) + hoff[l >> c0];
          sT[swz(l)] = tsrc[off];
          sP[swz(l)] = phi[off];
        }
      }
      __syncthreads();
    }
    if (active) {
      if (STAGED) {
#pragma unroll
        for (int x = 0; x < 16; ++x) v[x] = T::mk(to_c(lds[swz(roffl + xl[x])]));
      }
      for (int j = 0; j <= ngate; ++j) {
        if (j == lds_after) {
#pragma unroll
          for (int x = 0; x < 16; ++x)
            lds[swz(roffl + xl[x])] = from_c<S>(make_double2(v[x].x, v[x].y));
        }
        if (j < ngate) grp_apply<C>(gcode[g0 + j], v, gmat[g0 + j]);
      }
      if (!STAGED && store_dst) {
#pragma unroll
        for (int x = 0; x < 16; ++x) dst[gb + xg[x]] = from_c<S>(make_double2(v[x].x, v[x].y));
      }
      if (!STAGED) {
        if (nxt < ntiles) load_group((uint32_t)wdeposit((uint64_t)nxt, a.qmask) + roffg);
      }
      if (fast_stage && tside && nxt < ntiles) load_stage((uint32_t)wdeposit((uint64_t)nxt, a.qmask));
    }
    __syncthreads();
    // R += conj(T') phi^T over this tile: A[x][k] = conj(T'[x, r]), B[k][y] = phi[y, r]
#pragma unroll
    for (int k = 0; k < 8; ++k) {
      if (4 * k >= kper) break;
      const int li = swz(rlane | rdep[k]);
      double2 tv = to_c(sT[li]);
      double2 pv = to_c(sP[li]);
      // Re R += Re T Re phi + Im T Im phi,  Im R += Re T Im phi - Im T Re phi
      acc_r = __builtin_amdgcn_mfma_f64_16x16x4f64(tv.x, pv.x, acc_r, 0, 0, 0);
      acc_r = __builtin_amdgcn_mfma_f64_16x16x4f64(tv.y, pv.y, acc_r, 0, 0, 0);
      acc_i = __builtin_amdgcn_mfma_f64_16x16x4f64(tv.x, pv.y, acc_i, 0, 0, 0);
      acc_i = __builtin_amdgcn_mfma_f64_16x16x4f64(-tv.y, pv.x, acc_i, 0, 0, 0);
    }
    __syncthreads();
    if (STAGED) {
      // phi's final values (all window gates) replace its R operand in LDS,
      // then both tiles go back to HBM in tile order (T' holds its final
      // values since the group phase)
      if (a.store && active && !tside) {
#pragma unroll
        for (int x = 0; x < 16; ++x) sP[swz(roffl + xl[x])] = from_c<S>(make_double2(v[x].x, v[x].y));
      }
      if (fast_stage && !tside && nxt < ntiles) load_stage((uint32_t)wdeposit((uint64_t)nxt, a.qmask));
    }
    if (STAGED && a.store) {
      __syncthreads();
      if (fast_stage) {
        if (store_dst) {
#pragma unroll
          for (int i = 0; i < 16; ++i) dst[base + ust[i] + sl0] = lds[swz(sl0 + 128 * i)];
        }
      } else {
        for (int l = threadIdx.x; l < (1 << t); l += blockDim.x) {
          const uint32_t off = base + ((uint32_t)l & lomask) + hoff[l >> c0];
          if (!first) tdst[off] = sT[swz(l)];
          phi[off] = sP[swz(l)];
        }
      }
      __syncthreads();
    }
  }
  // combine waves: D[i][j] at lane j + 16 (i % 4), item i / 4
  double* red = reinterpret_cast<double*>(wsm);   // 4 waves x 512
  {
    const int j = lane & 15, ib = lane >> 4;
#pragma unroll
    for (int it = 0; it < 4; ++it) {
      int i = it * 4 + ib;
      red[wave * 512 + (i * 16 + j) * 2] = acc_r[it];
      red[wave * 512 + (i * 16 + j) * 2 + 1] = acc_i[it];
    }
  }
  __syncthreads();
  double* pb = a.part + ((size_t)b * gridDim.x + blockIdx.x) * 512;
  for (int e = threadIdx.x; e < 512; e += blockDim.x)
    pb[e] = ((red[e] + red[512 + e]) + (red[1024 + e] + red[1536 + e]));
}

// ---- forward circuit pass ------------------------------------------------
// Applies the gates of up to kMaxStages stages (stage k: forward circuit
// positions fpos[gstart[k] .. gstart[k+1]), gate (layer L_b - 1 - f / N,
// site f % N), positions f >= L_b N skipped) to every state.  Stage 0 reads
// HBM, the last stage writes HBM, intermediate stages exchange through the
// LDS tile.
// Support: the circuit starts from |0..0>, and a gate only spreads the state
// over its own bits, so before the pass every amplitude at an index with a
// bit outside live_in is zero (not read: the memory there may be stale) and
// after it every amplitude outside live_out is.  Only tiles inside live_out
// are visited (tmask: the tile-index bits), and only amplitudes inside
// live_out are written, except in a full pass, which writes every amplitude.
// from_basis0: the input is |0..0> itself (psi is not read).
template <typename S>
__global__ __launch_bounds__(256) void fwd_pass_kernel(FwdArgs<S> a) {
  using C = typename StoreC<S>::C;
  using T = CT<C>;
  using c2 = typename T::c2;
  extern __shared__ char wsm[];
  __shared__ c2 gmat[kMaxFwdGates][16];
  __shared__ int gcode[kMaxFwdGates];
  const int b = a.list ? a.list[blockIdx.y] : blockIdx.y;
  const int n = a.n;
  if (a.skip && a.skip[b]) return;
  const int lb = a.layers[b];
  const int ng = a.gstart[a.nstage];
  // one gate entry per thread (kMaxFwdGates * 16 <= 4 * 256); positions past
  // this state's circuit get code -1 (no gate)
  for (int t = threadIdx.x; t < ng * 16; t += blockDim.x) {
    const int e = t & 15, j = t >> 4;
    const int f = a.fpos[j];
    if (f < lb * n) {
      int st = 0;
      while (a.gstart[st + 1] <= j) ++st;
      const int layer = lb - 1 - f / n, site = f % n;
      const size_t gi = ((size_t)b * a.lmax + layer) * n + site;
      // grp_apply applies the adjoint of what it is given: stage G^dagger
      const double2 gv = a.gates[gi * 16 + (e & 3) * 4 + (e >> 2)];
      gmat[j][e] = T::mk(make_double2(gv.x, -gv.y));
      if (e == 0) gcode[j] = gate_code(a.gs[st].mask, n, site, wsane(n, site, a.widths[gi]));
    } else if (e == 0) {
      gcode[j] = -1;
    }
  }
  __syncthreads();
  int total = 0;
  for (int j = 0; j < ng; ++j) total += a.fpos[j] < lb * n ? 1 : 0;
  const uint64_t all = (1ull << n) - 1ull;
  // a dense pass without gates leaves the state as it is
  if (total == 0 && !a.from_basis0 && a.live_in == all) return;
  const int t = a.t;
  const size_t len = (size_t)1 << n;
  S* lds = reinterpret_cast<S*>(wsm);
  S* psi = a.psi + (size_t)b * len;
  const long ntiles = 1l << __popcll(a.tmask);
  const int ngroups = 1 << (t - 4);
  const uint32_t dead_in = (uint32_t)(~a.live_in & all), dead_out = (uint32_t)(~a.live_out & all);
  // One group per thread (ngroups <= 256 = blockDim for tiles of <= 2^12
  // amplitudes).  Stage 0 always uses the same group map, so the next tile's
  // stage-0 loads are issued as soon as this tile's are consumed and stay in
  // flight while this tile's stages run.
  const bool has = (int)threadIdx.x < ngroups;
  const GroupMap& g0m = a.gs[0];
  const uint32_t roffg0 = has ? group_roffg(g0m, threadIdx.x) : 0u;
  c2 pre[16];
  auto load0 = [&](uint32_t bse) {
#pragma unroll
    for (int x = 0; x < 16; ++x) {
      const uint32_t ad = bse + roffg0 + xoffs(g0m.ubit, x);
      if (ad & dead_in)
        pre[x] = T::mk(make_double2(0.0, 0.0));
      else if (a.from_basis0)
        pre[x] = T::mk(make_double2(ad == 0u ? 1.0 : 0.0, 0.0));
      else
        pre[x] = T::mk(to_c(psi[ad]));
    }
  };
  if (has && (long)blockIdx.x < ntiles) load0((uint32_t)wdeposit((uint64_t)blockIdx.x, a.tmask));
  for (long tile = blockIdx.x; tile < ntiles; tile += gridDim.x) {
    const uint32_t base = (uint32_t)wdeposit((uint64_t)tile, a.tmask);
    for (int st = 0; st < a.nstage; ++st) {
      const GroupMap& gm = a.gs[st];
      const bool last = (st + 1 == a.nstage);
      if (has) {
        const int grp = threadIdx.x;
        const uint32_t roffg = st == 0 ? roffg0 : group_roffg(gm, grp);
        const int roffl = group_roffl(gm, grp);
        c2 v[16];
        if (st == 0) {
#pragma unroll
          for (int x = 0; x < 16; ++x) v[x] = pre[x];
          const long nxt = tile + gridDim.x;
          if (nxt < ntiles) load0((uint32_t)wdeposit((uint64_t)nxt, a.tmask));
        } else {
#pragma unroll
          for (int x = 0; x < 16; ++x) v[x] = T::mk(to_c(lds[swz(roffl + (int)xoffs(gm.ux, x))]));
        }
        for (int j = a.gstart[st]; j < a.gstart[st + 1]; ++j) grp_apply<C>(gcode[j], v, gmat[j]);
        if (last) {
#pragma unroll
          for (int x = 0; x < 16; ++x) {
            const uint32_t ad = base + roffg + xoffs(gm.ubit, x);
            if (a.full || !(ad & dead_out)) psi[ad] = from_c<S>(make_double2(v[x].x, v[x].y));
          }
        } else {
#pragma unroll
          for (int x = 0; x < 16; ++x) lds[swz(roffl + (int)xoffs(gm.ux, x))] = from_c<S>(make_double2(v[x].x, v[x].y));
        }
      }
      // stage exchange, and the next tile's stage 0 must not overwrite LDS early
      if (a.nstage > 1) __syncthreads();
    }
  }
}

// The forward pass on the matrix cores (tiles of 2^12 amplitudes; complex64
// states with N <= 28, complex128 with N <= 27): each stage's gates are fused
// into one 16 x 16 complex matrix M (column c = the stage's gates applied in
// order to the unit vector e_c of the stage's 4-bit group index), and each
// wave applies M to its groups, 16 at a time, as complex products on
// v_mfma_f32_16x16x4f32 or v_mfma_f64_16x16x4f64 (three real products per
// k-step, below).  Operand layouts (checked on gfx950): A[i][k] in lane
// i + 16 k and B[k][j] in lane j + 16 k for both; D[i][j] in lane
// j + 16 (i / 4), item i % 4 for f32 and in lane j + 16 (i % 4), item i / 4
// for f64.  Lane jl + 16 kq of a wave therefore holds, for column block cb,
// the amplitudes P(kq, s) (s < 4) of one group as B operands and receives the
// amplitudes P(kq, r) (r < 4) of the same group in item r, when the rows of M
// are ordered by MfmaT::row; P and the group's rest bits follow the stage's
// lane roles (MfmaMap, chosen on the host).  Loads, support (live_in /
// live_out) and the LDS exchange between stages are those of fwd_pass_kernel.
typedef float f4v __attribute__((ext_vector_type(4)));

// Per-type parts of the matrix-core pass.  A tile holds 256 groups, run by
// 8 waves of 2 column blocks (16 groups each): complex64 at 6 waves per SIMD
// (3 blocks per CU while their LDS fits; 4 waves of 4 column blocks ran 3%
// slower at 4 per SIMD), complex128 at 2 (its 64 KiB tile leaves room for
// one block per CU).
template <typename S> struct MfmaT;
template <> struct MfmaT<float2> {
  using R = float;
  using V4 = f4v;
  static constexpr int kWaves = 8;
  static constexpr int kShift = 3;   // log2 bytes per amplitude
  static QMX_DEV V4 mfma(R x, R y, V4 c) { return __builtin_amdgcn_mfma_f32_16x16x4f32(x, y, c, 0, 0, 0); }
  static QMX_DEV float2 mk(double x, double y) { return make_float2((float)x, (float)y); }
  static QMX_DEV int row(int q, int r) { return 4 * q + r; }
};
template <> struct MfmaT<double2> {
  using R = double;
  using V4 = d4v;
  static constexpr int kWaves = 8;
  static constexpr int kShift = 4;
  static QMX_DEV V4 mfma(R x, R y, V4 c) { return __builtin_amdgcn_mfma_f64_16x16x4f64(x, y, c, 0, 0, 0); }
  static QMX_DEV double2 mk(double x, double y) { return make_double2(x, y); }
  static QMX_DEV int row(int q, int r) { return q + 4 * r; }
};

// index parts of the matrix-core pass (positions from MfmaMap words, one
// per byte): the lane's (jl at j[0..3], kq at ks[0..1]), the wave's (w at
// h[NB..3]), and the uniform (cb at h[0..NB-1], s at ks[2..3]), NB = 2 with 4
// column blocks per wave and 1 with 2; the three are disjoint bit fields
QMX_DEV uint32_t pos8(uint32_t w, int i) { return (w >> (8 * i)) & 255u; }
QMX_DEV uint32_t lanepart(uint32_t j, uint32_t ks, int jl, int kq) {
  return ((uint32_t)(jl & 1) << pos8(j, 0)) | ((uint32_t)((jl >> 1) & 1) << pos8(j, 1)) |
         ((uint32_t)((jl >> 2) & 1) << pos8(j, 2)) | ((uint32_t)(jl >> 3) << pos8(j, 3)) |
         ((uint32_t)(kq & 1) << pos8(ks, 0)) | ((uint32_t)(kq >> 1) << pos8(ks, 1));
}
template <int NB> QMX_DEV uint32_t wavepart(uint32_t h, int w) {
  uint32_t r = 0;
#pragma unroll
  for (int i = NB; i < 4; ++i) r |= (uint32_t)((w >> (i - NB)) & 1) << pos8(h, i);
  return r;
}
template <int NB> QMX_DEV uint32_t unipart(uint32_t h, uint32_t ks, int cb, int s) {
  uint32_t r = ((uint32_t)(s & 1) << pos8(ks, 2)) | ((uint32_t)(s >> 1) << pos8(ks, 3));
#pragma unroll
  for (int i = 0; i < NB; ++i) r |= (uint32_t)((cb >> i) & 1) << pos8(h, i);
  return r;
}

template <typename S>
__attribute__((always_inline)) QMX_DEV void fwd_mfma_body(const FwdArgs<S>& a) {
  using M = MfmaT<S>;
  using R = typename M::R;
  using V4 = typename M::V4;
  constexpr int kCB = 16 / M::kWaves;       // column blocks of 16 groups per wave
  constexpr int kNB = kCB == 4 ? 2 : 1;     // their index bits
  extern __shared__ char wsm[];
  __shared__ int gcode[kMaxFwdGates];
  __shared__ uint32_t mmw[kMaxStages][3];  // per stage: jl, hl, ksl of MfmaMap
  const int b = a.list ? a.list[blockIdx.y] : blockIdx.y;
  const int n = a.n;
  if (a.skip && a.skip[b]) return;
  const int lb = a.layers[b];
  const int ng = a.gstart[a.nstage];
  const int t = a.t;   // kFwdTileBits: 256 groups of 16, 16 * kCB per wave
  S* lds = reinterpret_cast<S*>(wsm);
  // the stage matrices after the tile, in A-operand order: with P(q, s) the
  // amplitude index whose bits kx are q and bits sx are s (lane roles),
  // mst[st][s][row(q, r) + 16 kq] = M[P(q, r)][P(kq, s)], so that a wave's
  // A load is contiguous
  S* mst = lds + ((size_t)1 << t);
  // gates (as G^dagger, which grp_apply applies as G) in the tile area, which
  // stage 0 only writes after the matrices are built
  S(*gmat)[16] = reinterpret_cast<S(*)[16]>(wsm);
  for (int e2 = threadIdx.x; e2 < ng * 16; e2 += blockDim.x) {
    const int e = e2 & 15, j = e2 >> 4;
    const int f = a.fpos[j];
    if (f < lb * n) {
      int st = 0;
      while (a.gstart[st + 1] <= j) ++st;
      const int layer = lb - 1 - f / n, site = f % n;
      const size_t gi = ((size_t)b * a.lmax + layer) * n + site;
      const double2 gv = a.gates[gi * 16 + (e & 3) * 4 + (e >> 2)];
      gmat[j][e] = M::mk(gv.x, -gv.y);
      if (e == 0) gcode[j] = gate_code(a.gs[st].mask, n, site, wsane(n, site, a.widths[gi]));
    } else if (e == 0) {
      gcode[j] = -1;
    }
  }
  __syncthreads();
  int total = 0;
  for (int j = 0; j < ng; ++j) total += a.fpos[j] < lb * n ? 1 : 0;
  const uint64_t all = (1ull << n) - 1ull;
  if (total == 0 && !a.from_basis0 && a.live_in == all) return;
  if ((int)threadIdx.x < a.nstage * 16) {
    const int st = threadIdx.x >> 4, c = threadIdx.x & 15;
    const MfmaMap& mm = a.mm[st];
    S v[16];
#pragma unroll
    for (int x = 0; x < 16; ++x) v[x] = M::mk(x == c ? 1.0 : 0.0, 0.0);
    for (int j = a.gstart[st]; j < a.gstart[st + 1]; ++j) grp_apply<R>(gcode[j], v, gmat[j]);
    // amplitude index x <-> (q, s): q = its bits kx, s = its bits sx
    const uint32_t kx0 = pos8(mm.kxsx, 0), kx1 = pos8(mm.kxsx, 1), sx0 = pos8(mm.kxsx, 2), sx1 = pos8(mm.kxsx, 3);
    auto qof = [&](int x) { return (int)(((x >> kx0) & 1) | (((x >> kx1) & 1) << 1)); };
    auto sof = [&](int x) { return (int)(((x >> sx0) & 1) | (((x >> sx1) & 1) << 1)); };
    const int kc = qof(c), sc = sof(c);
#pragma unroll
    for (int x = 0; x < 16; ++x) mst[(st * 4 + sc) * 64 + 16 * kc + M::row(qof(x), sof(x))] = v[x];
    if (c == 0) {
      mmw[st][0] = mm.jl;
      mmw[st][1] = mm.hl;
      mmw[st][2] = mm.ksl;
    }
  }
  __syncthreads();
  const size_t len = (size_t)1 << n;
  // 32-bit byte offsets from a scalar base (below 2^32 for N <= 28 in
  // complex64 and N <= 27 in complex128)
  char* pbytes = reinterpret_cast<char*>(a.psi + (size_t)b * len);
  const long ntiles = 1l << __popcll(a.tmask);
  const uint32_t dead_in = (uint32_t)(~a.live_in & all), dead_out = (uint32_t)(~a.live_out & all);
  const int lane = threadIdx.x & 63;
  const int wave = __builtin_amdgcn_readfirstlane(threadIdx.x >> 6);
  const int jl = lane & 15, kq = lane >> 4;
  // Every index below is an OR of disjoint bit fields, split into a per-lane
  // part (jl: the 4 low bits of the group index, kq: 2 bits of the amplitude
  // index) and a part that is uniform over the wave (wave, cb, s4 / r).  The
  // deposits and xoffs are linear over disjoint fields and the LDS swizzle is
  // linear over XOR, so an access costs one OR or XOR of the two parts.
  const MfmaMap& m0 = a.mm[0];
  const MfmaMap& mL = a.mm[a.nstage - 1];
  const uint32_t h0g = __builtin_amdgcn_readfirstlane(m0.hg), k0g = __builtin_amdgcn_readfirstlane(m0.ksg);
  const uint32_t hLg = __builtin_amdgcn_readfirstlane(mL.hg), kLg = __builtin_amdgcn_readfirstlane(mL.ksg);
  const uint32_t g0lane = lanepart(m0.jg, k0g, jl, kq);
  const uint32_t gLlane = lanepart(mL.jg, kLg, jl, kq);
  const uint32_t g0w = wavepart<kNB>(h0g, wave), gLw = wavepart<kNB>(hLg, wave);
  // B operands of stage 0 for one tile: [cb * 4 + s] = amplitude P(kq, s)
  S pre[4 * kCB];
  auto load0 = [&](uint32_t bse) {
#pragma unroll
    for (int cb = 0; cb < kCB; ++cb)
#pragma unroll
      for (int s4 = 0; s4 < 4; ++s4) {
        const uint32_t ad = bse + (g0w | unipart<kNB>(h0g, k0g, cb, s4)) + g0lane;
        S v;
        if (ad & dead_in)
          v = M::mk(0.0, 0.0);
        else if (a.from_basis0)
          v = M::mk(ad == 0u ? 1.0 : 0.0, 0.0);
        else
          v = *reinterpret_cast<const S*>(pbytes + (ad << M::kShift));
        pre[cb * 4 + s4] = v;
      }
  };

  for (long tile = blockIdx.x; tile < ntiles; tile += gridDim.x) {
    const uint32_t base = (uint32_t)wdeposit((uint64_t)tile, a.tmask);
    // stage 0's operands come from HBM at the tile's start: a prefetch of
    // the next tile's would cost the occupancy that hides the load latency
    // (measured with 4 column blocks per wave: 3 instead of 4 waves per
    // SIMD, slower)
    load0(base);
    for (int st = 0; st < a.nstage; ++st) {
      // lane roles of the stage (uniform: LDS broadcast, then scalar)
      const uint32_t mjl = __builtin_amdgcn_readfirstlane(mmw[st][0]);
      const uint32_t mhl = __builtin_amdgcn_readfirstlane(mmw[st][1]);
      const uint32_t mksl = __builtin_amdgcn_readfirstlane(mmw[st][2]);
      const bool last = (st + 1 == a.nstage);
      // tile-local offsets: per lane (jl, kq, swizzled once per stage) and
      // per (wave, cb, s or r)
      const int lsw = swz((int)lanepart(mjl, mksl, jl, kq));
      const uint32_t lw = wavepart<kNB>(mhl, wave);
      S xv[4 * kCB];
      if (st == 0) {
#pragma unroll
        for (int q = 0; q < 4 * kCB; ++q) xv[q] = pre[q];
      } else {
#pragma unroll
        for (int cb = 0; cb < kCB; ++cb)
#pragma unroll
          for (int s4 = 0; s4 < 4; ++s4) xv[cb * 4 + s4] = lds[lsw ^ swz((int)(lw | unipart<kNB>(mhl, mksl, cb, s4)))];
      }
      // A operands: M[jl][4 s + kq]
      S am[4];
#pragma unroll
      for (int s4 = 0; s4 < 4; ++s4) am[s4] = mst[(st * 4 + s4) * 64 + lane];
      // one column block at a time, each complex product as three real ones
      // (Gauss): k1 = Re M Re x, k2 = Im M Im x, k3 = (Re M + Im M)(Re x +
      // Im x); Re y = k1 - k2, Im y = k3 - k1 - k2.  The three accumulators
      // alternate, so consecutive products are independent.
      R asum[4];
#pragma unroll
      for (int s4 = 0; s4 < 4; ++s4) asum[s4] = am[s4].x + am[s4].y;
#pragma unroll
      for (int cb = 0; cb < kCB; ++cb) {
        V4 k1 = {0, 0, 0, 0}, k2 = {0, 0, 0, 0}, k3 = {0, 0, 0, 0};
#pragma unroll
        for (int s4 = 0; s4 < 4; ++s4) {
          const S x = xv[cb * 4 + s4];
          k1 = M::mfma(am[s4].x, x.x, k1);
          k2 = M::mfma(am[s4].y, x.y, k2);
          k3 = M::mfma(asum[s4], x.x + x.y, k3);
        }
        const V4 yr = k1 - k2, yi = k3 - k1 - k2;
        if (last) {
#pragma unroll
          for (int r = 0; r < 4; ++r) {
            const uint32_t ad = base + (gLw | unipart<kNB>(hLg, kLg, cb, r)) + gLlane;
            if (a.full || !(ad & dead_out))
              *reinterpret_cast<S*>(pbytes + (ad << M::kShift)) = M::mk(yr[r], yi[r]);
          }
        } else {
#pragma unroll
          for (int r = 0; r < 4; ++r)
            lds[lsw ^ swz((int)(lw | unipart<kNB>(mhl, mksl, cb, r)))] = M::mk(yr[r], yi[r]);
        }
      }
      // stage exchange, and the next tile's stage 0 must not overwrite LDS early
      if (a.nstage > 1) __syncthreads();
    }
  }
}

__global__ __launch_bounds__(512) __attribute__((amdgpu_waves_per_eu(6, 6))) void fwd_mfma_kernel(FwdArgs<float2> a) {
  fwd_mfma_body(a);
}
__global__ __launch_bounds__(512) __attribute__((amdgpu_waves_per_eu(2, 2))) void fwd_mfma64_kernel(FwdArgs<double2> a) {
  fwd_mfma_body(a);
}

// Replay the window on the 16-dim space of U (one block per state).
// It runs beside the other half-batch's window pass, whose waves fill the
// SIMDs; its own waves take the highest issue priority so that this short,
// latency-bound chain is not slowed down by them.
__global__ __launch_bounds__(256) void window_finalize_kernel(WinFinArgs a) {
  __builtin_amdgcn_s_setprio(3);
  __shared__ double2 q[256];
  __shared__ double2 qn[256];
  __shared__ double2 gm[kMaxWindow][16];
  __shared__ int gw[kMaxWindow], gh[kMaxWindow], gl[kMaxWindow];
  __shared__ double2 gnew[16];
  __shared__ double2 es[16];
  __shared__ double2 rm[16];
  __shared__ double part_sum[512];
  __shared__ double s_nuc;
  __shared__ int s_ok;
  const int b = a.list ? a.list[blockIdx.x] : blockIdx.x;
  const int n = a.n;
  if (a.skip && a.skip[b]) return;
  const int kb = a.layers[b] * n;
  if (a.s0 >= kb) return;
  const int nc = min(a.s1, kb) - a.s0;
  // the window's gates, one entry per thread (nc * 16 <= 128), and the
  // positions of each gate's bits inside U (rank among U's bits); issued
  // before the partial sum so that both latencies overlap
  if ((int)threadIdx.x < nc * 16) {
    const int j = threadIdx.x >> 4, e = threadIdx.x & 15;
    int layer, site;
    step_gate(n, a.s0 + j, layer, site);
    const size_t gi = ((size_t)b * a.lmax + layer) * n + site;
    gm[j][e] = a.gates[gi * 16 + e];
    if (e == 0) {
      const int w = wsane(n, site, a.widths[gi]);
      gw[j] = w;
      const int hb = n - 1 - site;
      gh[j] = __popcll(a.umask & ((1ull << hb) - 1ull));
      gl[j] = w == 2 ? __popcll(a.umask & ((1ull << (hb - 1)) - 1ull)) : 0;
    }
  }
  // R = sum of the block partials (slots 0, ps, 2 ps, ...) in a fixed order
  const double* pb = a.part + (size_t)b * a.nblk * 512;
  const int ps = a.pstride > 0 ? a.pstride : 1;
  const size_t st4 = (size_t)ps * 512;
  // (8 partials per step, all loads issued before the adds: the partials
  // come from L2 / HBM and the sum is latency-bound)
  for (int e = threadIdx.x; e < 512; e += blockDim.x) {
    double s = 0.0;
    int qq = 0;
#pragma unroll 2
    for (; qq + 7 * ps < a.nblk; qq += 8 * ps) {
      const double* q0 = pb + (size_t)qq * 512 + e;
      double v[8];
#pragma unroll
      for (int u = 0; u < 8; ++u) v[u] = q0[u * st4];
      s += (((v[0] + v[1]) + (v[2] + v[3])) + ((v[4] + v[5]) + (v[6] + v[7])));
    }
    for (; qq < a.nblk; qq += ps) s += pb[(size_t)qq * 512 + e];
    part_sum[e] = s;
  }
  __syncthreads();
  const int x = threadIdx.x >> 4, y = threadIdx.x & 15;
  q[threadIdx.x] = make_double2(part_sum[2 * threadIdx.x], part_sum[2 * threadIdx.x + 1]);
  __syncthreads();
  for (int j = 0; j < nc; ++j) {
    const int w = gw[j], hb = gh[j], lb = gl[j];
    // E[o][i] = sum_z Q[(o,z)][(i,z)] over the U bits outside gate j (16 lanes)
    if (threadIdx.x < 16) {
      const int o = threadIdx.x >> 2, i = threadIdx.x & 3;
      const int d = w == 2 ? 4 : 2;
      double2 acc = cz();
      if (o < d && i < d) {
        for (int xx = 0; xx < 16; ++xx) {
          int go = w == 2 ? (((xx >> hb) & 1) << 1) | ((xx >> lb) & 1) : ((xx >> hb) & 1);
          if (go != o) continue;
          int yy;
          if (w == 2) yy = (xx & ~((1 << hb) | (1 << lb))) | (((i >> 1) & 1) << hb) | ((i & 1) << lb);
          else yy = (xx & ~(1 << hb)) | ((i & 1) << hb);
          acc = cadd(acc, q[xx * 16 + yy]);
        }
      }
      es[threadIdx.x] = acc;
    }
    __syncthreads();
    const int d = w == 2 ? 4 : 2;
    const bool last_gate = a.s0 + j == kb - 1 && a.overlap;
    if (threadIdx.x < 16) {
      // polar(E) by the scaled Newton iteration on 16 lanes (one entry
      // each); the iterate is E, or diag(E, |E|_F I) for a 2 x 2 problem
      const int l = threadIdx.x, r = l >> 2, c = l & 3;
      const bool in = r < d && c < d;
      double2 xv = in ? es[l] : cz();
      const double fro = sum16(cabs2(xv));
      if (d == 2 && (l == 10 || l == 15)) xv = make_double2(fro > 0.0 ? sqrt(fro) : 1.0, 0.0);
      const bool ok = fro > 0.0 && polar4_newton_lanes(xv, l);
      if (ok) {
        const double2 uv = in ? xv : cz();
        // Re tr(U^H E) = sum of the singular values
        const double tr = sum16(in ? uv.x * es[l].x + uv.y * es[l].y : 0.0);
        gnew[l] = cconj(uv);
        if (l == 0) {
          s_nuc = tr;
          if (last_gate && a.omega_m1 == 0.0) a.overlap[(size_t)b * a.overlap_stride] = tr;
        }
      }
      if (l == 0) s_ok = ok ? 1 : 0;
    }
    __syncthreads();
    if (!s_ok || a.omega_m1 != 0.0) {
      // ill-conditioned E (Jacobi SVD, one lane), and the over-relaxed step
      if (threadIdx.x == 0) {
        double2 u[16];
        double nuc = s_nuc;
        if (!s_ok) {
          polar4_jacobi(d, es, u, &nuc);
#pragma unroll
          for (int k = 0; k < 16; ++k) u[k] = cconj(u[k]);
        } else {
#pragma unroll
          for (int k = 0; k < 16; ++k) u[k] = gnew[k];
        }
        if (a.omega_m1 != 0.0) nuc = relax_update(d, a.omega_m1, gm[j], u, es, rm);
#pragma unroll
        for (int k = 0; k < 16; ++k) gnew[k] = u[k];
        if (last_gate) a.overlap[(size_t)b * a.overlap_stride] = nuc;
      }
      __syncthreads();
    }
    if (threadIdx.x < 16) {
      int layer, site;
      step_gate(n, a.s0 + j, layer, site);
      size_t gi = ((size_t)b * a.lmax + layer) * n + site;
      a.gates[gi * 16 + threadIdx.x] = gnew[threadIdx.x];
    }
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
void launch_fwd_pass(const FwdArgs<S>& a, int nblk, int batch, hipStream_t st) {
  size_t shm = sizeof(S) << a.t;
  // the host fills a.mm under the same condition (fwd_mfma_applies)
  if (fwd_mfma_applies(a.mfma, a.t, a.n, sizeof(S))) {
    // the tile, then one 16 x 16 complex matrix per stage
    shm += (size_t)a.nstage * 256 * sizeof(S);
    if constexpr (sizeof(S) == 8) {
      hipLaunchKernelGGL(fwd_mfma_kernel, dim3(nblk, batch), dim3(512), shm, st, a);
      return;
    } else {
      // above the default 64 KiB of dynamic LDS; the limit is always set to
      // the kernel's largest need (tile + kMaxStages matrices), so that
      // launches from several host threads never lower it under each other
      const size_t most = (sizeof(S) << kFwdTileBits) + (size_t)kMaxStages * 256 * sizeof(S);
      if (hipFuncSetAttribute(reinterpret_cast<const void*>(fwd_mfma64_kernel),
                              hipFuncAttributeMaxDynamicSharedMemorySize, (int)most) == hipSuccess) {
        hipLaunchKernelGGL(fwd_mfma64_kernel, dim3(nblk, batch), dim3(512), shm, st, a);
        return;
      }
      shm = sizeof(S) << a.t;
    }
  }
  hipLaunchKernelGGL(fwd_pass_kernel<S>, dim3(nblk, batch), dim3(256), shm, st, a);
}
template void launch_fwd_pass<float2>(const FwdArgs<float2>&, int, int, hipStream_t);
template void launch_fwd_pass<double2>(const FwdArgs<double2>&, int, int, hipStream_t);

template <typename S>
void launch_window_pass(const WinPassArgs<S>& a, int nblk, int batch, hipStream_t st) {
  int t = a.t;
  size_t shm = 2 * (sizeof(S) << t);
  const size_t red = 4 * 512 * sizeof(double);
  if (shm < red) shm = red;
  constexpr int wpe = sizeof(S) == 8 ? 4 : 2;
  if (a.staged)
    hipLaunchKernelGGL((window_pass_kernel<S, wpe, true>), dim3(nblk, batch), dim3(256), shm, st, a);
  else
    hipLaunchKernelGGL((window_pass_kernel<S, wpe, false>), dim3(nblk, batch), dim3(256), shm, st, a);
}
template void launch_window_pass<float2>(const WinPassArgs<float2>&, int, int, hipStream_t);
template void launch_window_pass<double2>(const WinPassArgs<double2>&, int, int, hipStream_t);

void launch_window_finalize(const WinFinArgs& a, int batch, hipStream_t st) {
  hipLaunchKernelGGL(window_finalize_kernel, dim3(batch), dim3(256), 0, st, a);
}

// block k of state b sums partial slots [k group, min((k + 1) group, nblk))
// into slot k group (each thread owns elements e, e + 256 of every slot)
__global__ __launch_bounds__(256) void window_reduce_kernel(double* part, int nblk, int group, const int* list) {
  double* pb = part + (size_t)(list ? list[blockIdx.y] : blockIdx.y) * nblk * 512;
  const int q0 = blockIdx.x * group, q1 = min(q0 + group, nblk);
  for (int e = threadIdx.x; e < 512; e += blockDim.x) {
    double s = 0.0;
    int q = q0;
    for (; q + 3 < q1; q += 4)
      s += ((pb[(size_t)q * 512 + e] + pb[(size_t)(q + 1) * 512 + e]) +
            (pb[(size_t)(q + 2) * 512 + e] + pb[(size_t)(q + 3) * 512 + e]));
    for (; q < q1; ++q) s += pb[(size_t)q * 512 + e];
    pb[(size_t)q0 * 512 + e] = s;
  }
}

void launch_window_reduce(double* part, int nblk, int group, int batch, const int* list, hipStream_t st) {
  hipLaunchKernelGGL(window_reduce_kernel, dim3((nblk + group - 1) / group, batch), dim3(256), 0, st, part,
                     nblk, group, list);
}

}  // namespace qmx
