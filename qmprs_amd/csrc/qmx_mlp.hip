// Learned warm start: BF16 MLP training/inference kernels and su(4) maps.
//
// The network maps a feature row per two-qubit gate to 15 su(4) generator
// coefficients c (Gell-Mann basis, tr(l_a l_b) = 2 delta_ab); the proposed
// gate is G exp(i sum_k c_k l_k).  It has no counterpart in the reference: it
// proposes warm-start gates for the environment sweeps (sequential.py:
// 509-541), which decide the final fidelity.
//
// GEMMs run on v_mfma_f32_16x16x16bf16_1k (BF16 operands, FP32 accumulation):
//   A[i][k] lane i + 16 (k / 4), element k % 4
//   B[k][j] lane j + 16 (k / 4), element k % 4
//   D[i][j] lane j + 16 (i / 4), element i % 4
// su(4) log / exp maps run in FP64 with the 4x4 Hermitian Jacobi solver.
#include <algorithm>

#include "qmx_device.h"
#include "qmx_internal.h"

namespace qmx {

typedef short s4v __attribute__((ext_vector_type(4)));
typedef float f4v __attribute__((ext_vector_type(4)));

QMX_DEV uint16_t f2bf(float f) {
  uint32_t u = __float_as_uint(f);
  if ((u & 0x7fffffffu) > 0x7f800000u) return (uint16_t)((u >> 16) | 0x40u);   // NaN stays NaN
  u += 0x7fffu + ((u >> 16) & 1u);   // round to nearest even
  return (uint16_t)(u >> 16);
}
QMX_DEV float bf2f(uint16_t b) { return __uint_as_float((uint32_t)b << 16); }

// ------------------------------------------------------------- BF16 GEMM --
// C[m][n] = sum_k A(m, k) B(k, n); A(m, k) = transA ? A[k lda + m] : A[m lda + k],
// B(k, n) = transB ? B[n ldb + k] : B[k ldb + n].  Epilogue: + bias[n], then
// ReLU, or a mask (mask(m, n) > 0 keeps the value); stores FP32 and/or BF16.
constexpr int kGT = 64;    // tile edge
constexpr int kGK = 16;    // k step

__global__ __launch_bounds__(256) void mlp_gemm_kernel(MlpGemmArgs a) {
  __shared__ uint16_t As[kGT][kGK + 4];
  __shared__ uint16_t Bs[kGK][kGT + 4];
  const int m0 = blockIdx.y * kGT, n0 = blockIdx.x * kGT;
  const int lane = threadIdx.x & 63, wave = threadIdx.x >> 6;
  const int wm = (wave >> 1) * 32, wn = (wave & 1) * 32;
  f4v acc[2][2];
#pragma unroll
  for (int i = 0; i < 2; ++i)
#pragma unroll
    for (int j = 0; j < 2; ++j) acc[i][j] = f4v{0.f, 0.f, 0.f, 0.f};
  for (int k0 = 0; k0 < a.K; k0 += kGK) {
    // stage A (64 x 16) and B (16 x 64): 1024 values each, 4 per thread
#pragma unroll
    for (int r = 0; r < 4; ++r) {
      const int e = threadIdx.x + r * 256;
      {
        const int mm = e / kGK, kk = e % kGK;
        const int gm = m0 + mm, gk = k0 + kk;
        uint16_t v = 0;
        if (gm < a.M && gk < a.K) v = a.transA ? a.A[(size_t)gk * a.lda + gm] : a.A[(size_t)gm * a.lda + gk];
        As[mm][kk] = v;
      }
      {
        const int kk = e / kGT, nn = e % kGT;
        const int gk = k0 + kk, gn = n0 + nn;
        uint16_t v = 0;
        if (gk < a.K && gn < a.N) v = a.transB ? a.B[(size_t)gn * a.ldb + gk] : a.B[(size_t)gk * a.ldb + gn];
        Bs[kk][nn] = v;
      }
    }
    __syncthreads();
    const int li = lane & 15, kq = (lane >> 4) * 4;
#pragma unroll
    for (int ti = 0; ti < 2; ++ti) {
      s4v av;
#pragma unroll
      for (int e = 0; e < 4; ++e) av[e] = (short)As[wm + ti * 16 + li][kq + e];
#pragma unroll
      for (int tj = 0; tj < 2; ++tj) {
        s4v bv;
#pragma unroll
        for (int e = 0; e < 4; ++e) bv[e] = (short)Bs[kq + e][wn + tj * 16 + li];
        acc[ti][tj] = __builtin_amdgcn_mfma_f32_16x16x16bf16_1k(av, bv, acc[ti][tj], 0, 0, 0);
      }
    }
    __syncthreads();
  }
  const int j_in = lane & 15, i_base = (lane >> 4) * 4;
#pragma unroll
  for (int ti = 0; ti < 2; ++ti)
#pragma unroll
    for (int tj = 0; tj < 2; ++tj)
#pragma unroll
      for (int it = 0; it < 4; ++it) {
        const int m = m0 + wm + ti * 16 + i_base + it;
        const int nn = n0 + wn + tj * 16 + j_in;
        if (m >= a.M || nn >= a.N) continue;
        float v = acc[ti][tj][it];
        if (a.bias) v += a.bias[nn];
        if (a.relu) v = fmaxf(v, 0.f);
        if (a.mask && !(bf2f(a.mask[(size_t)m * a.ldm + nn]) > 0.f)) v = 0.f;
        if (a.C) a.C[(size_t)m * a.ldc + nn] = v;
        if (a.Cb) a.Cb[(size_t)m * a.ldcb + nn] = f2bf(v);
      }
}

// db[n] = sum_m float(dz[m][n]) in a fixed order (one thread per column)
__global__ void colsum_bf16_kernel(const uint16_t* dz, int M, int N, int ld, float* db) {
  const int nn = blockIdx.x * blockDim.x + threadIdx.x;
  if (nn >= N) return;
  float s = 0.f;
  for (int m = 0; m < M; ++m) s += bf2f(dz[(size_t)m * ld + nn]);
  db[nn] = s;
}

__global__ void to_bf16_kernel(const float* x, uint16_t* y, size_t n) {
  for (size_t i = blockIdx.x * (size_t)blockDim.x + threadIdx.x; i < n; i += (size_t)gridDim.x * blockDim.x)
    y[i] = f2bf(x[i]);
}

// Adam (bias-corrected); master FP32 parameters and their BF16 shadow
__global__ void adam_kernel(float* p, uint16_t* pb, const float* g, float* m, float* v, size_t n,
                            float lr, float b1, float b2, float eps, float c1, float c2) {
  for (size_t i = blockIdx.x * (size_t)blockDim.x + threadIdx.x; i < n; i += (size_t)gridDim.x * blockDim.x) {
    const float gi = g[i];
    const float mi = b1 * m[i] + (1.f - b1) * gi;
    const float vi = b2 * v[i] + (1.f - b2) * gi * gi;
    m[i] = mi;
    v[i] = vi;
    const float pi = p[i] - lr * (mi / c1) / (sqrtf(vi / c2) + eps);
    p[i] = pi;
    if (pb) pb[i] = f2bf(pi);
  }
}

// Weighted MSE over the first `valid` of `ld` output columns:
// loss = sum_r w_r sum_c (y - t)^2 / (sum_r w_r * valid); dy = d loss / d y.
// Block partials of the loss go to part[blockIdx.x].
__global__ __launch_bounds__(256) void mse_kernel(const float* y, const float* t, const float* w, int rows,
                                                  int ld, int valid, float inv_norm, float* dy,
                                                  float* part) {
  __shared__ float red[256];
  float acc = 0.f;
  for (int r = blockIdx.x * blockDim.x + threadIdx.x; r < rows; r += gridDim.x * blockDim.x) {
    const float wr = w ? w[r] : 1.f;
    for (int c = 0; c < ld; ++c) {
      const size_t i = (size_t)r * ld + c;
      const float d = (c < valid) ? (y[i] - t[i]) : 0.f;
      acc += wr * d * d;
      dy[i] = 2.f * wr * d * inv_norm;
    }
  }
  red[threadIdx.x] = acc;
  __syncthreads();
  for (int s = 128; s > 0; s >>= 1) {
    if ((int)threadIdx.x < s) red[threadIdx.x] += red[threadIdx.x + s];
    __syncthreads();
  }
  if (threadIdx.x == 0) part[blockIdx.x] = red[0] * inv_norm;
}

__global__ void sum_kernel(const float* part, int n, float* out) {
  if (threadIdx.x != 0 || blockIdx.x != 0) return;
  float s = 0.f;
  for (int i = 0; i < n; ++i) s += part[i];
  *out = s;
}

// ------------------------------------------------------------ su(4) maps --
// Gell-Mann matrices l_0..l_14 of su(4): for j < k the symmetric (E_jk + E_kj)
// and antisymmetric (-i E_jk + i E_kj) pairs in (0,1), (0,2), (0,3), (1,2),
// (1,3), (2,3) order, then the diagonal ones sqrt(2 / (l (l + 1))) diag(1..1,
// -l, 0..) for l = 1, 2, 3.
QMX_DEV void gm_pair(int k, int& j, int& l) {
  const int pj[6] = {0, 0, 0, 1, 1, 2}, pl[6] = {1, 2, 3, 2, 3, 3};
  j = pj[k];
  l = pl[k];
}

// H (4x4 Hermitian) -> coefficients c_k = Re tr(H l_k) / 2
QMX_DEV void su4_coeffs(const double2* h, double* c) {
#pragma unroll
  for (int p = 0; p < 6; ++p) {
    int j, l;
    gm_pair(p, j, l);
    const double2 hjl = h[j * 4 + l], hlj = h[l * 4 + j];
    c[2 * p] = 0.5 * (hlj.x + hjl.x);                 // tr(H (E_jl + E_lj)) / 2
    c[2 * p + 1] = 0.5 * (hlj.y - hjl.y);             // Re tr(H (-i E_jl + i E_lj)) / 2 = -Im H_jl
  }
#pragma unroll
  for (int l = 1; l < 4; ++l) {
    const double s = sqrt(2.0 / (l * (l + 1.0)));
    double t = 0.0;
#pragma unroll
    for (int j = 0; j < l; ++j) t += h[j * 5].x;
    t -= l * h[l * 5].x;
    c[12 + l - 1] = 0.5 * s * t;
  }
}

// coefficients -> H = sum_k c_k l_k
QMX_DEV void su4_matrix(const float* c, double2* h) {
#pragma unroll
  for (int i = 0; i < 16; ++i) h[i] = cz();
#pragma unroll
  for (int p = 0; p < 6; ++p) {
    int j, l;
    gm_pair(p, j, l);
    const double a = c[2 * p], b = c[2 * p + 1];
    h[j * 4 + l] = make_double2(a, -b);
    h[l * 4 + j] = make_double2(a, b);
  }
#pragma unroll
  for (int l = 1; l < 4; ++l) {
    const double s = c[12 + l - 1] * sqrt(2.0 / (l * (l + 1.0)));
#pragma unroll
    for (int j = 0; j < l; ++j) h[j * 5].x += s;
    h[l * 5].x -= l * s;
  }
}

// principal Hermitian logarithm: U = exp(i H), eigenphases in (-pi, pi],
// traceless part returned.  U is normal, so the Hermitian combination
// (U + U^H) / 2 + x (U - U^H) / (2i) with an irrational x shares its
// eigenvectors and separates distinct eigenvalues.
QMX_DEV void su4_log(const double2* u, double2* h) {
  double2 m[16], v[16];
  double lam[4];
  const double x = 0.6180339887498949;
#pragma unroll
  for (int r = 0; r < 4; ++r)
#pragma unroll
    for (int c = 0; c < 4; ++c) {
      const double2 a = u[r * 4 + c], b = cconj(u[c * 4 + r]);   // U, U^H entries
      // (a + b) / 2 + x (a - b) / (2i) = (a + b) / 2 - i x (a - b) / 2
      const double2 s = make_double2(0.5 * (a.x + b.x), 0.5 * (a.y + b.y));
      const double2 d = make_double2(0.5 * (a.x - b.x), 0.5 * (a.y - b.y));
      m[r * 4 + c] = make_double2(s.x + x * d.y, s.y - x * d.x);
    }
  herm_eig4(m, v, lam);
  double th[4];
#pragma unroll
  for (int k = 0; k < 4; ++k) {
    double2 z = cz();   // v_k^H U v_k
#pragma unroll
    for (int r = 0; r < 4; ++r) {
      double2 uv = cz();
#pragma unroll
      for (int c = 0; c < 4; ++c) uv = cfma(u[r * 4 + c], v[c * 4 + k], uv);
      z = cfmac(v[r * 4 + k], uv, z);
    }
    th[k] = atan2(z.y, z.x);
  }
  double tr = 0.25 * (th[0] + th[1] + th[2] + th[3]);
#pragma unroll
  for (int r = 0; r < 4; ++r)
#pragma unroll
    for (int c = 0; c < 4; ++c) {
      double2 s = cz();
#pragma unroll
      for (int k = 0; k < 4; ++k) s = cfma(cscale(v[r * 4 + k], th[k] - tr), cconj(v[c * 4 + k]), s);
      h[r * 4 + c] = s;
    }
}

// exp(i H) for Hermitian H
QMX_DEV void su4_exp(const double2* h, double2* e) {
  double2 m[16], v[16];
  double lam[4];
#pragma unroll
  for (int i = 0; i < 16; ++i) m[i] = h[i];
  herm_eig4(m, v, lam);
#pragma unroll
  for (int r = 0; r < 4; ++r)
#pragma unroll
    for (int c = 0; c < 4; ++c) {
      double2 s = cz();
#pragma unroll
      for (int k = 0; k < 4; ++k) {
        double sn, cs;
        sincos(lam[k], &sn, &cs);
        s = cfma(cmul(v[r * 4 + k], make_double2(cs, sn)), cconj(v[c * 4 + k]), s);
      }
      e[r * 4 + c] = s;
    }
}

QMX_DEV bool gate_row(const GateRowArgs& a, int row, size_t& gi) {
  const int per = a.n - 1;
  const int b = row / (a.lmax * per);
  const int rem = row % (a.lmax * per);
  const int layer = rem / per, site = rem % per;
  gi = ((size_t)b * a.lmax + layer) * a.n + site;
  return layer < a.layers[b] && a.widths[gi] == 2;
}

// features of each two-qubit gate slot (zero rows for empty slots):
// [c(Ga^H Gb) (15), Re/Im Gb (32), layer / lmax, site / n, 1, 0 ...]
__global__ __launch_bounds__(64) void gate_features_kernel(GateRowArgs a, const double2* ga, const double2* gb, uint16_t* feat,
                                     int ld, float* weight) {
  const int row = blockIdx.x * blockDim.x + threadIdx.x;
  if (row >= a.rows) return;
  size_t gi;
  const bool ok = gate_row(a, row, gi);
  uint16_t* f = feat + (size_t)row * ld;
  for (int c = 0; c < ld; ++c) f[c] = 0;
  if (weight) weight[row] = ok ? 1.f : 0.f;
  if (!ok) return;
  double2 u[16], h[16];
  const double2* A = ga + gi * 16;
  const double2* B = gb + gi * 16;
#pragma unroll
  for (int r = 0; r < 4; ++r)
#pragma unroll
    for (int c = 0; c < 4; ++c) {
      double2 s = cz();
#pragma unroll
      for (int k = 0; k < 4; ++k) s = cfmac(A[k * 4 + r], B[k * 4 + c], s);
      u[r * 4 + c] = s;
    }
  su4_log(u, h);
  double cf[15];
  su4_coeffs(h, cf);
#pragma unroll
  for (int k = 0; k < 15; ++k) f[k] = f2bf((float)cf[k]);
#pragma unroll
  for (int e = 0; e < 16; ++e) {
    f[15 + 2 * e] = f2bf((float)B[e].x);
    f[16 + 2 * e] = f2bf((float)B[e].y);
  }
  const int per = a.n - 1;
  const int rem = row % (a.lmax * per);
  f[47] = f2bf((float)(rem / per) / (float)a.lmax);
  f[48] = f2bf((float)(rem % per) / (float)a.n);
  f[49] = f2bf(1.f);
}

// labels: c(Ga^H Gb) as FP32 rows of width ld (zeros beyond 15)
// ---- trajectory extrapolation (learned warm start) ----------------------
// su(4) coefficients c(A^H B) of the relative rotation of one gate slot
QMX_DEV void slot_log(const double2* A, const double2* B, double* cf) {
  double2 u[16], h[16];
#pragma unroll
  for (int r = 0; r < 4; ++r)
#pragma unroll
    for (int c = 0; c < 4; ++c) {
      double2 s = cz();
#pragma unroll
      for (int k = 0; k < 4; ++k) s = cfmac(A[k * 4 + r], B[k * 4 + c], s);
      u[r * 4 + c] = s;
    }
  su4_log(u, h);
  su4_coeffs(h, cf);
}

// Trajectory features of each two-qubit gate slot at sweep s, from the gates
// after sweeps s-2 (a), s-1 (b) and s (c): the last two updates d1 = c(Ga^H
// Gb), d2 = c(Gb^H Gc) in units of |d2| (the row scale, floored at 1e-12):
// [d1 / scale (15), d2 / scale (15), log10(scale) / 8, sweep_frac, layer /
// lmax, site / n, 1, 0 ...]
__global__ __launch_bounds__(64) void traj_features_kernel(GateRowArgs a, const double2* ga, const double2* gb,
                                                           const double2* gc, uint16_t* feat, int ld,
                                                           float* weight, float* scale, float sweep_frac) {
  const int row = blockIdx.x * blockDim.x + threadIdx.x;
  if (row >= a.rows) return;
  size_t gi;
  const bool ok = gate_row(a, row, gi);
  uint16_t* f = feat + (size_t)row * ld;
  for (int c = 0; c < ld; ++c) f[c] = 0;
  if (weight) weight[row] = ok ? 1.f : 0.f;
  if (scale) scale[row] = 1.f;
  if (!ok) return;
  double d1[15], d2[15];
  slot_log(ga + gi * 16, gb + gi * 16, d1);
  slot_log(gb + gi * 16, gc + gi * 16, d2);
  double nrm = 0.0;
#pragma unroll
  for (int k = 0; k < 15; ++k) nrm += d2[k] * d2[k];
  const double sc = fmax(sqrt(nrm), 1e-12);
  const float scf = (float)sc;
  if (scale) scale[row] = scf;
#pragma unroll
  for (int k = 0; k < 15; ++k) {
    f[k] = f2bf((float)(d1[k] / (double)scf));
    f[15 + k] = f2bf((float)(d2[k] / (double)scf));
  }
  const int per = a.n - 1;
  const int rem = row % (a.lmax * per);
  f[30] = f2bf(log10f(scf) / 8.f);
  f[31] = f2bf(sweep_frac);
  f[32] = f2bf((float)(rem / per) / (float)a.lmax);
  f[33] = f2bf((float)(rem % per) / (float)a.n);
  f[34] = f2bf(1.f);
}

// Orthonormal basis of the span of a slot's last two updates (in units of
// |d2|): e2 = d2 / |d2|, e1 = the part of d1 orthogonal to e2, normalised
// (zero when d1 is parallel to d2 to 1e-6).  Returns |d2| (floor 1e-12).
QMX_DEV double traj_basis(const double2* ga, const double2* gb, const double2* gc, double* e1, double* e2) {
  double d1[15], d2[15];
  slot_log(ga, gb, d1);
  slot_log(gb, gc, d2);
  double n2 = 0.0;
#pragma unroll
  for (int k = 0; k < 15; ++k) n2 += d2[k] * d2[k];
  n2 = sqrt(n2);
  const double sc = fmax(n2, 1e-12);
  double dot = 0.0;
#pragma unroll
  for (int k = 0; k < 15; ++k) {
    e2[k] = n2 > 0.0 ? d2[k] / n2 : 0.0;
    dot += d1[k] * e2[k];
  }
  double n1 = 0.0;
#pragma unroll
  for (int k = 0; k < 15; ++k) {
    e1[k] = d1[k] - dot * e2[k];
    n1 += e1[k] * e1[k];
  }
  n1 = sqrt(n1);
  const bool keep = n1 > 1e-6 * sc;
#pragma unroll
  for (int k = 0; k < 15; ++k) e1[k] = keep ? e1[k] / n1 : 0.0;
  return sc;
}

// targets: the update c(Gc^H Gd) in units of |d2|, projected on (e2, e1):
// [t.e2, t.e1, 0 ...] (FP32, zero-padded to ld)
__global__ __launch_bounds__(64) void traj_target_kernel(GateRowArgs a, const double2* ga, const double2* gb,
                                                         const double2* gc, const double2* gd, float* out, int ld) {
  const int row = blockIdx.x * blockDim.x + threadIdx.x;
  if (row >= a.rows) return;
  size_t gi;
  const bool ok = gate_row(a, row, gi);
  float* o = out + (size_t)row * ld;
  for (int c = 0; c < ld; ++c) o[c] = 0.f;
  if (!ok) return;
  double e1[15], e2[15], t[15];
  const double sc = traj_basis(ga + gi * 16, gb + gi * 16, gc + gi * 16, e1, e2);
  slot_log(gc + gi * 16, gd + gi * 16, t);
  double p2 = 0.0, p1 = 0.0;
#pragma unroll
  for (int k = 0; k < 15; ++k) {
    p2 += t[k] * e2[k];
    p1 += t[k] * e1[k];
  }
  o[0] = (float)(p2 / sc);
  o[1] = (float)(p1 / sc);
}

// jump: G <- G exp(i |d2| sum_k (coef0 e2_k + coef1 e1_k) l_k) on live rows of
// states with skip == 0, the basis taken from (Ga, Gb, G)
__global__ __launch_bounds__(64) void traj_jump_kernel(GateRowArgs a, const double2* ga, const double2* gb,
                                                       const float* coef, int ld, double2* g, const int* skip) {
  const int row = blockIdx.x * blockDim.x + threadIdx.x;
  if (row >= a.rows) return;
  size_t gi;
  if (!gate_row(a, row, gi)) return;
  const int b = row / (a.lmax * (a.n - 1));
  if (skip && skip[b]) return;
  double2* G = g + gi * 16;
  double e1[15], e2[15];
  const double sc = traj_basis(ga + gi * 16, gb + gi * 16, G, e1, e2);
  const double c0 = coef[(size_t)row * ld], c1 = coef[(size_t)row * ld + 1];
  float cs[15];
#pragma unroll
  for (int k = 0; k < 15; ++k) cs[k] = (float)(sc * (c0 * e2[k] + c1 * e1[k]));
  double2 h[16], e[16], out[16];
  su4_matrix(cs, h);
  su4_exp(h, e);
#pragma unroll
  for (int r = 0; r < 4; ++r)
#pragma unroll
    for (int c = 0; c < 4; ++c) {
      double2 s = cz();
#pragma unroll
      for (int k = 0; k < 4; ++k) s = cfma(G[r * 4 + k], e[k * 4 + c], s);
      out[r * 4 + c] = s;
    }
#pragma unroll
  for (int i = 0; i < 16; ++i) G[i] = out[i];
}

__global__ __launch_bounds__(64) void gate_log_kernel(GateRowArgs a, const double2* ga, const double2* gb, float* out, int ld) {
  const int row = blockIdx.x * blockDim.x + threadIdx.x;
  if (row >= a.rows) return;
  size_t gi;
  const bool ok = gate_row(a, row, gi);
  float* o = out + (size_t)row * ld;
  for (int c = 0; c < ld; ++c) o[c] = 0.f;
  if (!ok) return;
  double2 u[16], h[16];
  const double2* A = ga + gi * 16;
  const double2* B = gb + gi * 16;
#pragma unroll
  for (int r = 0; r < 4; ++r)
#pragma unroll
    for (int c = 0; c < 4; ++c) {
      double2 s = cz();
#pragma unroll
      for (int k = 0; k < 4; ++k) s = cfmac(A[k * 4 + r], B[k * 4 + c], s);
      u[r * 4 + c] = s;
    }
  su4_log(u, h);
  double cf[15];
  su4_coeffs(h, cf);
#pragma unroll
  for (int k = 0; k < 15; ++k) o[k] = (float)cf[k];
}

// G <- G exp(i sum_k c_k l_k) on every live two-qubit slot; states with
// skip[b] set are left alone
__global__ __launch_bounds__(64) void gate_expmap_kernel(GateRowArgs a, const float* coef, int ld, double2* g, const int* skip) {
  const int row = blockIdx.x * blockDim.x + threadIdx.x;
  if (row >= a.rows) return;
  size_t gi;
  if (!gate_row(a, row, gi)) return;
  const int b = row / (a.lmax * (a.n - 1));
  if (skip && skip[b]) return;
  double2 h[16], e[16], out[16];
  su4_matrix(coef + (size_t)row * ld, h);
  su4_exp(h, e);
  double2* G = g + gi * 16;
#pragma unroll
  for (int r = 0; r < 4; ++r)
#pragma unroll
    for (int c = 0; c < 4; ++c) {
      double2 s = cz();
#pragma unroll
      for (int k = 0; k < 4; ++k) s = cfma(G[r * 4 + k], e[k * 4 + c], s);
      out[r * 4 + c] = s;
    }
#pragma unroll
  for (int i = 0; i < 16; ++i) G[i] = out[i];
}

// -------------------------------------------------------------- launchers --
void launch_mlp_gemm(const MlpGemmArgs& a, hipStream_t st) {
  dim3 grid((a.N + kGT - 1) / kGT, (a.M + kGT - 1) / kGT);
  hipLaunchKernelGGL(mlp_gemm_kernel, grid, dim3(256), 0, st, a);
}
void launch_colsum_bf16(const uint16_t* dz, int M, int N, int ld, float* db, hipStream_t st) {
  hipLaunchKernelGGL(colsum_bf16_kernel, dim3((N + 255) / 256), dim3(256), 0, st, dz, M, N, ld, db);
}
void launch_to_bf16(const float* x, uint16_t* y, size_t n, hipStream_t st) {
  int blocks = (int)std::min<size_t>((n + 255) / 256, 4096);
  hipLaunchKernelGGL(to_bf16_kernel, dim3(blocks), dim3(256), 0, st, x, y, n);
}
void launch_adam(float* p, uint16_t* pb, const float* g, float* m, float* v, size_t n, float lr, float b1,
                 float b2, float eps, int step, hipStream_t st) {
  const float c1 = 1.f - powf(b1, (float)step), c2 = 1.f - powf(b2, (float)step);
  int blocks = (int)std::min<size_t>((n + 255) / 256, 4096);
  hipLaunchKernelGGL(adam_kernel, dim3(blocks), dim3(256), 0, st, p, pb, g, m, v, n, lr, b1, b2, eps, c1, c2);
}
void launch_mse(const float* y, const float* t, const float* w, int rows, int ld, int valid, float inv_norm,
                float* dy, float* part, int nblk, float* loss, hipStream_t st) {
  hipLaunchKernelGGL(mse_kernel, dim3(nblk), dim3(256), 0, st, y, t, w, rows, ld, valid, inv_norm, dy, part);
  hipLaunchKernelGGL(sum_kernel, dim3(1), dim3(64), 0, st, part, nblk, loss);
}
void launch_gate_features(const GateRowArgs& a, const double2* ga, const double2* gb, uint16_t* feat, int ld,
                          float* weight, hipStream_t st) {
  hipLaunchKernelGGL(gate_features_kernel, dim3((a.rows + 63) / 64), dim3(64), 0, st, a, ga, gb, feat, ld,
                     weight);
}
void launch_gate_log(const GateRowArgs& a, const double2* ga, const double2* gb, float* out, int ld,
                     hipStream_t st) {
  hipLaunchKernelGGL(gate_log_kernel, dim3((a.rows + 63) / 64), dim3(64), 0, st, a, ga, gb, out, ld);
}
void launch_traj_features(const GateRowArgs& a, const double2* ga, const double2* gb, const double2* gc,
                          uint16_t* feat, int ld, float* weight, float* scale, float sweep_frac, hipStream_t st) {
  hipLaunchKernelGGL(traj_features_kernel, dim3((a.rows + 63) / 64), dim3(64), 0, st, a, ga, gb, gc, feat, ld,
                     weight, scale, sweep_frac);
}
void launch_traj_targets(const GateRowArgs& a, const double2* ga, const double2* gb, const double2* gc,
                         const double2* gd, float* out, int ld, hipStream_t st) {
  hipLaunchKernelGGL(traj_target_kernel, dim3((a.rows + 63) / 64), dim3(64), 0, st, a, ga, gb, gc, gd, out, ld);
}
void launch_traj_jump(const GateRowArgs& a, const double2* ga, const double2* gb, const float* coef, int ld,
                      double2* g, const int* skip, hipStream_t st) {
  hipLaunchKernelGGL(traj_jump_kernel, dim3((a.rows + 63) / 64), dim3(64), 0, st, a, ga, gb, coef, ld, g, skip);
}
void launch_gate_expmap(const GateRowArgs& a, const float* coef, int ld, double2* g, const int* skip,
                        hipStream_t st) {
  hipLaunchKernelGGL(gate_expmap_kernel, dim3((a.rows + 63) / 64), dim3(64), 0, st, a, coef, ld, g, skip);
}

}  // namespace qmx
