// Device-side helpers shared by the qmx kernels (gfx950, wave64).
//
// Complex numbers are double2 (re, im) in registers and LDS; dense states are
// stored either as float2 (complex64) or double2 (complex128) in HBM.
// All small linear algebra here (4x4 Hermitian eigensolver, 4x4 polar factor,
// Householder completion) runs serially in one lane in FP64: the matrices are
// tiny and the results feed every later step, so accuracy wins over width.
#pragma once
#include <hip/hip_runtime.h>
#include <stdint.h>

#define QMX_DEV __device__ __forceinline__

// ---------------------------------------------------------------- complex --
QMX_DEV double2 cz() { return make_double2(0.0, 0.0); }
QMX_DEV double2 cadd(double2 a, double2 b) { return make_double2(a.x + b.x, a.y + b.y); }
QMX_DEV double2 csub(double2 a, double2 b) { return make_double2(a.x - b.x, a.y - b.y); }
QMX_DEV double2 cmul(double2 a, double2 b) {
  return make_double2(fma(a.x, b.x, -a.y * b.y), fma(a.x, b.y, a.y * b.x));
}
// conj(a) * b
QMX_DEV double2 cmulc(double2 a, double2 b) {
  return make_double2(fma(a.x, b.x, a.y * b.y), fma(a.x, b.y, -a.y * b.x));
}
// acc + a * b
QMX_DEV double2 cfma(double2 a, double2 b, double2 acc) {
  acc.x = fma(a.x, b.x, acc.x);
  acc.x = fma(-a.y, b.y, acc.x);
  acc.y = fma(a.x, b.y, acc.y);
  acc.y = fma(a.y, b.x, acc.y);
  return acc;
}
// acc + conj(a) * b
QMX_DEV double2 cfmac(double2 a, double2 b, double2 acc) {
  acc.x = fma(a.x, b.x, acc.x);
  acc.x = fma(a.y, b.y, acc.x);
  acc.y = fma(a.x, b.y, acc.y);
  acc.y = fma(-a.y, b.x, acc.y);
  return acc;
}
QMX_DEV double2 cconj(double2 a) { return make_double2(a.x, -a.y); }
QMX_DEV double2 cscale(double2 a, double s) { return make_double2(a.x * s, a.y * s); }
QMX_DEV double cabs2(double2 a) { return fma(a.x, a.x, a.y * a.y); }
QMX_DEV double cabs_(double2 a) { return sqrt(cabs2(a)); }

// storage <-> compute conversion
QMX_DEV double2 to_c(float2 v) { return make_double2((double)v.x, (double)v.y); }
QMX_DEV double2 to_c(double2 v) { return v; }
template <typename S> QMX_DEV S from_c(double2 v);
template <> QMX_DEV float2 from_c<float2>(double2 v) { return make_float2((float)v.x, (float)v.y); }
template <> QMX_DEV double2 from_c<double2>(double2 v) { return v; }

// ------------------------------------------------------------- reductions --
QMX_DEV double wave_sum(double v) {
#pragma unroll
  for (int off = 32; off > 0; off >>= 1) v += __shfl_xor(v, off, 64);
  return v;
}

// ------------------------------------------------- 4x4 Hermitian Jacobi ----
// Cyclic two-sided Jacobi on an n x n (n <= 4) Hermitian matrix a (row-major,
// stride 4).  On return lam[] holds eigenvalues sorted in decreasing order and
// the columns of v (stride 4) the matching orthonormal eigenvectors.
QMX_DEV void herm_eig4(int n, double2* a, double2* v, double* lam) {
  for (int i = 0; i < 4; ++i)
    for (int j = 0; j < 4; ++j) v[i * 4 + j] = make_double2(i == j ? 1.0 : 0.0, 0.0);
  for (int sweep = 0; sweep < 40; ++sweep) {
    double off = 0.0, diag = 0.0;
    for (int i = 0; i < n; ++i) {
      diag += a[i * 4 + i].x * a[i * 4 + i].x;
      for (int j = i + 1; j < n; ++j) off += cabs2(a[i * 4 + j]);
    }
    if (off <= 1e-32 * diag || off == 0.0) break;
    for (int p = 0; p < n - 1; ++p) {
      for (int q = p + 1; q < n; ++q) {
        double2 g = a[p * 4 + q];
        double ag = cabs_(g);
        if (ag == 0.0) continue;
        double al = a[p * 4 + p].x, be = a[q * 4 + q].x;
        double2 ph = make_double2(g.x / ag, g.y / ag);  // e^{i phi}
        double tau = (be - al) / (2.0 * ag);
        double t = (tau >= 0 ? 1.0 : -1.0) / (fabs(tau) + sqrt(1.0 + tau * tau));
        double c = 1.0 / sqrt(1.0 + t * t), s = t * c;
        // J = [[c, s], [-s e^{-i phi}, c e^{-i phi}]] acting on columns (p, q)
        double2 em = cconj(ph);
        double2 j00 = make_double2(c, 0), j01 = make_double2(s, 0);
        double2 j10 = cscale(em, -s), j11 = cscale(em, c);
        // A <- A J (columns)
        for (int i = 0; i < n; ++i) {
          double2 x = a[i * 4 + p], y = a[i * 4 + q];
          a[i * 4 + p] = cadd(cmul(x, j00), cmul(y, j10));
          a[i * 4 + q] = cadd(cmul(x, j01), cmul(y, j11));
          double2 vx = v[i * 4 + p], vy = v[i * 4 + q];
          v[i * 4 + p] = cadd(cmul(vx, j00), cmul(vy, j10));
          v[i * 4 + q] = cadd(cmul(vx, j01), cmul(vy, j11));
        }
        // A <- J^H A (rows)
        for (int j = 0; j < n; ++j) {
          double2 x = a[p * 4 + j], y = a[q * 4 + j];
          a[p * 4 + j] = cadd(cmulc(j00, x), cmulc(j10, y));
          a[q * 4 + j] = cadd(cmulc(j01, x), cmulc(j11, y));
        }
        a[p * 4 + q] = cz();
        a[q * 4 + p] = cz();
        a[p * 4 + p].y = 0.0;
        a[q * 4 + q].y = 0.0;
      }
    }
  }
  for (int i = 0; i < n; ++i) lam[i] = a[i * 4 + i].x;
  // selection sort, decreasing, stable on ties
  for (int i = 0; i < n; ++i) {
    int m = i;
    for (int j = i + 1; j < n; ++j)
      if (lam[j] > lam[m]) m = j;
    if (m != i) {
      double tl = lam[i]; lam[i] = lam[m]; lam[m] = tl;
      for (int r = 0; r < 4; ++r) {
        double2 tv = v[r * 4 + i]; v[r * 4 + i] = v[r * 4 + m]; v[r * 4 + m] = tv;
      }
    }
  }
}

// rsum2 cutoff + max_bond keep count (decreasing eigenvalues lam[0..n-1]).
QMX_DEV int keep_count(const double* lam, int n, int max_bond, double cutoff) {
  double total = 0.0;
  for (int i = 0; i < n; ++i) total += fmax(lam[i], 0.0);
  int k = n;
  if (total > 0.0 && cutoff > 0.0) {
    double tail = 0.0;
    for (int j = n - 1; j > 0; --j) {
      tail += fmax(lam[j], 0.0);
      if (tail <= cutoff * total) k = j; else break;
    }
  }
  if (max_bond > 0 && k > max_bond) k = max_bond;
  return k < 1 ? 1 : k;
}

// Canonical orthonormal basis of span(v[:, 0:k]) (columns, stride 4):
// two-pass Gram-Schmidt over the columns of the projector P = V V^H in index
// order, skipping residuals with norm <= tol.  Writes w (n x k, stride 4).
// Returns the number of basis vectors found (== k unless degenerate input).
QMX_DEV int canonical_basis4(int n, int k, const double2* v, double2* w, double tol) {
  double2 p[16];
  for (int i = 0; i < n; ++i)
    for (int j = 0; j < n; ++j) {
      double2 s = cz();
      for (int m = 0; m < k; ++m) s = cfma(v[i * 4 + m], cconj(v[j * 4 + m]), s);
      p[i * 4 + j] = s;
    }
  int found = 0;
  for (int j = 0; j < n && found < k; ++j) {
    double2 x[4];
    for (int i = 0; i < n; ++i) x[i] = p[i * 4 + j];
    for (int pass = 0; pass < 2; ++pass) {
      for (int b = 0; b < found; ++b) {
        double2 d = cz();
        for (int i = 0; i < n; ++i) d = cfmac(w[i * 4 + b], x[i], d);
        for (int i = 0; i < n; ++i) x[i] = csub(x[i], cmul(w[i * 4 + b], d));
      }
    }
    double nrm = 0.0;
    for (int i = 0; i < n; ++i) nrm += cabs2(x[i]);
    nrm = sqrt(nrm);
    if (nrm > tol) {
      for (int i = 0; i < n; ++i) w[i * 4 + found] = cscale(x[i], 1.0 / nrm);
      ++found;
    }
  }
  return found;
}

// ------------------------------------------------------- polar (d <= 4) ----
// Unitary polar factor U V^H of the d x d matrix e (row-major, stride 4) via a
// one-sided (Hestenes) Jacobi SVD.  Singular pairs below 1e-13 of the largest
// are replaced by a canonical completion (see below).  Returns the sum of
// singular values (the nuclear norm) in *nuc.
QMX_DEV void polar4(int d, const double2* e, double2* out, double* nuc) {
  double2 a[16], v[16];
  for (int i = 0; i < 16; ++i) { a[i] = e[i]; v[i] = cz(); }
  for (int i = 0; i < 4; ++i) v[i * 4 + i] = make_double2(1.0, 0.0);
  for (int sweep = 0; sweep < 60; ++sweep) {
    bool rotated = false;
    for (int p = 0; p < d - 1; ++p)
      for (int q = p + 1; q < d; ++q) {
        double al = 0.0, be = 0.0;
        double2 g = cz();
        for (int i = 0; i < d; ++i) {
          al += cabs2(a[i * 4 + p]);
          be += cabs2(a[i * 4 + q]);
          g = cfmac(a[i * 4 + p], a[i * 4 + q], g);  // a_p^H a_q
        }
        double ag = cabs_(g);
        if (ag <= 1e-16 * sqrt(al * be) || ag == 0.0) continue;
        rotated = true;
        double2 ph = make_double2(g.x / ag, g.y / ag);
        double zeta = (be - al) / (2.0 * ag);
        double t = (zeta >= 0 ? 1.0 : -1.0) / (fabs(zeta) + sqrt(1.0 + zeta * zeta));
        double c = 1.0 / sqrt(1.0 + t * t), s = t * c;
        // [a_p, a_q] <- [a_p, a_q] [[c, s e^{i phi}], [-s e^{-i phi}, c]]
        double2 sp = cscale(ph, s), sm = cscale(cconj(ph), -s);
        for (int i = 0; i < d; ++i) {
          double2 x = a[i * 4 + p], y = a[i * 4 + q];
          a[i * 4 + p] = cadd(cscale(x, c), cmul(y, sm));
          a[i * 4 + q] = cadd(cmul(x, sp), cscale(y, c));
          double2 vx = v[i * 4 + p], vy = v[i * 4 + q];
          v[i * 4 + p] = cadd(cscale(vx, c), cmul(vy, sm));
          v[i * 4 + q] = cadd(cmul(vx, sp), cscale(vy, c));
        }
      }
    if (!rotated) break;
  }
  double sig[4], smax = 0.0, ssum = 0.0;
  for (int j = 0; j < d; ++j) {
    double s2 = 0.0;
    for (int i = 0; i < d; ++i) s2 += cabs2(a[i * 4 + j]);
    sig[j] = sqrt(s2);
    smax = fmax(smax, sig[j]);
    ssum += sig[j];
  }
  *nuc = ssum;
  // well-determined pairs (u_j, v_j); the rest are completed canonically:
  // both sides by two-pass Gram-Schmidt of e_0, e_1, ... in order, paired in
  // the order found.  The factor then depends only on the kept subspaces.
  double2 us[16], vs[16];
  int r = 0;
  for (int j = 0; j < d; ++j) {
    if (smax > 0.0 && sig[j] > 1e-13 * smax) {
      for (int i = 0; i < d; ++i) {
        us[i * 4 + r] = cscale(a[i * 4 + j], 1.0 / sig[j]);
        vs[i * 4 + r] = v[i * 4 + j];
      }
      ++r;
    }
  }
  int ru = r, rv = r;
  for (int side = 0; side < 2; ++side) {
    double2* basis = side == 0 ? us : vs;
    int& cnt = side == 0 ? ru : rv;
    for (int e0 = 0; e0 < d && cnt < d; ++e0) {
      double2 x[4];
      for (int i = 0; i < d; ++i) x[i] = make_double2(i == e0 ? 1.0 : 0.0, 0.0);
      for (int pass = 0; pass < 2; ++pass)
        for (int b = 0; b < cnt; ++b) {
          double2 dd = cz();
          for (int i = 0; i < d; ++i) dd = cfmac(basis[i * 4 + b], x[i], dd);
          for (int i = 0; i < d; ++i) x[i] = csub(x[i], cmul(basis[i * 4 + b], dd));
        }
      double nr = 0.0;
      for (int i = 0; i < d; ++i) nr += cabs2(x[i]);
      nr = sqrt(nr);
      if (nr > 0.5) {
        for (int i = 0; i < d; ++i) basis[i * 4 + cnt] = cscale(x[i], 1.0 / nr);
        ++cnt;
      }
    }
  }
  // out = sum_m us_m vs_m^H
  for (int i = 0; i < d; ++i)
    for (int j = 0; j < d; ++j) {
      double2 s = cz();
      for (int m = 0; m < d; ++m) s = cfma(us[i * 4 + m], cconj(vs[j * 4 + m]), s);
      out[i * 4 + j] = s;
    }
}

// --------------------------------------------- Householder null space -----
// Trailing columns of the full Householder QR of rows^T (n x m, n <= 4,
// m < n), LAPACK zgeqr2 + zung2r conventions.  rows is m x n (stride 4).
// Writes q (n x n, stride 4); the null space is q[:, m:].
QMX_DEV void householder_q(int m, int n, const double2* rows, double2* q) {
  double2 a[16];
  for (int i = 0; i < n; ++i)
    for (int j = 0; j < m; ++j) a[i * 4 + j] = rows[j * 4 + i];
  double2 vs[4][4];
  double2 taus[4];
  for (int j = 0; j < m; ++j) {
    double2 alpha = a[j * 4 + j];
    double xn2 = 0.0;
    for (int i = j + 1; i < n; ++i) xn2 += cabs2(a[i * 4 + j]);
    double2 vv[4];
    for (int i = 0; i < 4; ++i) vv[i] = cz();
    vv[0] = make_double2(1.0, 0.0);
    double2 tau;
    if (xn2 == 0.0 && alpha.y == 0.0) {
      tau = cz();
    } else {
      double nrm = sqrt(alpha.x * alpha.x + alpha.y * alpha.y + xn2);
      double beta = signbit(alpha.x) ? nrm : -nrm;   // -sign(Re alpha) * norm
      tau = make_double2((beta - alpha.x) / beta, -alpha.y / beta);
      // scale = 1 / (alpha - beta)
      double2 den = make_double2(alpha.x - beta, alpha.y);
      double dd = cabs2(den);
      double2 scale = make_double2(den.x / dd, -den.y / dd);
      for (int i = j + 1; i < n; ++i) vv[i - j] = cmul(a[i * 4 + j], scale);
      a[j * 4 + j] = make_double2(beta, 0.0);
      for (int i = j + 1; i < n; ++i) a[i * 4 + j] = cz();
    }
    // apply H^H = I - conj(tau) v v^H to trailing columns
    for (int c = j + 1; c < m; ++c) {
      double2 w = cz();
      for (int i = j; i < n; ++i) w = cfmac(vv[i - j], a[i * 4 + c], w);
      double2 f = cmul(cconj(tau), w);
      for (int i = j; i < n; ++i) a[i * 4 + c] = csub(a[i * 4 + c], cmul(vv[i - j], f));
    }
    for (int i = 0; i < 4; ++i) vs[j][i] = vv[i];
    taus[j] = tau;
  }
  for (int i = 0; i < 4; ++i)
    for (int j = 0; j < 4; ++j) q[i * 4 + j] = make_double2(i == j ? 1.0 : 0.0, 0.0);
  for (int j = m - 1; j >= 0; --j) {
    // q[j:, :] -= tau v (v^H q[j:, :])
    for (int c = 0; c < n; ++c) {
      double2 w = cz();
      for (int i = j; i < n; ++i) w = cfmac(vs[j][i - j], q[i * 4 + c], w);
      double2 f = cmul(taus[j], w);
      for (int i = j; i < n; ++i) q[i * 4 + c] = csub(q[i * 4 + c], cmul(vs[j][i - j], f));
    }
  }
}

// ----------------------------------------------------------- bit helpers --
// Insert a zero bit at position pos of x.
QMX_DEV uint64_t insert_zero(uint64_t x, int pos) {
  uint64_t lo = x & ((1ull << pos) - 1ull);
  return ((x >> pos) << (pos + 1)) | lo;
}
