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
// Cyclic two-sided Jacobi on a 4 x 4 Hermitian matrix a (row-major; pad a
// smaller matrix with zeros).  On return lam[] holds the eigenvalues in
// decreasing order and the columns of v the matching orthonormal eigenvectors.
// All indices are compile-time so the arrays stay in registers.
QMX_DEV void herm_eig4(double2* a, double2* v, double* lam) {
#pragma unroll
  for (int i = 0; i < 16; ++i) v[i] = make_double2((i % 5) == 0 ? 1.0 : 0.0, 0.0);
  for (int sweep = 0; sweep < 40; ++sweep) {
    double off = 0.0, diag = 0.0;
#pragma unroll
    for (int i = 0; i < 4; ++i) {
      diag += a[i * 4 + i].x * a[i * 4 + i].x;
#pragma unroll
      for (int j = i + 1; j < 4; ++j) off += cabs2(a[i * 4 + j]);
    }
    if (off <= 1e-32 * diag || off == 0.0) break;
#pragma unroll
    for (int p = 0; p < 3; ++p) {
#pragma unroll
      for (int q = p + 1; q < 4; ++q) {
        double2 g = a[p * 4 + q];
        double ag = cabs_(g);
        if (ag != 0.0) {
          double al = a[p * 4 + p].x, be = a[q * 4 + q].x;
          double2 ph = make_double2(g.x / ag, g.y / ag);  // e^{i phi}
          double tau = (be - al) / (2.0 * ag);
          double t = (tau >= 0 ? 1.0 : -1.0) / (fabs(tau) + sqrt(1.0 + tau * tau));
          double c = 1.0 / sqrt(1.0 + t * t), sn = t * c;
          // J = [[c, s], [-s e^{-i phi}, c e^{-i phi}]] acting on columns (p, q)
          double2 em = cconj(ph);
          double2 j00 = make_double2(c, 0), j01 = make_double2(sn, 0);
          double2 j10 = cscale(em, -sn), j11 = cscale(em, c);
#pragma unroll
          for (int i = 0; i < 4; ++i) {
            double2 x = a[i * 4 + p], y = a[i * 4 + q];
            a[i * 4 + p] = cadd(cmul(x, j00), cmul(y, j10));
            a[i * 4 + q] = cadd(cmul(x, j01), cmul(y, j11));
            double2 vx = v[i * 4 + p], vy = v[i * 4 + q];
            v[i * 4 + p] = cadd(cmul(vx, j00), cmul(vy, j10));
            v[i * 4 + q] = cadd(cmul(vx, j01), cmul(vy, j11));
          }
#pragma unroll
          for (int j = 0; j < 4; ++j) {
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
  }
#pragma unroll
  for (int i = 0; i < 4; ++i) lam[i] = a[i * 4 + i].x;
  // decreasing order by a fixed compare-exchange network (stable on ties)
#define QMX_CX(i, j)                                                        \
  if (lam[j] > lam[i]) {                                                    \
    double tl = lam[i]; lam[i] = lam[j]; lam[j] = tl;                       \
    _Pragma("unroll") for (int r = 0; r < 4; ++r) {                         \
      double2 tv = v[r * 4 + i]; v[r * 4 + i] = v[r * 4 + j]; v[r * 4 + j] = tv; \
    }                                                                       \
  }
  QMX_CX(0, 1) QMX_CX(2, 3) QMX_CX(0, 2) QMX_CX(1, 3) QMX_CX(1, 2)
#undef QMX_CX
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

// Canonical orthonormal basis of span(v[:, 0:k]) (columns of a 4 x 4 v;
// pad smaller problems with zero rows): two-pass Gram-Schmidt over the
// columns of the projector P = V V^H in index order, skipping residuals with
// norm <= tol.  Writes w (columns 0..found-1).  Returns found (== k unless
// the input is degenerate).  Compile-time indices only.
QMX_DEV int canonical_basis4(int k, const double2* v, double2* w, double tol) {
  double2 p[16];
#pragma unroll
  for (int i = 0; i < 4; ++i)
#pragma unroll
    for (int j = 0; j < 4; ++j) {
      double2 s = cz();
#pragma unroll
      for (int m = 0; m < 4; ++m)
        if (m < k) s = cfma(v[i * 4 + m], cconj(v[j * 4 + m]), s);
      p[i * 4 + j] = s;
    }
#pragma unroll
  for (int i = 0; i < 16; ++i) w[i] = cz();
  int found = 0;
#pragma unroll
  for (int j = 0; j < 4; ++j) {
    if (found < k) {
      double2 x[4];
#pragma unroll
      for (int i = 0; i < 4; ++i) x[i] = p[i * 4 + j];
#pragma unroll
      for (int pass = 0; pass < 2; ++pass) {
#pragma unroll
        for (int b = 0; b < 4; ++b) {
          if (b < found) {
            double2 d = cz();
#pragma unroll
            for (int i = 0; i < 4; ++i) d = cfmac(w[i * 4 + b], x[i], d);
#pragma unroll
            for (int i = 0; i < 4; ++i) x[i] = csub(x[i], cmul(w[i * 4 + b], d));
          }
        }
      }
      double nrm = 0.0;
#pragma unroll
      for (int i = 0; i < 4; ++i) nrm += cabs2(x[i]);
      nrm = sqrt(nrm);
      if (nrm > tol) {
#pragma unroll
        for (int c = 0; c < 4; ++c)
          if (c == found) {
#pragma unroll
            for (int i = 0; i < 4; ++i) w[i * 4 + c] = cscale(x[i], 1.0 / nrm);
          }
        ++found;
      }
    }
  }
  return found;
}

// ------------------------------------------------------- polar (d <= 4) ----
// In-place inverse of a 4 x 4 complex matrix m by Gauss-Jordan elimination
// with partial (row) pivoting: row swaps are recorded and undone as column
// swaps at the end.  False when a pivot is zero or the result is not finite.
// Compile-time indices only (swaps are selects).
QMX_DEV bool inv4_inplace(double2* m) {
  int perm[4];
#pragma unroll
  for (int k = 0; k < 4; ++k) {
    int p = k;
    double best = cabs2(m[k * 4 + k]);
#pragma unroll
    for (int i = k + 1; i < 4; ++i) {
      const double v = cabs2(m[i * 4 + k]);
      if (v > best) {
        best = v;
        p = i;
      }
    }
    if (!(best > 0.0)) return false;
    perm[k] = p;
#pragma unroll
    for (int i = k + 1; i < 4; ++i) {
      if (p == i) {
#pragma unroll
        for (int j = 0; j < 4; ++j) {
          const double2 t = m[k * 4 + j];
          m[k * 4 + j] = m[i * 4 + j];
          m[i * 4 + j] = t;
        }
      }
    }
    const double ib = 1.0 / best;
    const double2 ip = make_double2(m[k * 4 + k].x * ib, -m[k * 4 + k].y * ib);   // 1 / pivot
    m[k * 4 + k] = make_double2(1.0, 0.0);
#pragma unroll
    for (int j = 0; j < 4; ++j) m[k * 4 + j] = cmul(m[k * 4 + j], ip);
#pragma unroll
    for (int i = 0; i < 4; ++i) {
      if (i != k) {
        const double2 f = m[i * 4 + k];
        m[i * 4 + k] = cz();
#pragma unroll
        for (int j = 0; j < 4; ++j) m[i * 4 + j] = csub(m[i * 4 + j], cmul(f, m[k * 4 + j]));
      }
    }
  }
#pragma unroll
  for (int k = 3; k >= 0; --k) {
#pragma unroll
    for (int c = k + 1; c < 4; ++c) {
      if (perm[k] == c) {
#pragma unroll
        for (int i = 0; i < 4; ++i) {
          const double2 t = m[i * 4 + k];
          m[i * 4 + k] = m[i * 4 + c];
          m[i * 4 + c] = t;
        }
      }
    }
  }
  double s = 0.0;
#pragma unroll
  for (int i = 0; i < 16; ++i) s += cabs2(m[i]);
  return s < 1e300;   // false for inf / nan
}

// Unitary polar factor of a nonsingular 4 x 4 matrix by the scaled Newton
// iteration (N. J. Higham, "Computing the polar decomposition -- with
// applications", SIAM J. Sci. Stat. Comput. 7 (1986)), in place on u (which
// holds the matrix on entry):
//   X_0 = u,  X_{k+1} = (z X_k + X_k^{-H} / z) / 2,
//   z = (||X_k^{-1}||_F / ||X_k||_F)^{1/2}  (z = 1 once ||X_{k+1} - X_k||_F < 1e-2),
// which converges quadratically; it stops when ||X_{k+1} - X_k||_F <= 1e-8
// (the last iterate is then accurate to rounding).  Returns false, with u
// unspecified, when the matrix is ill-conditioned (||X_0||_F ||X_0^{-1}||_F >
// 1e6), an inverse fails, or 40 iterations do not converge.
QMX_DEV bool polar4_newton(double2* u) {
  double2 xi[16];
  bool scale = true;
  for (int it = 0; it < 40; ++it) {
#pragma unroll
    for (int i = 0; i < 16; ++i) xi[i] = u[i];
    if (!inv4_inplace(xi)) return false;
    double nx = 0.0, ni = 0.0;
#pragma unroll
    for (int i = 0; i < 16; ++i) {
      nx += cabs2(u[i]);
      ni += cabs2(xi[i]);
    }
    if (it == 0 && nx * ni > 1e12) return false;
    const double z = scale ? sqrt(sqrt(ni / nx)) : 1.0;
    const double zi = 1.0 / z;
    double dl = 0.0;
    // X_{k+1}[i][j] = (z X[i][j] + conj(X^{-1}[j][i]) / z) / 2
#pragma unroll
    for (int i = 0; i < 4; ++i)
#pragma unroll
      for (int j = 0; j < 4; ++j) {
        const double2 a = u[i * 4 + j], b = xi[j * 4 + i];
        const double2 v = make_double2(0.5 * (z * a.x + zi * b.x), 0.5 * (z * a.y - zi * b.y));
        dl += cabs2(csub(v, a));
        u[i * 4 + j] = v;
      }
    if (dl < 1e-4) scale = false;
    if (dl <= 1e-16) return true;
  }
  return false;
}

// ---- the same Newton iteration on 16 lanes -----------------------------------
// Lanes 0..15 of one wave (all active, the rest of the wave idle) hold one
// entry each: lane l = 4 i + j holds X[i][j].  Cross-lane reads are __shfl
// within the 16 lanes; sums are xor butterflies, in which both partners add
// the same two values at every level, so every lane holds the bitwise same
// sum and every branch below is uniform over the 16 lanes.
QMX_DEV double2 shfl16(double2 v, int src) {
  return make_double2(__shfl(v.x, src, 16), __shfl(v.y, src, 16));
}
QMX_DEV double sum16(double v) {
#pragma unroll
  for (int m = 1; m < 16; m <<= 1) v += __shfl_xor(v, m, 16);
  return v;
}

// In-place inverse by Gauss-Jordan elimination with row pivoting (largest
// |entry| in the pivot column, first on ties), the operations of
// inv4_inplace entry by entry.  Returns false (on every lane) for a zero
// pivot or a non-finite result.
QMX_DEV bool inv4_lanes(double2& m, int l) {
  const int i = l >> 2, j = l & 3;
  int perm[4];
#pragma unroll
  for (int k = 0; k < 4; ++k) {
    const double c = cabs2(m);
    int p = k;
    double best = __shfl(c, k * 4 + k, 16);
#pragma unroll
    for (int r = k + 1; r < 4; ++r) {
      const double v = __shfl(c, r * 4 + k, 16);
      if (v > best) {
        best = v;
        p = r;
      }
    }
    if (!(best > 0.0)) return false;
    perm[k] = p;
    if (p != k) m = shfl16(m, i == k ? p * 4 + j : (i == p ? k * 4 + j : l));
    const double2 piv = shfl16(m, k * 4 + k);
    const double ib = 1.0 / best;
    const double2 ip = make_double2(piv.x * ib, -piv.y * ib);   // 1 / pivot
    const double2 rk = shfl16(m, k * 4 + j);                      // row k, column j
    const double2 f = shfl16(m, i * 4 + k);                       // row i, column k
    const double2 rkj = cmul(j == k ? make_double2(1.0, 0.0) : rk, ip);
    m = (i == k) ? rkj : csub(j == k ? cz() : m, cmul(f, rkj));
  }
#pragma unroll
  for (int k = 3; k >= 0; --k) {
    const int p = perm[k];
    if (p != k) m = shfl16(m, j == k ? i * 4 + p : (j == p ? i * 4 + k : l));
  }
  return sum16(cabs2(m)) < 1e300;   // false for inf / nan
}

// polar4_newton on 16 lanes (lane l holds X[l / 4][l % 4] on entry and the
// unitary factor on a true return).
QMX_DEV bool polar4_newton_lanes(double2& x, int l) {
  const int i = l >> 2, j = l & 3;
  bool scale = true;
  for (int it = 0; it < 40; ++it) {
    double2 xi = x;
    if (!inv4_lanes(xi, l)) return false;
    const double nx = sum16(cabs2(x)), ni = sum16(cabs2(xi));
    if (it == 0 && nx * ni > 1e12) return false;
    const double z = scale ? sqrt(sqrt(ni / nx)) : 1.0;
    const double zi = 1.0 / z;
    // X_{k+1}[i][j] = (z X[i][j] + conj(X^{-1}[j][i]) / z) / 2
    const double2 b = shfl16(xi, j * 4 + i);
    const double2 v = make_double2(0.5 * (z * x.x + zi * b.x), 0.5 * (z * x.y - zi * b.y));
    const double dl = sum16(cabs2(csub(v, x)));
    x = v;
    if (dl < 1e-4) scale = false;
    if (dl <= 1e-16) return true;
  }
  return false;
}

// Unitary polar factor U V^H of the d x d (d = 2 or 4) matrix e (row-major,
// stride 4).  Well-conditioned inputs (see polar4_newton) take the Newton
// iteration; the others (and any input it rejects) a one-sided (Hestenes) Jacobi SVD on a 4 x 4 working matrix (a
// 2 x 2 problem is embedded as diag(E, |E| I), whose padded columns are
// orthogonal to E's and never rotate).  Rotations: cyclic pairs p < q, when
// |a_p^H a_q| > 1e-16 |a_p| |a_q|.  Pairs (u_j, v_j) with sigma_j >
// 1e-13 sigma_max are kept; each other column j (in increasing j) receives
// the next unit vector e_0, e_1, ... whose two-pass Gram-Schmidt residual
// against the columns already set has norm > 1/2, on both the U and V sides.
// The result then depends only on the well-determined singular subspaces.
// *nuc = sum of E's singular values.  Compile-time indices only.
QMX_DEV void polar4_jacobi(int d, const double2* e, double2* out, double* nuc) {
  double2 a[16], v[16];
  double fro = 0.0;
#pragma unroll
  for (int i = 0; i < 16; ++i) {
    a[i] = e[i];
    fro += cabs2(e[i]);
    v[i] = make_double2((i % 5) == 0 ? 1.0 : 0.0, 0.0);
  }
  if (d == 2) {
    double s = sqrt(fro);
    if (s == 0.0) s = 1.0;
#pragma unroll
    for (int i = 0; i < 16; ++i) {
      int r = i / 4, c = i % 4;
      if (r >= 2 || c >= 2) a[i] = make_double2((r == c) ? s : 0.0, 0.0);
    }
  }
  for (int sweep = 0; sweep < 60; ++sweep) {
    bool rotated = false;
#pragma unroll
    for (int p = 0; p < 3; ++p)
#pragma unroll
      for (int q = p + 1; q < 4; ++q) {
        double al = 0.0, be = 0.0;
        double2 g = cz();
#pragma unroll
        for (int i = 0; i < 4; ++i) {
          al += cabs2(a[i * 4 + p]);
          be += cabs2(a[i * 4 + q]);
          g = cfmac(a[i * 4 + p], a[i * 4 + q], g);  // a_p^H a_q
        }
        double ag = cabs_(g);
        if (ag != 0.0 && ag > 1e-16 * sqrt(al * be)) {
          rotated = true;
          double2 ph = make_double2(g.x / ag, g.y / ag);
          double zeta = (be - al) / (2.0 * ag);
          double t = (zeta >= 0 ? 1.0 : -1.0) / (fabs(zeta) + sqrt(1.0 + zeta * zeta));
          double c = 1.0 / sqrt(1.0 + t * t), sn = t * c;
          // [a_p, a_q] <- [a_p, a_q] [[c, s e^{i phi}], [-s e^{-i phi}, c]]
          double2 sp = cscale(ph, sn), sm = cscale(cconj(ph), -sn);
#pragma unroll
          for (int i = 0; i < 4; ++i) {
            double2 x = a[i * 4 + p], y = a[i * 4 + q];
            a[i * 4 + p] = cadd(cscale(x, c), cmul(y, sm));
            a[i * 4 + q] = cadd(cmul(x, sp), cscale(y, c));
            double2 vx = v[i * 4 + p], vy = v[i * 4 + q];
            v[i * 4 + p] = cadd(cscale(vx, c), cmul(vy, sm));
            v[i * 4 + q] = cadd(cmul(vx, sp), cscale(vy, c));
          }
        }
      }
    if (!rotated) break;
  }
  double sig[4], smax = 0.0, ssum = 0.0;
#pragma unroll
  for (int j = 0; j < 4; ++j) {
    double s2 = 0.0;
#pragma unroll
    for (int i = 0; i < 4; ++i) s2 += cabs2(a[i * 4 + j]);
    sig[j] = sqrt(s2);
    if (j < d) {
      smax = fmax(smax, sig[j]);
      ssum += sig[j];
    }
  }
  *nuc = ssum;
  bool good[4];
#pragma unroll
  for (int j = 0; j < 4; ++j) {
    good[j] = (j >= d) || (smax > 0.0 && sig[j] > 1e-13 * smax);
    if (good[j]) {
#pragma unroll
      for (int i = 0; i < 4; ++i) a[i * 4 + j] = cscale(a[i * 4 + j], 1.0 / sig[j]);
    }
  }
  // canonical completion of the other columns, U side (a) then V side (v)
#pragma unroll
  for (int side = 0; side < 2; ++side) {
    double2* m = side == 0 ? a : v;
    bool set[4];
#pragma unroll
    for (int j = 0; j < 4; ++j) set[j] = good[j];
    int next_e = 0;
#pragma unroll
    for (int j = 0; j < 4; ++j) {
      if (!set[j]) {
#pragma unroll
        for (int e0 = 0; e0 < 4; ++e0) {
          if (!set[j] && e0 >= next_e) {
            double2 x[4];
#pragma unroll
            for (int i = 0; i < 4; ++i) x[i] = make_double2(i == e0 ? 1.0 : 0.0, 0.0);
#pragma unroll
            for (int pass = 0; pass < 2; ++pass)
#pragma unroll
              for (int b = 0; b < 4; ++b) {
                if (set[b]) {
                  double2 dd = cz();
#pragma unroll
                  for (int i = 0; i < 4; ++i) dd = cfmac(m[i * 4 + b], x[i], dd);
#pragma unroll
                  for (int i = 0; i < 4; ++i) x[i] = csub(x[i], cmul(m[i * 4 + b], dd));
                }
              }
            double nr = 0.0;
#pragma unroll
            for (int i = 0; i < 4; ++i) nr += cabs2(x[i]);
            nr = sqrt(nr);
            if (nr > 0.5) {
#pragma unroll
              for (int i = 0; i < 4; ++i) m[i * 4 + j] = cscale(x[i], 1.0 / nr);
              set[j] = true;
              next_e = e0 + 1;
            }
          }
        }
      }
    }
  }
  // out = U V^H (the d x d block)
#pragma unroll
  for (int i = 0; i < 4; ++i)
#pragma unroll
    for (int j = 0; j < 4; ++j) {
      double2 s = cz();
#pragma unroll
      for (int mm = 0; mm < 4; ++mm) s = cfma(a[i * 4 + mm], cconj(v[j * 4 + mm]), s);
      out[i * 4 + j] = (i < d && j < d) ? s : cz();
    }
}

QMX_DEV void polar4(int d, const double2* e, double2* out, double* nuc) {
  // the Newton iterate, in out: e, or diag(E, |E|_F I) for a 2 x 2 problem
  double fro = 0.0;
#pragma unroll
  for (int i = 0; i < 16; ++i) {
    const int r = i / 4, c = i % 4;
    out[i] = (r < d && c < d) ? e[i] : cz();
    fro += cabs2(out[i]);
  }
  if (d == 2) {
    const double s = fro > 0.0 ? sqrt(fro) : 1.0;
    out[10] = make_double2(s, 0.0);
    out[15] = make_double2(s, 0.0);
  }
  // e is read again below (from LDS where the kernels keep it) instead of
  // being held in registers through the iteration
  asm volatile("" ::: "memory");
  if (fro > 0.0 && polar4_newton(out)) {
    asm volatile("" ::: "memory");
    double tr = 0.0;   // Re tr(U^H E) = sum of the singular values
#pragma unroll
    for (int i = 0; i < 16; ++i) {
      const int r = i / 4, c = i % 4;
      if (r < d && c < d) {
        tr += out[i].x * e[i].x + out[i].y * e[i].y;
      } else {
        out[i] = cz();
      }
    }
    *nuc = tr;
    return;
  }
  asm volatile("" ::: "memory");
  polar4_jacobi(d, e, out, nuc);
}

// Over-relaxed local update (no reference counterpart).  g holds the plain
// update conj(polar(E)) of a gate whose value before the update was g_old;
// with omega_m1 = omega - 1 its d x d block becomes
//   polar(g + omega_m1 (g - g_old)),
// a step past the local optimum along the direction of the update, projected
// back onto the unitaries (omega = 1 leaves g as it is).  Returns the
// overlap after the update, |sum_{o,i} g[o][i] E[o][i]| over the block.
QMX_DEV double relax_update(int d, double omega_m1, const double2* g_old, double2* g, const double2* e,
                            double2* m) {
  // m (16 entries of scratch, LDS in the kernels: the polar reads it from
  // there rather than holding a second copy in registers)
#pragma unroll
  for (int k = 0; k < 16; ++k) {
    const bool in = (k / 4) < d && (k % 4) < d;
    m[k] = in ? cadd(g[k], cscale(csub(g[k], g_old[k]), omega_m1)) : cz();
  }
  double unused = 0.0;
  polar4(d, m, g, &unused);
  asm volatile("" ::: "memory");
  double2 ov = cz();
#pragma unroll
  for (int k = 0; k < 16; ++k)
    if ((k / 4) < d && (k % 4) < d) ov = cfma(g[k], e[k], ov);
  return cabs_(ov);
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
