// C-ABI entry points and host orchestration of libqmx (see include/qmx.h).
//
// The host code only enqueues kernels on the caller's stream; every shape that
// depends on data (bond dimensions after a cutoff, layers kept after the early
// exit, states that reached their fidelity goal) stays on the device and the
// kernels read it, so a whole prepare_state runs without host round trips
// except the error checks at the end of the chain and layer phases.
#include <hip/hip_runtime.h>

#include <algorithm>
#include <cmath>
#include <cstring>
#include <string>
#include <vector>

#include "qmx.h"
#include "qmx_internal.h"

// Phase timers (qmx_profile_*): event pairs recorded on the caller's stream
// around each phase when profiling is enabled.
struct ProfRec {
  int phase;
  hipEvent_t start, end;
  long long launches;
};

struct qmx_handle {
  int device;
  int cus;
  std::string err;
  bool profiling = false;
  int sweep_mode = QMX_SWEEP_AUTO;
  std::vector<ProfRec> recs;
  std::vector<hipEvent_t> pool;
  bool split_sweeps = true;             // QMX_OPT_SPLIT_SWEEPS
  int win_blocks_per_cu = 16;           // QMX_OPT_WINDOW_BLOCKS_PER_CU
  int fwd_blocks_per_cu = 4;            // QMX_OPT_FWD_BLOCKS_PER_CU
  int fwd_mfma = 1;                     // QMX_OPT_FWD_MFMA
  int staged_windows = 1;               // QMX_OPT_STAGED_WINDOWS (0, 1, 2)
  int relax_milli = 1000;               // QMX_OPT_SWEEP_RELAXATION (omega * 1000)
  int jacobi_sweep_cap = 0;             // QMX_OPT_JACOBI_SWEEP_CAP (0: the solvers' own limits)
  bool fail_unitarity = false;          // QMX_OPT_FAIL_UNITARITY (failure injection)
  hipStream_t aux[2] = {nullptr, nullptr};   // the two half-batch jobs
  hipEvent_t pass_ev[2] = {nullptr, nullptr};
  hipEvent_t join_ev[2] = {nullptr, nullptr};
  hipEvent_t call_ev = nullptr;
  void* scratch = nullptr;              // this device's qmx_sweep_noalloc scratch (qmx_scratch.hip)
  hipError_t scratch_status = hipSuccess;
};

namespace {
using namespace qmx;

constexpr int kVersion = 1;

struct Carver {
  char* base;
  size_t off = 0;
  explicit Carver(void* b) : base(static_cast<char*>(b)) {}
  template <typename T>
  T* take(size_t count) {
    off = (off + 255) & ~static_cast<size_t>(255);
    T* p = base ? reinterpret_cast<T*>(base + off) : nullptr;
    off += count * sizeof(T);
    return p;
  }
};

int fail(qmx_handle* h, int code, const std::string& msg) {
  if (h) h->err = msg;
  return code;
}

int hip_fail(qmx_handle* h, hipError_t e, const char* what) {
  return fail(h, QMX_ERR_HIP, std::string(what) + ": " + hipGetErrorString(e));
}

#define QMX_HIP(h, expr)                                  \
  do {                                                    \
    hipError_t e_ = (expr);                               \
    if (e_ != hipSuccess) return hip_fail(h, e_, #expr);  \
  } while (0)

#define QMX_LAUNCHED(h, what)                             \
  do {                                                    \
    hipError_t e_ = hipGetLastError();                    \
    if (e_ != hipSuccess) return hip_fail(h, e_, what);   \
  } while (0)

#define QMX_CALL(expr)            \
  do {                            \
    int s_ = (expr);              \
    if (s_ != QMX_OK) return s_;  \
  } while (0)

inline size_t state_len(int n) { return static_cast<size_t>(1) << n; }

hipEvent_t pool_event(qmx_handle* h) {
  if (!h->pool.empty()) {
    hipEvent_t e = h->pool.back();
    h->pool.pop_back();
    return e;
  }
  hipEvent_t e = nullptr;
  if (hipEventCreate(&e) != hipSuccess) return nullptr;
  return e;
}

// Scoped phase timer: records a start event now and an end event on stop().
struct PhaseTimer {
  qmx_handle* h;
  hipStream_t st;
  int phase;
  hipEvent_t start = nullptr;
  PhaseTimer(qmx_handle* h_, hipStream_t st_, int phase_) : h(h_), st(st_), phase(phase_) {
    if (h->profiling) {
      start = pool_event(h);
      if (start) (void)hipEventRecord(start, st);
    }
  }
  void stop(long long launches) {
    if (!start) return;
    hipEvent_t end = pool_event(h);
    if (!end) return;
    (void)hipEventRecord(end, st);
    h->recs.push_back({phase, start, end, launches});
    start = nullptr;
  }
};

int blocks_for(long ntiles, int batch, int cus) {
  long target = 4L * cus;
  long per = (target + batch - 1) / batch;
  if (per < 1) per = 1;
  if (per > ntiles) per = ntiles;
  if (per > 65535) per = 65535;
  return static_cast<int>(per);
}

int check_common(qmx_handle* h, int n, int batch) {
  if (!h) return QMX_ERR_INVALID;
  if (n < 2 || n > kMaxQubits) return fail(h, QMX_ERR_INVALID, "num_qubits must be in [2, 30]");
  if (batch < 1 || batch > 65535) return fail(h, QMX_ERR_INVALID, "batch must be in [1, 65535]");
  return QMX_OK;
}

int device_error_status(qmx_handle* h, int* err_dev, hipStream_t st) {
  int err = 0;
  QMX_HIP(h, hipMemcpyAsync(&err, err_dev, sizeof(int), hipMemcpyDeviceToHost, st));
  QMX_HIP(h, hipStreamSynchronize(st));
  if (err & kErrNotUnitary) return fail(h, QMX_ERR_NOT_UNITARY, "All the generated unitaries must be unitary.");
  if (err & kErrEigNoConv) return fail(h, QMX_ERR_NO_CONVERGENCE, "Jacobi eigensolver did not converge");
  if (err & kErrBasis) return fail(h, QMX_ERR_BASIS, "degenerate projector in the canonical gauge");
  return QMX_OK;
}

// Worst-case gate bits of site i (width 2 unless it is the last site).
uint64_t worst_bits(int n, int site) {
  uint64_t m = 1ull << (n - 1 - site);
  if (site < n - 1) m |= 1ull << (n - 2 - site);
  return m;
}

// Greedy grouping of a gate sequence into passes whose bit union fits a tile.
std::vector<std::vector<int>> plan_passes(int n, const std::vector<int>& pos,
                                          const std::vector<int>& sites, int max_union) {
  std::vector<std::vector<int>> passes;
  std::vector<int> cur;
  uint64_t um = 0;
  for (size_t j = 0; j < pos.size(); ++j) {
    uint64_t nb = um | worst_bits(n, sites[j]);
    if (!cur.empty() && (__builtin_popcountll(nb) > max_union ||
                         static_cast<int>(cur.size()) >= kMaxPassGates)) {
      passes.push_back(cur);
      cur.clear();
      nb = worst_bits(n, sites[j]);
    }
    cur.push_back(pos[j]);
    um = nb;
  }
  if (!cur.empty()) passes.push_back(cur);
  return passes;
}

// The Jacobi kernels handle Gram matrices up to 1024 x 1024, i.e. chi <= 512
// whenever the chain truncates (2^(N-1) > chi).
bool chi_supported(int n, int chi) { return chi <= 512 || (1l << (n - 1)) <= chi; }

constexpr int kJacobiBlockMax = 512;  // one-block global Jacobi up to 512 x 512

// ------------------------------------------------------------ layouts -------
struct ChiLayout {
  double2* psi;      // only when the caller gives no psi_out
  double2* xa;
  double2* xb;
  double* nrm;
  double* npart;
  int* ncnt;
  double2* a;
  double2* v;
  double2* w;        // [i_first + 1][B][nmax^2]
  double2* part;
  int* nn;
  int* kk;
  int* ncur;
  int* error;
  double* jlog;      // LDS Jacobi rotation log (nmax <= 128)
  int* jstate;
  char* jblock;      // block Jacobi state (nmax > 512)
  int i_first, nmax, nchunks_max, nblk_norm;
};

ChiLayout chi_layout(Carver& c, int n, int batch, int chi, bool need_psi) {
  ChiLayout L{};
  size_t len = state_len(n);
  int i_first = 0;
  for (int i = n - 1; i >= 1; --i)
    if ((1l << (n - i)) > chi) { i_first = i; break; }
  L.i_first = i_first;
  // n_i <= 2 k_{i+1} <= 2 chi, and n_{i_first} = 2^(N - i_first) <= 2 chi
  L.nmax = i_first ? 2 * chi : 2;
  L.nchunks_max = 16;
  L.nblk_norm = 256;
  L.psi = need_psi ? c.take<double2>(len * batch) : nullptr;
  L.xa = i_first ? c.take<double2>(len * batch) : nullptr;
  L.xb = i_first ? c.take<double2>(len * batch) : nullptr;
  L.nrm = c.take<double>(batch);
  L.npart = c.take<double>(static_cast<size_t>(batch) * L.nblk_norm * 2);
  L.ncnt = c.take<int>(batch);
  size_t nm2 = static_cast<size_t>(L.nmax) * L.nmax;
  if (i_first) {
    L.a = c.take<double2>(nm2 * batch);
    L.v = c.take<double2>(nm2 * batch);
    L.w = c.take<double2>(nm2 * batch * (i_first + 1));
    L.part = c.take<double2>(nm2 * batch * L.nchunks_max);
    if (jacobi_lds_fits(L.nmax)) {
      L.jlog = c.take<double>(jacobi_lds_log_doubles(L.nmax) * batch);
      L.jstate = c.take<int>(2 * static_cast<size_t>(batch));
    }
    if (L.nmax > kJacobiBlockMax) L.jblock = c.take<char>(jacobi_block_bytes(L.nmax, batch));
  }
  L.nn = c.take<int>(static_cast<size_t>(batch) * (n + 1));
  L.kk = c.take<int>(static_cast<size_t>(batch) * (n + 1));
  L.ncur = c.take<int>(batch);
  L.error = c.take<int>(1);
  return L;
}

struct LayerLayout {
  double2* xa;
  double2* xb;
  int* r;
  double2* w;
  double* part;
  int* counter;
  int* skip;
  int* error;
};
constexpr int kChain2MaxBlocks = 256;
constexpr int kJacobiSweeps = 30;   // sweep launches of the LDS Jacobi per bond (16-17 run at 2r = 128)
// sweeps of the block Jacobi (2r > 512): 1 to 34 per bond (16 on average)
// on the Gram matrices of an N=26 random-MPS state of bond 1024 at chi = 512
constexpr int kJacobiBlockSweeps = 80;
constexpr int kJacobiGlobalSweeps = 40;   // one-block Jacobi on global memory (2r <= 512)
constexpr int kWinMaxBlocks = 256;
// the finalize sums up to this many block partials per state itself; above
// it, window_reduce_kernel pre-sums groups of 32
constexpr int kWinFinalizePartials = 64;
constexpr int kWinMaxBlocksLarge = 1024;

// window-pass blocks per state: about win_blocks_per_cu resident blocks per CU
// over the whole batch.  Each block writes a 512-double partial; per state the
// sweep layout holds max(2^(n-10) * 32, kWinMaxBlocks * 512) doubles, so large
// states (few per batch) may use up to kWinMaxBlocksLarge blocks (more would
// make the one-block finalize's partial sum the bottleneck).
int window_blocks(const qmx_handle* h, int n, int batch) {
  const long ntiles = 1l << (n - std::min(n, kWinTileBits));
  long per = ((long)h->win_blocks_per_cu * h->cus + batch - 1) / batch;
  per = std::max(1l, std::min(per, ntiles));
  const long held = std::max<long>(kWinMaxBlocks, ((1l << (n - std::min(n, kSweepTileBits))) * 32) / 512);
  return (int)std::min<long>(per, std::min<long>(held, kWinMaxBlocksLarge));
}


struct Window {
  int s0, s1;
  uint64_t umask;
};

// ---- tile geometry (host side) ---------------------------------------------
uint64_t tile_bits(int n, int t, uint64_t need) {
  uint64_t p = need;
  for (int bit = 0; bit < n && __builtin_popcountll(p) < t; ++bit) p |= 1ull << bit;
  return p;
}

GroupMap make_group(int n, uint64_t p, uint64_t u) {
  GroupMap g{};
  g.mask = u;
  int m = 0;
  for (int bit = 0; bit < n && m < 4; ++bit)
    if ((u >> bit) & 1ull) g.ubit[m++] = bit;
  auto local = [&](int bit) { return __builtin_popcountll(p & ((1ull << bit) - 1ull)); };
  for (int q = 0; q < 4; ++q) g.ux[q] = local(g.ubit[q]);
  for (int bit = 0; bit < n; ++bit)
    if (((p >> bit) & 1ull) && !((u >> bit) & 1ull)) {
      g.restg |= 1ull << bit;
      g.restl |= 1u << local(bit);
    }
  return g;
}

// Lane roles of one stage of the matrix-core forward pass (tiles of 2^12
// amplitudes).  Within an LDS access, the 32 lanes of a phase vary the 4 jl
// bits and kq's low bit; with the tile swizzle (low 5 bits ^= bits 5-9 ^
// bits 10-11) they fall on distinct banks when the 5 tile-local positions
// have distinct residues (b mod 5 below 10, b - 10 above).  A stage that
// reads or writes HBM (the first and the last) keeps jl on the 4 lowest rest
// bits, whose accesses coalesce; only kq is chosen there.
MfmaMap make_mfma_map(int n, uint64_t p, const GroupMap& g, bool hbm) {
  int restl[8], restg[8], m = 0;
  for (int bit = 0; bit < n && m < 8; ++bit)
    if (((p >> bit) & 1ull) && !((g.mask >> bit) & 1ull)) {
      restg[m] = bit;
      restl[m] = __builtin_popcountll(p & ((1ull << bit) - 1ull));
      ++m;
    }
  auto residue = [](int b) { return b < 10 ? b % 5 : b - 10; };
  int best = -1, bsel = 0, bk = 0;
  for (int sel = 0; sel < 256; ++sel) {
    if (__builtin_popcount(sel) != 4 || (hbm && sel != 15)) continue;
    for (int k0 = 0; k0 < 4; ++k0) {
      int seen = 1 << residue(g.ux[k0]);
      for (int i = 0; i < 8; ++i)
        if ((sel >> i) & 1) seen |= 1 << residue(restl[i]);
      const int score = __builtin_popcount(seen);
      if (score > best) best = score, bsel = sel, bk = k0;
    }
  }
  auto put = [](uint32_t& w, int i, int v) { w |= (uint32_t)v << (8 * i); };
  MfmaMap mm{};
  int nj = 0, nh = 0;
  for (int i = 0; i < 8; ++i) {
    if ((bsel >> i) & 1) {
      put(mm.jl, nj, restl[i]);
      put(mm.jg, nj++, restg[i]);
    } else {
      put(mm.hl, nh, restl[i]);
      put(mm.hg, nh++, restg[i]);
    }
  }
  int kx[2] = {bk, bk == 0 ? 1 : 0}, sx[2], ns = 0;
  for (int q = 0; q < 4; ++q)
    if (q != kx[0] && q != kx[1]) sx[ns++] = q;
  for (int i = 0; i < 2; ++i) {
    put(mm.ksl, i, g.ux[kx[i]]);
    put(mm.ksl, 2 + i, g.ux[sx[i]]);
    put(mm.ksg, i, g.ubit[kx[i]]);
    put(mm.ksg, 2 + i, g.ubit[sx[i]]);
    put(mm.kxsx, i, kx[i]);
    put(mm.kxsx, 2 + i, sx[i]);
  }
  return mm;
}

// Forward (circuit-order) passes over the light cone of a bit window.  Two
// gates conflict when their worst-case bits intersect, and a gate may run
// once every earlier conflicting gate has run.  Each pass picks the window of
// W consecutive state bits (W = kFwdMaxUnion, or all n bits when the tile is
// the whole state) whose closure is largest: every remaining gate inside the
// window whose earlier conflicting remaining gates are taken too.  Across
// layers this is a band of the staircase that runs ahead of the window's
// lower edge, so one HBM pass applies up to W(W-1)/2 gates instead of one
// layer's W-1.  The pass is then cut the same way into stages over
// contiguous 4-bit sets (register groups, LDS exchange between stages).
struct FwdPass {
  int nstage;
  uint64_t mask[kMaxStages];
  int gstart[kMaxStages + 1];
  int fpos[kMaxFwdGates];
};

// gates of rem (forward order) inside `allowed` whose earlier conflicting
// gates of rem are taken as well
void cone_take(int n, const std::vector<int>& rem, uint64_t allowed, std::vector<int>& took) {
  took.clear();
  uint64_t blocked = 0;
  for (int f : rem) {
    const uint64_t b = worst_bits(n, f % n);
    if (!(b & ~allowed) && !(b & blocked))
      took.push_back(f);
    else
      blocked |= b;
  }
}

std::vector<FwdPass> plan_fwd(int n, int kmax) {
  const int W = n <= kFwdTileBits ? n : kFwdMaxUnion;
  std::vector<int> rem(kmax), took, best, prem, sbest;
  for (int f = 0; f < kmax; ++f) rem[f] = f;
  std::vector<FwdPass> out;
  while (!rem.empty()) {
    best.clear();
    int a0 = 0;
    for (int a = 0; a + W <= n; ++a) {
      cone_take(n, rem, ((1ull << W) - 1ull) << a, took);
      if (took.size() > best.size()) {
        best = took;
        a0 = a;
      }
    }
    FwdPass p{};
    int ng = 0;
    prem = best;
    while (!prem.empty() && p.nstage < kMaxStages && ng < kMaxFwdGates) {
      sbest.clear();
      uint64_t smask = 0;
      for (int c = a0; c + 4 <= a0 + W; ++c) {
        cone_take(n, prem, 15ull << c, took);
        if (took.size() > sbest.size()) {
          sbest = took;
          smask = 15ull << c;
        }
      }
      // a prefix of a stage is still closed
      if (ng + (int)sbest.size() > kMaxFwdGates) sbest.resize(kMaxFwdGates - ng);
      p.mask[p.nstage] = smask;
      p.gstart[p.nstage] = ng;
      for (int f : sbest) p.fpos[ng++] = f;
      ++p.nstage;
      std::vector<int> keep;
      size_t k = 0;
      for (int f : prem) {
        if (k < sbest.size() && sbest[k] == f)
          ++k;
        else
          keep.push_back(f);
      }
      prem.swap(keep);
    }
    p.gstart[p.nstage] = ng;
    // gates of the window left over (caps) stay in rem for the next passes
    std::vector<int> keep;
    std::vector<char> taken(kmax, 0);
    for (int j = 0; j < ng; ++j) taken[p.fpos[j]] = 1;
    for (int f : rem)
      if (!taken[f]) keep.push_back(f);
    rem.swap(keep);
    out.push_back(p);
  }
  return out;
}

// Sweep windows: maximal runs of consecutive steps (reverse circuit order)
// whose worst-case qubits fit in 4 bits; each bit set is padded to exactly 4
// bits with the lowest free bits.
std::vector<Window> plan_windows(int n, int kmax) {
  std::vector<Window> out;
  int s = 0;
  while (s < kmax) {
    Window w{s, s, 0};
    while (w.s1 < kmax && w.s1 - w.s0 < kMaxWindow) {
      int site = n - 1 - w.s1 % n;
      uint64_t nb = w.umask | worst_bits(n, site);
      if (__builtin_popcountll(nb) > 4) break;
      w.umask = nb;
      ++w.s1;
    }
    for (int bit = 0; bit < n && __builtin_popcountll(w.umask) < 4; ++bit) w.umask |= 1ull << bit;
    out.push_back(w);
    s = w.s1;
  }
  return out;
}

LayerLayout layer_layout(Carver& c, int n, int batch) {
  LayerLayout L{};
  size_t len = state_len(n);
  L.xa = c.take<double2>(len * batch);
  L.xb = c.take<double2>(len * batch);
  L.r = c.take<int>(static_cast<size_t>(batch) * (n + 1));
  L.w = c.take<double2>(static_cast<size_t>(batch) * (n + 1) * 16);
  L.part = c.take<double>(static_cast<size_t>(batch) * kChain2MaxBlocks * 32);
  L.counter = c.take<int>(batch);
  L.skip = c.take<int>(batch);
  L.error = c.take<int>(1);
  return L;
}

template <typename S>
struct SweepLayout {
  S* phi;
  S* tprime;
  double* part;
  int* counter;
  double* dot_part;
  int* dot_counter;
  int* skip;
  int* active;
  int* list;   // live states after the last goal check
  int nblk_max;
};

template <typename S>
SweepLayout<S> sweep_layout(Carver& c, int n, int batch) {
  SweepLayout<S> L{};
  size_t len = state_len(n);
  L.nblk_max = 65535;
  L.phi = c.take<S>(len * batch);
  L.tprime = c.take<S>(len * batch);
  // partial buffers sized for the largest grid any pass can use
  long ntiles = 1l << (n - std::min(n, kSweepTileBits));
  long ntiles_apply = 1l << (n - std::min(n, kApplyTileBits));
  long mx = std::max(ntiles, ntiles_apply);
  L.part = c.take<double>(static_cast<size_t>(batch) * std::max<long>(mx * 32, kWinMaxBlocks * 512));
  L.counter = c.take<int>(batch);
  L.dot_part = c.take<double>(static_cast<size_t>(batch) * mx * 2);
  L.dot_counter = c.take<int>(batch);
  L.skip = c.take<int>(batch);
  L.active = c.take<int>(1);
  L.list = c.take<int>(batch);
  return L;
}

// ------------------------------------------------------------- phases -------
int run_chi(qmx_handle* h, int n, int batch, int chi, double cutoff, const void* sv, int sv_dtype,
            double2* psi, double2* target, double* spectra, const ChiLayout& L, hipStream_t st) {
  const size_t len = state_len(n);
  launch_fill_int(L.ncnt, 0, batch, st);
  launch_fill_int(L.error, 0, 1, st);
  int nb = static_cast<int>(std::min<size_t>(L.nblk_norm, std::max<size_t>(1, len / 1024)));
  if (sv_dtype == QMX_C64) {
    launch_norm2(static_cast<const float2*>(sv), n, batch, L.npart, L.ncnt, L.nrm, nb, st);
    launch_scale_convert(static_cast<const float2*>(sv), psi, n, batch, L.nrm, st);
  } else {
    launch_norm2(static_cast<const double2*>(sv), n, batch, L.npart, L.ncnt, L.nrm, nb, st);
    launch_scale_convert(static_cast<const double2*>(sv), psi, n, batch, L.nrm, st);
  }
  QMX_LAUNCHED(h, "normalise");
  PhaseTimer timer(h, st, QMX_PHASE_CHI);
  const int i_first = L.i_first;
  if (i_first == 0) {
    launch_scale_convert(static_cast<const double2*>(psi), target, n, batch, nullptr, st);
    QMX_LAUNCHED(h, "copy");
    timer.stop(1);
    return QMX_OK;
  }
  const int nmax = L.nmax;
  const size_t nm2 = static_cast<size_t>(nmax) * nmax;
  const int kmax = std::min(chi, nmax);
  auto xbuf = [&](int i) { return ((i_first - i) % 2 == 0) ? L.xa : L.xb; };
  // forward chain: X_i = Y_i W_i
  for (int i = i_first; i >= 1; --i) {
    const double2* y = (i == i_first) ? psi : xbuf(i + 1);
    launch_chain_dims(L.nn, L.ncur, L.kk, n, i, (i == i_first) ? (1 << (n - i)) : 0, batch, st);
    int rows = 1 << i;
    int nchunks = std::max(1, std::min(L.nchunks_max, rows / 256));
    launch_gram(y, len, L.ncur, i, nmax, L.part, nchunks, batch, st);
    launch_gram_reduce(L.part, nchunks, L.ncur, nmax, L.a, nm2, batch, st);
    // sweep limits of the three eigensolvers; QMX_OPT_JACOBI_SWEEP_CAP lowers them
    auto cap = [&](int limit) { return h->jacobi_sweep_cap > 0 ? std::min(limit, h->jacobi_sweep_cap) : limit; };
    if (L.jblock) {
      std::vector<int> conv(batch);
      if (launch_jacobi_block(L.a, L.v, nm2, L.ncur, nmax, batch, L.jblock, conv.data(), cap(kJacobiBlockSweeps),
                              L.error, st) != 0)
        return fail(h, QMX_ERR_HIP, "block Jacobi");
    } else if (L.jlog && jacobi_lds_setup(nmax) == hipSuccess) {
      launch_jacobi_lds(L.a, L.v, nm2, L.ncur, nmax, batch, L.jlog, L.jstate, cap(kJacobiSweeps), L.error, st);
    } else {
      launch_jacobi(L.a, L.v, nm2, L.ncur, nmax, batch, cap(kJacobiGlobalSweeps), L.error, st);
    }
    double2* wi = L.w + static_cast<size_t>(i) * batch * nm2;
    launch_select(L.a, L.v, wi, nm2, L.ncur, nmax, chi, cutoff, L.kk + i, n + 1,
                  spectra ? spectra + static_cast<size_t>(i) * 2 * chi : nullptr,
                  static_cast<size_t>(n + 1) * 2 * chi, batch, st);
    GemmDims d{i, L.nn + i, L.kk + i, n + 1, nmax, kmax, nmax};
    launch_gemm(y, len, wi, nm2, xbuf(i), len, d, 0, batch, st);
    QMX_LAUNCHED(h, "chi chain step");
  }
  // contraction back: Z_{i+1} = Z_i W_i^H, Z_1 = X_1
  const double2* z = xbuf(1);
  for (int i = 1; i <= i_first; ++i) {
    double2* out = (i == i_first) ? target : (z == L.xa ? L.xb : L.xa);
    const double2* wi = L.w + static_cast<size_t>(i) * batch * nm2;
    GemmDims d{i, L.kk + i, L.nn + i, n + 1, kmax, nmax, nmax};
    launch_gemm(z, len, wi, nm2, out, len, d, 1, batch, st);
    QMX_LAUNCHED(h, "chi contraction step");
    z = out;
  }
  timer.stop(7L * i_first);
  return QMX_OK;
}

int run_generate_layer(qmx_handle* h, int n, int batch, int lmax, int layer, double cutoff,
                       const double2* psi_w, double2* gates, int* widths, int* bonds,
                       const LayerLayout& L, const int* skip, hipStream_t st) {
  for (int i = n - 1; i >= 1; --i) {
    Chain2Args a{};
    a.n = n;
    a.step = i;
    a.cutoff = cutoff;
    a.psi = psi_w;
    a.xa = L.xa;
    a.xb = L.xb;
    a.r = L.r;
    a.w = L.w;
    a.spectra = nullptr;
    a.skip = skip;
    a.part = L.part;
    a.counter = L.counter;
    a.error = L.error;
    long rows = 1l << i;
    // about 1024 blocks over the batch, each thread >= 4 rows
    const long per = std::max(1l, (1024l + batch - 1) / batch);
    int nblk = static_cast<int>(std::max(1l, std::min<long>({(long)kChain2MaxBlocks, rows / 1024, per})));
    launch_chain2_step(a, nblk, batch, st);
  }
  RanArgs r{};
  r.n = n;
  r.lmax = lmax;
  r.layer = layer;
  r.psi = psi_w;
  r.xa = L.xa;
  r.xb = L.xb;
  r.r = L.r;
  r.w = L.w;
  r.gates = gates;
  r.widths = widths;
  r.bonds = bonds;
  r.skip = skip;
  r.error = L.error;
  r.unitary_atol = h->fail_unitarity ? -1.0 : 1e-8;
  launch_ran_gates(r, batch, st);
  QMX_LAUNCHED(h, "generate layer");
  return QMX_OK;
}

template <typename S>
int run_passes(qmx_handle* h, int n, int batch, int lmax, int mode, int layer,
               const std::vector<std::vector<int>>& passes, S* psi, const double2* gates,
               const int* widths, const int* layers, const int* skip, const double2* ref,
               double* dot_part, int* dot_counter, double2* dot_out, int dot_stride, int cus,
               hipStream_t st) {
  const int t = std::min(n, kApplyTileBits);
  const long ntiles = 1l << (n - t);
  const int nblk = blocks_for(ntiles, batch, cus);
  for (size_t p = 0; p < passes.size(); ++p) {
    PassArgs<S> a{};
    a.n = n;
    a.lmax = lmax;
    a.tbits = kApplyTileBits;
    a.mode = mode;
    a.layer = layer;
    a.count = static_cast<int>(passes[p].size());
    for (int j = 0; j < a.count; ++j) a.pos[j] = passes[p][j];
    a.psi = psi;
    a.gates = gates;
    a.widths = widths;
    a.layers = layers;
    a.skip = skip;
    bool last = (p + 1 == passes.size());
    a.ref = last ? ref : nullptr;
    a.dot_part = dot_part;
    a.dot_counter = dot_counter;
    a.dot_out = dot_out;
    a.dot_stride = dot_stride;
    launch_apply_pass(a, nblk, batch, st);
  }
  QMX_LAUNCHED(h, "apply passes");
  return QMX_OK;
}

std::vector<std::vector<int>> inverse_layer_passes(int n) {
  std::vector<int> pos, sites;
  for (int i = n - 1; i >= 0; --i) { pos.push_back(i); sites.push_back(i); }
  return plan_passes(n, pos, sites, kApplyMaxUnion);
}

std::vector<std::vector<int>> forward_circuit_passes(int n, int lmax) {
  std::vector<int> pos, sites;
  for (int f = 0; f < lmax * n; ++f) { pos.push_back(f); sites.push_back(f % n); }
  return plan_passes(n, pos, sites, kApplyMaxUnion);
}

// Forward circuit passes (register-group stages) from |0..0>.  Needs n >= 5.
// The support of the state grows with the bits the gates touch; passes
// before it covers every bit visit only the tiles inside it.
template <typename S>
int run_fwd(qmx_handle* h, int n, int batch, int lmax, const std::vector<FwdPass>& passes, S* psi,
            const double2* gates, const int* widths, const int* layers, const int* skip,
            const double2* ref, double* dot_part, int* dot_counter, double2* dot_out,
            int dot_stride, hipStream_t st, const int* list = nullptr, int rows = 0) {
  if (passes.empty()) return fail(h, QMX_ERR_INVALID, "empty forward-pass plan");
  // grid rows: the batch, or the `rows` live states of list
  const int nrow = list ? rows : batch;
  if (nrow < 1) return QMX_OK;
  const int t = std::min(n, kFwdTileBits);
  const uint64_t all = (1ull << n) - 1ull;
  uint64_t live = 0;
  for (size_t pi = 0; pi < passes.size(); ++pi) {
    const FwdPass& fp = passes[pi];
    uint64_t um = 0, touched = 0;
    for (int k = 0; k < fp.nstage; ++k) um |= fp.mask[k];
    for (int j = 0; j < fp.gstart[fp.nstage]; ++j) touched |= worst_bits(n, fp.fpos[j] % n);
    const uint64_t p = tile_bits(n, t, um);
    FwdArgs<S> a{};
    a.n = n;
    a.lmax = lmax;
    a.t = t;
    a.nstage = fp.nstage;
    a.from_basis0 = pi == 0 ? 1 : 0;
    a.live_in = live;
    live |= touched;
    a.live_out = live;
    a.full = (live == all || pi + 1 == passes.size()) ? 1 : 0;
    const uint64_t qmask = all & ~p;
    a.tmask = a.full ? qmask : (qmask & live);
    for (int k = 0; k < fp.nstage; ++k) a.gs[k] = make_group(n, p, fp.mask[k]);
    if (fwd_mfma_applies(h->fwd_mfma, t, n, sizeof(S)))
      for (int k = 0; k < fp.nstage; ++k) a.mm[k] = make_mfma_map(n, p, a.gs[k], k == 0 || k + 1 == fp.nstage);
    for (int k = 0; k <= fp.nstage; ++k) a.gstart[k] = fp.gstart[k];
    for (int j = 0; j < fp.gstart[fp.nstage]; ++j) a.fpos[j] = fp.fpos[j];
    a.psi = psi;
    a.gates = gates;
    a.widths = widths;
    a.layers = layers;
    a.skip = skip;
    a.list = list;
    a.mfma = h->fwd_mfma;
    const long ntiles = 1l << __builtin_popcountll(a.tmask);
    const int nblk =
        (int)std::max(1l, std::min(ntiles, ((long)h->fwd_blocks_per_cu * h->cus + nrow - 1) / nrow));
    launch_fwd_pass(a, nblk, nrow, st);
    if (a.full) live = all;
  }
  // the overlap with ref is a separate coalesced pass: in the last circuit
  // pass the gate bits are the lowest ones, where a fused read of the
  // complex128 reference would be strided
  if (ref) launch_overlap<S>(psi, ref, n, batch, dot_part, dot_counter, dot_out, dot_stride, skip, st, list, nrow);
  QMX_LAUNCHED(h, "forward passes");
  return QMX_OK;
}

int run_extract(qmx_handle* h, int n, int batch, int lmax, double thr, double cutoff,
                double2* psi_w, double2* gates, int* widths, int* bonds, int* layers_used,
                double2* amp0, const LayerLayout& L, hipStream_t st) {
  launch_fill_int(L.skip, 0, batch, st);
  launch_fill_int(L.counter, 0, batch, st);
  launch_fill_int(L.error, 0, 1, st);
  launch_fill_int(layers_used, 0, batch, st);
  launch_fill_int(widths, 1, batch * lmax * n, st);
  auto inv = inverse_layer_passes(n);
  PhaseTimer timer(h, st, QMX_PHASE_LAYERS);
  for (int l = 0; l < lmax; ++l) {
    QMX_CALL(run_generate_layer(h, n, batch, lmax, l, cutoff, psi_w, gates, widths, bonds, L,
                                L.skip, st));
    QMX_CALL(run_passes<double2>(h, n, batch, lmax, 0, l, inv, psi_w, gates, widths, nullptr,
                                 L.skip, nullptr, nullptr, nullptr, nullptr, 1, h->cus, st));
    Amp0Args a{};
    a.n = n;
    a.batch = batch;
    a.lmax = lmax;
    a.layer = l;
    a.atol = 1.0 - thr;
    a.psi = psi_w;
    a.skip = L.skip;
    a.layers = layers_used;
    a.amp0 = amp0;
    launch_amp0(a, st);
    QMX_LAUNCHED(h, "early exit check");
  }
  timer.stop(static_cast<long long>(lmax) * (n + 1 + inv.size()));
  return device_error_status(h, L.error, st);
}

// One batch of states swept on one stream.  qmx_sweep runs one job, or two
// jobs (the two halves of the batch) on the handle's two auxiliary streams;
// their window passes alternate, so each window's latency-bound finalize
// overlaps the other half's bandwidth-bound pass.
template <typename S>
struct SweepJob {
  int n, batch, lmax, nsweeps;
  const int* layers_used;
  const S* target;
  double2* gates;
  const int* widths;
  const double2* psi_ref;
  double2* fidelity;
  double* overlap;
  const double* goal;
  int* sweeps_done;
  S* circuit_out;
  SweepLayout<S> L;
  hipStream_t st;
  const int* skip = nullptr;
  bool use_goal = false;
  bool active = true;
  // After a goal check with states done: the live states' indices (device)
  // and their count; the sweep's launches then cover the live states only,
  // with as many blocks per state as a full grid holds.  Otherwise null and
  // the batch.
  const int* list = nullptr;
  int live = 0;
  // gate snapshots (qmx_sweep_snapshots): this job's gates after snap_first
  // and snap_first + 1 completed sweeps of the call go to snaps and
  // snaps + snap_stride (null: none)
  int snap_first = -1;
  double2* snaps = nullptr;
  size_t snap_stride = 0;
  std::vector<FwdPass> fwd_plan;   // forward passes of the circuit rebuild
};

template <typename S>
int job_begin(qmx_handle* h, SweepJob<S>& J) {
  const hipStream_t st = J.st;
  launch_fill_int(J.L.counter, 0, J.batch, st);
  launch_fill_int(J.L.dot_counter, 0, J.batch, st);
  launch_fill_int(J.L.skip, 0, J.batch, st);
  if (J.fidelity)
    launch_fill_int(reinterpret_cast<int*>(J.fidelity), 0, J.batch * (J.nsweeps + 1) * 4, st);
  if (J.sweeps_done) launch_fill_int(J.sweeps_done, J.goal ? 0 : J.nsweeps, J.batch, st);
  J.use_goal = J.goal && J.psi_ref && J.fidelity;
  J.skip = J.use_goal ? J.L.skip : nullptr;
  J.list = nullptr;
  J.live = J.batch;
  QMX_LAUNCHED(h, "sweep set-up");
  return QMX_OK;
}

// C|0..0> (before sweep s, or the final state when s == nsweeps) and the
// fidelity entry s
template <typename S>
int job_rebuild(qmx_handle* h, SweepJob<S>& J, int s, bool final_state) {
  const int n = J.n;
  const int* skip = final_state ? nullptr : J.skip;
  const double2* ref = J.fidelity ? J.psi_ref : nullptr;
  double2* dot = J.fidelity ? J.fidelity + s : nullptr;
  if (n >= 5 && h->sweep_mode != QMX_SWEEP_STEP) {
    if (J.fwd_plan.empty()) J.fwd_plan = plan_fwd(n, J.lmax * n);
    const int* list = final_state ? nullptr : J.list;
    QMX_CALL(run_fwd<S>(h, n, J.batch, J.lmax, J.fwd_plan, J.L.phi, J.gates, J.widths, J.layers_used, skip,
                        ref, J.L.dot_part, J.L.dot_counter, dot, J.nsweeps + 1, J.st, list,
                        list ? J.live : J.batch));
  } else {
    launch_basis0(J.L.phi, n, J.batch, skip, J.st);
    QMX_CALL(run_passes<S>(h, n, J.batch, J.lmax, 1, 0, forward_circuit_passes(n, J.lmax), J.L.phi,
                           J.gates, J.widths, J.layers_used, skip, ref, J.L.dot_part, J.L.dot_counter,
                           dot, J.nsweeps + 1, h->cus, J.st));
  }
  return QMX_OK;
}

template <typename S>
int job_goal(qmx_handle* h, SweepJob<S>& J, int s) {
  launch_fill_int(J.L.active, 0, 1, J.st);
  launch_goal_check(J.fidelity + s, J.nsweeps + 1, J.goal, J.L.skip, J.L.active, J.sweeps_done, J.batch,
                    J.st);
  launch_live_list(J.L.skip, J.batch, J.L.list, J.st);
  int active = 0;
  QMX_HIP(h, hipMemcpyAsync(&active, J.L.active, sizeof(int), hipMemcpyDeviceToHost, J.st));
  QMX_HIP(h, hipStreamSynchronize(J.st));
  J.active = active != 0;
  J.live = active;
  J.list = active < J.batch ? J.L.list : nullptr;
  return QMX_OK;
}

template <typename S>
void job_window_pass(qmx_handle* h, SweepJob<S>& J, const std::vector<Window>& wins, size_t wi) {
  const int n = J.n;
  const Window& w = wins[wi];
  const int wblk = window_blocks(h, n, J.live);
  WinPassArgs<S> pa{};
  pa.n = n;
  pa.lmax = J.lmax;
  pa.t = std::min(n, kWinTileBits);
  pa.s0 = w.s0;
  pa.s1 = w.s1;
  pa.p0 = wi ? wins[wi - 1].s0 : 0;
  pa.p1 = wi ? wins[wi - 1].s1 : 0;
  const uint64_t pmask = wi ? wins[wi - 1].umask : w.umask;
  const uint64_t pbits = tile_bits(n, pa.t, w.umask | pmask);
  pa.qmask = ((1ull << n) - 1ull) & ~pbits;
  pa.gu = make_group(n, pbits, w.umask);
  pa.gp = make_group(n, pbits, pmask);
  pa.target = J.target;
  pa.tprime = J.L.tprime;
  pa.phi = J.L.phi;
  pa.gates = J.gates;
  pa.widths = J.widths;
  pa.layers = J.layers_used;
  pa.skip = J.skip;
  pa.list = J.list;
  pa.part = J.L.part;
  // after a sweep's last window neither T' nor phi is read again (the next
  // sweep rebuilds phi from |0..0> and restarts T' from the target)
  pa.store = wi + 1 < wins.size() ? 1 : 0;
  // When a side's groups stride the state by >= 4 amplitudes per lane (its
  // gate bits are the lowest state bits), staging the tiles through LDS in
  // tile order pays if the tile's contiguous low run is long enough.
  // Measured at N=20 (us per launch, direct -> staged): both sides strided
  // 322 -> 244, one side by 16 amplitudes 355 -> 287 (low run 7), one side
  // by 4 amplitudes 287 -> 271 (low run 9) but 251 -> 273 (low run 6).
  // Mode 1 (default) stages from a low run of 7 bits, mode 2 from 3.
  const int low_run = __builtin_ctzll(~pbits);
  const int stride_bits = std::max(__builtin_ctzll(pa.gu.restg | (1ull << 63)),
                                   __builtin_ctzll(pa.gp.restg | (1ull << 63)));
  const int min_run = h->staged_windows == 2 ? 3 : 7;
  pa.staged = (h->staged_windows && stride_bits >= 2 && low_run >= min_run && pa.t - low_run <= 8) ? 1 : 0;
  PhaseTimer pt(h, J.st, QMX_PHASE_WINDOW_PASS);
  launch_window_pass(pa, wblk, J.live, J.st);
  pt.stop(1);
}

template <typename S>
void job_window_finalize(qmx_handle* h, SweepJob<S>& J, const std::vector<Window>& wins, size_t wi, int s) {
  const int n = J.n;
  const Window& w = wins[wi];
  WinFinArgs fa{};
  fa.n = n;
  fa.lmax = J.lmax;
  fa.nblk = window_blocks(h, n, J.live);
  fa.s0 = w.s0;
  fa.s1 = w.s1;
  fa.umask = w.umask;
  fa.gates = J.gates;
  fa.widths = J.widths;
  fa.layers = J.layers_used;
  fa.skip = J.skip;
  fa.list = J.list;
  fa.part = J.L.part;
  fa.overlap = J.overlap ? J.overlap + s : nullptr;
  fa.overlap_stride = J.nsweeps;
  fa.omega_m1 = (h->relax_milli - 1000) * 1e-3;
  fa.pstride = 1;
  if (fa.nblk > kWinFinalizePartials) {
    // large grids (few states, or few live ones): pre-sum groups of 32
    // partials with a grid of blocks, so the one-block finalize reads
    // nblk / 32 of them
    launch_window_reduce(J.L.part, fa.nblk, 32, J.live, J.list, J.st);
    fa.pstride = 32;
  }
  launch_window_finalize(fa, J.live, J.st);
}

template <typename S>
void job_step_sweep(qmx_handle* h, SweepJob<S>& J, int s) {
  const int n = J.n;
  const int kmax = J.lmax * n;
  const int nblk = blocks_for(1l << (n - std::min(n, kSweepTileBits)), J.batch, h->cus);
  for (int k = 0; k < kmax; ++k) {
    SweepArgs<S> a{};
    a.n = n;
    a.lmax = J.lmax;
    a.tbits = kSweepTileBits;
    a.step = k;
    a.target = J.target;
    a.tprime = J.L.tprime;
    a.phi = J.L.phi;
    a.gates = J.gates;
    a.widths = J.widths;
    a.layers = J.layers_used;
    a.skip = J.skip;
    a.part = J.L.part;
    a.counter = J.L.counter;
    a.overlap = J.overlap ? J.overlap + s : nullptr;
    a.overlap_stride = J.nsweeps;
    a.omega_m1 = (h->relax_milli - 1000) * 1e-3;
    launch_sweep_step(a, nblk, J.batch, J.st);
  }
}

// the job's gates after `done` completed sweeps of the call, when requested
template <typename S>
int job_snapshot(qmx_handle* h, SweepJob<S>& J, int done) {
  if (!J.snaps || done < J.snap_first || done > J.snap_first + 1) return QMX_OK;
  const size_t bytes = (size_t)J.batch * J.lmax * J.n * 16 * sizeof(double2);
  QMX_HIP(h, hipMemcpyAsync(J.snaps + (size_t)(done - J.snap_first) * J.snap_stride, J.gates, bytes,
                            hipMemcpyDeviceToDevice, J.st));
  return QMX_OK;
}

// record `ev` on `from` and make `to` wait for it
int stream_join(qmx_handle* h, hipEvent_t ev, hipStream_t from, hipStream_t to) {
  QMX_HIP(h, hipEventRecord(ev, from));
  QMX_HIP(h, hipStreamWaitEvent(to, ev, 0));
  return QMX_OK;
}

template <typename S>
int run_sweep_jobs(qmx_handle* h, std::vector<SweepJob<S>>& jobs) {
  const int n = jobs[0].n, lmax = jobs[0].lmax, nsweeps = jobs[0].nsweeps;
  const int kmax = lmax * n;
  const bool windowed = (h->sweep_mode == QMX_SWEEP_WINDOW || h->sweep_mode == QMX_SWEEP_AUTO) &&
                        n >= kWinMinQubits;
  const auto wins = plan_windows(n, kmax);
  const size_t nj = jobs.size();
  const hipStream_t s0 = jobs[0].st;
  for (auto& J : jobs) QMX_CALL(job_begin(h, J));
  for (auto& J : jobs) QMX_CALL(job_snapshot(h, J, 0));
  const long long rebuild_launches =
      (n >= 5 && h->sweep_mode != QMX_SWEEP_STEP) ? (long long)plan_fwd(n, kmax).size()
                                                   : (long long)forward_circuit_passes(n, lmax).size();
  for (int s = 0; s < nsweeps; ++s) {
    PhaseTimer rt(h, s0, QMX_PHASE_REBUILD);
    for (auto& J : jobs)
      if (J.active) QMX_CALL(job_rebuild(h, J, s, false));
    for (size_t j = 1; j < nj; ++j) QMX_CALL(stream_join(h, h->join_ev[j], jobs[j].st, s0));
    rt.stop(rebuild_launches);
    bool any = false;
    for (auto& J : jobs) {
      if (J.use_goal && J.active) QMX_CALL(job_goal(h, J, s));
      any = any || J.active;
    }
    if (!any) break;
    PhaseTimer swt(h, s0, QMX_PHASE_SWEEP);
    if (windowed) {
      for (size_t wi = 0; wi < wins.size(); ++wi) {
        for (size_t j = 0; j < nj; ++j) {
          SweepJob<S>& J = jobs[j];
          if (!J.active) continue;
          // passes of the two jobs never overlap: wait for the other job's last pass
          if (nj > 1) QMX_HIP(h, hipStreamWaitEvent(J.st, h->pass_ev[1 - j], 0));
          job_window_pass(h, J, wins, wi);
          if (nj > 1) QMX_HIP(h, hipEventRecord(h->pass_ev[j], J.st));
          job_window_finalize(h, J, wins, wi, s);
        }
      }
      QMX_LAUNCHED(h, "windowed sweep");
    } else {
      for (auto& J : jobs)
        if (J.active) job_step_sweep(h, J, s);
      QMX_LAUNCHED(h, "sweep");
    }
    for (auto& J : jobs) QMX_CALL(job_snapshot(h, J, s + 1));
    for (size_t j = 1; j < nj; ++j) QMX_CALL(stream_join(h, h->join_ev[j], jobs[j].st, s0));
    swt.stop(windowed ? (long long)wins.size() : (long long)kmax);
  }
  // final circuit state and fidelity
  for (auto& J : jobs) {
    QMX_CALL(job_rebuild(h, J, nsweeps, true));
    if (J.circuit_out) launch_scale_convert<S, S>(J.L.phi, J.circuit_out, n, J.batch, nullptr, J.st);
  }
  QMX_LAUNCHED(h, "final circuit state");
  return QMX_OK;
}

bool sweep_windowed(const qmx_handle* h, int n) {
  return (h->sweep_mode == QMX_SWEEP_WINDOW || h->sweep_mode == QMX_SWEEP_AUTO) && n >= kWinMinQubits;
}

template <typename S>
int sweep_dispatch(qmx_handle* h, int n, int batch, int lmax, const int* layers_used, int nsweeps,
                   const S* target, double2* gates, const int* widths, const double2* psi_ref,
                   double2* fidelity, double* overlap, const double* goal, int* sweeps_done,
                   S* circuit_out, Carver& c, hipStream_t st, int snap_first = -1,
                   double2* snapshots = nullptr) {
  const bool split = h->split_sweeps && batch >= 2 && sweep_windowed(h, n);
  const size_t len = state_len(n);
  auto job = [&](int b0, int nb, hipStream_t js) {
    SweepJob<S> J{};
    J.n = n;
    J.batch = nb;
    J.lmax = lmax;
    J.nsweeps = nsweeps;
    J.layers_used = layers_used + b0;
    J.target = target + (size_t)b0 * len;
    J.gates = gates + (size_t)b0 * lmax * n * 16;
    J.widths = widths + (size_t)b0 * lmax * n;
    J.psi_ref = psi_ref ? psi_ref + (size_t)b0 * len : nullptr;
    J.fidelity = fidelity ? fidelity + (size_t)b0 * (nsweeps + 1) : nullptr;
    J.overlap = overlap ? overlap + (size_t)b0 * nsweeps : nullptr;
    J.goal = goal ? goal + b0 : nullptr;
    J.sweeps_done = sweeps_done ? sweeps_done + b0 : nullptr;
    J.circuit_out = circuit_out ? circuit_out + (size_t)b0 * len : nullptr;
    if (snapshots && snap_first >= 0) {
      J.snap_first = snap_first;
      J.snaps = snapshots + (size_t)b0 * lmax * n * 16;
      J.snap_stride = (size_t)batch * lmax * n * 16;
    }
    J.L = sweep_layout<S>(c, n, nb);
    J.st = js;
    return J;
  };
  std::vector<SweepJob<S>> jobs;
  if (!split) {
    jobs.push_back(job(0, batch, st));
    return run_sweep_jobs<S>(h, jobs);
  }
  for (int i = 0; i < 2; ++i) {
    if (!h->aux[i]) QMX_HIP(h, hipStreamCreateWithFlags(&h->aux[i], hipStreamNonBlocking));
    if (!h->pass_ev[i]) QMX_HIP(h, hipEventCreateWithFlags(&h->pass_ev[i], hipEventDisableTiming));
  }
  for (int i = 0; i < 2; ++i)
    if (!h->join_ev[i]) QMX_HIP(h, hipEventCreateWithFlags(&h->join_ev[i], hipEventDisableTiming));
  if (!h->call_ev) QMX_HIP(h, hipEventCreateWithFlags(&h->call_ev, hipEventDisableTiming));
  const int b1 = (batch + 1) / 2;
  jobs.push_back(job(0, b1, h->aux[0]));
  jobs.push_back(job(b1, batch - b1, h->aux[1]));
  // the auxiliary streams start after the caller's earlier work ...
  QMX_HIP(h, hipEventRecord(h->call_ev, st));
  for (int i = 0; i < 2; ++i) QMX_HIP(h, hipStreamWaitEvent(h->aux[i], h->call_ev, 0));
  const int rc = run_sweep_jobs<S>(h, jobs);
  // ... and the caller's later work waits for both
  for (int i = 0; i < 2; ++i) QMX_CALL(stream_join(h, h->join_ev[i], h->aux[i], st));
  return rc;
}

template <typename S>
size_t sweep_bytes(int n, int batch) {
  Carver one(nullptr);
  sweep_layout<S>(one, n, batch);
  Carver two(nullptr);
  if (batch >= 2) {
    sweep_layout<S>(two, n, (batch + 1) / 2);
    sweep_layout<S>(two, n, batch / 2);
  }
  return std::max(one.off, two.off) + 256;
}

// ---- learned warm start: MLP parameter and activation layout -------------
struct MlpDims {
  int in, w, out;
  size_t oW1() const { return 0; }
  size_t ob1() const { return oW1() + (size_t)in * w; }
  size_t oW2() const { return ob1() + w; }
  size_t ob2() const { return oW2() + (size_t)w * w; }
  size_t oW3() const { return ob2() + w; }
  size_t ob3() const { return oW3() + (size_t)w * out; }
  size_t count() const { return ob3() + out; }
};

bool mlp_dims_ok(int rows, int in, int w, int out) {
  return rows >= 1 && in >= 1 && w >= 1 && out >= 1 && in <= 4096 && w <= 8192 && out <= 4096;
}

struct MlpLayout {
  uint16_t *a1, *a2, *dyb, *dz2, *dz1;
  float* part;
};
constexpr int kMseBlocks = 64;

MlpLayout mlp_layout(Carver& c, int rows, const MlpDims& d) {
  MlpLayout L{};
  L.a1 = c.take<uint16_t>((size_t)rows * d.w);
  L.a2 = c.take<uint16_t>((size_t)rows * d.w);
  L.dyb = c.take<uint16_t>((size_t)rows * d.out);
  L.dz2 = c.take<uint16_t>((size_t)rows * d.w);
  L.dz1 = c.take<uint16_t>((size_t)rows * d.w);
  L.part = c.take<float>(kMseBlocks);
  return L;
}

MlpGemmArgs gemm(int M, int N, int K, const uint16_t* A, int lda, int ta, const uint16_t* B, int ldb, int tb) {
  MlpGemmArgs g{};
  g.M = M; g.N = N; g.K = K;
  g.A = A; g.lda = lda; g.transA = ta;
  g.B = B; g.ldb = ldb; g.transB = tb;
  return g;
}

int gate_rows_check(qmx_handle* h, int n, int batch, int num_layers, GateRowArgs& r,
                    const int* layers_used, const int* widths) {
  QMX_CALL(check_common(h, n, batch));
  if (num_layers < 1 || !layers_used || !widths) return fail(h, QMX_ERR_INVALID, "gate rows: bad arguments");
  const long rows = (long)batch * num_layers * (n - 1);
  if (rows > (1l << 30)) return fail(h, QMX_ERR_INVALID, "gate rows: too many gates");
  r.n = n;
  r.lmax = num_layers;
  r.rows = (int)rows;
  r.layers = layers_used;
  r.widths = widths;
  return QMX_OK;
}

}  // namespace

void qmx_set_error_text(qmx_handle* h, const char* msg) {
  if (h) h->err = msg ? msg : "";
}
int qmx_handle_device(const qmx_handle* h) { return h ? h->device : 0; }
// the qmx_set_option settings of `from`, for handles that act on its behalf
void qmx_copy_options(const qmx_handle* from, qmx_handle* to) {
  if (!from || !to) return;
  to->sweep_mode = from->sweep_mode;
  to->split_sweeps = from->split_sweeps;
  to->win_blocks_per_cu = from->win_blocks_per_cu;
  to->fwd_blocks_per_cu = from->fwd_blocks_per_cu;
  to->staged_windows = from->staged_windows;
  to->relax_milli = from->relax_milli;
  to->fwd_mfma = from->fwd_mfma;
  to->jacobi_sweep_cap = from->jacobi_sweep_cap;
  to->fail_unitarity = from->fail_unitarity;
}

// =================================================================== C-ABI ==
extern "C" {

int qmx_version(void) { return kVersion; }

int qmx_create(int device, qmx_handle** out) {
  if (!out) return QMX_ERR_INVALID;
  int count = 0;
  if (hipGetDeviceCount(&count) != hipSuccess || device < 0 || device >= count) {
    *out = nullptr;
    return QMX_ERR_HIP;
  }
  hipDeviceProp_t prop;
  if (hipGetDeviceProperties(&prop, device) != hipSuccess) return QMX_ERR_HIP;
  qmx_handle* h = new qmx_handle();
  h->device = device;
  h->cus = prop.multiProcessorCount > 0 ? prop.multiProcessorCount : 256;
  // the allocation-free entry point's scratch: its device address, looked up
  // with the handle's device current (the caller's current device restored)
  int prev = 0;
  if (hipGetDevice(&prev) == hipSuccess && hipSetDevice(device) == hipSuccess) {
    h->scratch_status = scratch_address(device, &h->scratch);
    (void)hipSetDevice(prev);
  } else {
    h->scratch_status = hipErrorInvalidDevice;
  }
  if (h->scratch_status != hipSuccess) h->scratch = nullptr;
  *out = h;
  return QMX_OK;
}

int qmx_destroy(qmx_handle* h) {
  if (!h) return QMX_OK;
  for (auto& r : h->recs) {
    (void)hipEventDestroy(r.start);
    (void)hipEventDestroy(r.end);
  }
  for (auto e : h->pool) (void)hipEventDestroy(e);
  for (int i = 0; i < 2; ++i) {
    if (h->aux[i]) (void)hipStreamDestroy(h->aux[i]);
    if (h->pass_ev[i]) (void)hipEventDestroy(h->pass_ev[i]);
  }
  for (int i = 0; i < 2; ++i)
    if (h->join_ev[i]) (void)hipEventDestroy(h->join_ev[i]);
  if (h->call_ev) (void)hipEventDestroy(h->call_ev);
  delete h;
  return QMX_OK;
}

const char* qmx_last_error(const qmx_handle* h) { return h ? h->err.c_str() : "null handle"; }

int qmx_set_option(qmx_handle* h, int option, int value) {
  if (!h) return QMX_ERR_INVALID;
  if (option == QMX_OPT_SWEEP_MODE) {
    if (value < QMX_SWEEP_AUTO || value > QMX_SWEEP_WINDOW) return fail(h, QMX_ERR_INVALID, "sweep mode");
    h->sweep_mode = value;
    return QMX_OK;
  }
  if (option == QMX_OPT_WINDOW_BLOCKS_PER_CU) {
    if (value < 1 || value > 16) return fail(h, QMX_ERR_INVALID, "window blocks per CU: 1..16");
    h->win_blocks_per_cu = value;
    return QMX_OK;
  }
  if (option == QMX_OPT_STAGED_WINDOWS) {
    if (value < 0 || value > 2) return fail(h, QMX_ERR_INVALID, "staged windows: 0, 1 or 2");
    h->staged_windows = value;
    return QMX_OK;
  }
  if (option == QMX_OPT_FWD_MFMA) {
    if (value < 0 || value > 1) return fail(h, QMX_ERR_INVALID, "forward passes on the matrix cores: 0 or 1");
    h->fwd_mfma = value;
    return QMX_OK;
  }
  if (option == QMX_OPT_FWD_BLOCKS_PER_CU) {
    if (value < 1 || value > 16) return fail(h, QMX_ERR_INVALID, "forward blocks per CU: 1..16");
    h->fwd_blocks_per_cu = value;
    return QMX_OK;
  }
  if (option == QMX_OPT_SWEEP_RELAXATION) {
    if (value < 1000 || value > 1950) return fail(h, QMX_ERR_INVALID, "sweep relaxation: 1000..1950 (omega * 1000)");
    h->relax_milli = value;
    return QMX_OK;
  }
  if (option == QMX_OPT_JACOBI_SWEEP_CAP) {
    if (value < 0 || value > 1000) return fail(h, QMX_ERR_INVALID, "Jacobi sweep cap: 0..1000 (0: no cap)");
    h->jacobi_sweep_cap = value;
    return QMX_OK;
  }
  if (option == QMX_OPT_FAIL_UNITARITY) {
    if (value != 0 && value != 1) return fail(h, QMX_ERR_INVALID, "fail unitarity: 0 or 1");
    h->fail_unitarity = value != 0;
    return QMX_OK;
  }
  if (option == QMX_OPT_SPLIT_SWEEPS) {
    if (value != 0 && value != 1) return fail(h, QMX_ERR_INVALID, "split sweeps: 0 or 1");
    h->split_sweeps = value != 0;
    return QMX_OK;
  }
  return fail(h, QMX_ERR_INVALID, "unknown option");
}

int qmx_forward_plan(int num_qubits, int num_layers, int* records, size_t record_capacity,
                     unsigned long long* stage_masks, size_t mask_capacity, int* num_records,
                     int* num_stages) {
  if (num_qubits < 5 || num_qubits > kMaxQubits || num_layers < 1 || !num_records || !num_stages)
    return QMX_ERR_INVALID;
  const auto plan = plan_fwd(num_qubits, num_layers * num_qubits);
  int nr = 0, ns = 0;
  for (const FwdPass& p : plan) {
    nr += p.gstart[p.nstage];
    ns += p.nstage;
  }
  *num_records = nr;
  *num_stages = ns;
  if (!records || !stage_masks || record_capacity < static_cast<size_t>(3) * nr ||
      mask_capacity < static_cast<size_t>(ns))
    return QMX_ERR_WORKSPACE;
  int r = 0, sg = 0;
  for (size_t pi = 0; pi < plan.size(); ++pi) {
    const FwdPass& p = plan[pi];
    for (int k = 0; k < p.nstage; ++k, ++sg) {
      stage_masks[sg] = p.mask[k];
      for (int j = p.gstart[k]; j < p.gstart[k + 1]; ++j, ++r) {
        records[3 * r] = static_cast<int>(pi);
        records[3 * r + 1] = sg;
        records[3 * r + 2] = p.fpos[j];
      }
    }
  }
  return QMX_OK;
}

int qmx_profile_enable(qmx_handle* h, int enable) {
  if (!h) return QMX_ERR_INVALID;
  h->profiling = enable != 0;
  return QMX_OK;
}

int qmx_profile_read(qmx_handle* h, double* ms, long long* launches) {
  if (!h || !ms || !launches) return QMX_ERR_INVALID;
  for (int p = 0; p < QMX_PHASE_COUNT; ++p) { ms[p] = 0.0; launches[p] = 0; }
  for (auto& r : h->recs) {
    QMX_HIP(h, hipEventSynchronize(r.end));
    float t = 0.0f;
    QMX_HIP(h, hipEventElapsedTime(&t, r.start, r.end));
    if (r.phase >= 0 && r.phase < QMX_PHASE_COUNT) {
      ms[r.phase] += t;
      launches[r.phase] += r.launches;
    }
    h->pool.push_back(r.start);
    h->pool.push_back(r.end);
  }
  h->recs.clear();
  return QMX_OK;
}

size_t qmx_chi_truncate_workspace(int num_qubits, int batch, int bond_dimension) {
  if (num_qubits < 2 || num_qubits > kMaxQubits || batch < 1 || bond_dimension < 1) return 0;
  Carver c(nullptr);
  chi_layout(c, num_qubits, batch, bond_dimension, true);
  return c.off + 256;
}

int qmx_chi_truncate(qmx_handle* h, int num_qubits, int batch, int bond_dimension, double cutoff,
                     const void* statevectors, int sv_dtype, void* psi_out, void* target_out,
                     double* spectra, void* workspace, size_t workspace_bytes, void* stream) {
  QMX_CALL(check_common(h, num_qubits, batch));
  if (bond_dimension < 1) return fail(h, QMX_ERR_INVALID, "`bond_dimension` must be an integer greater than 0.");
  if (!statevectors || !target_out || !workspace) return fail(h, QMX_ERR_INVALID, "null pointer");
  if (sv_dtype != QMX_C64 && sv_dtype != QMX_C128) return fail(h, QMX_ERR_INVALID, "sv_dtype");
  if (!chi_supported(num_qubits, bond_dimension))
    return fail(h, QMX_ERR_INVALID, "bond_dimension above 512 with truncation is not supported");
  if (workspace_bytes < qmx_chi_truncate_workspace(num_qubits, batch, bond_dimension))
    return fail(h, QMX_ERR_WORKSPACE, "workspace too small");
  QMX_HIP(h, hipSetDevice(h->device));
  hipStream_t st = static_cast<hipStream_t>(stream);
  Carver c(workspace);
  ChiLayout L = chi_layout(c, num_qubits, batch, bond_dimension, true);
  double2* psi = psi_out ? static_cast<double2*>(psi_out) : L.psi;
  QMX_CALL(run_chi(h, num_qubits, batch, bond_dimension, cutoff, statevectors, sv_dtype, psi,
                   static_cast<double2*>(target_out), spectra, L, st));
  return device_error_status(h, L.error, st);
}

size_t qmx_layer_workspace(int num_qubits, int batch) {
  if (num_qubits < 2 || num_qubits > kMaxQubits || batch < 1) return 0;
  Carver c(nullptr);
  layer_layout(c, num_qubits, batch);
  return c.off + 256;
}

int qmx_generate_layer(qmx_handle* h, int num_qubits, int batch, int num_layers, int layer,
                       double cutoff, const void* psi_work, void* gates, int* widths, int* bonds,
                       void* workspace, size_t workspace_bytes, void* stream) {
  QMX_CALL(check_common(h, num_qubits, batch));
  if (num_layers < 1 || layer < 0 || layer >= num_layers) return fail(h, QMX_ERR_INVALID, "layer index");
  if (!psi_work || !gates || !widths || !workspace) return fail(h, QMX_ERR_INVALID, "null pointer");
  if (workspace_bytes < qmx_layer_workspace(num_qubits, batch))
    return fail(h, QMX_ERR_WORKSPACE, "workspace too small");
  QMX_HIP(h, hipSetDevice(h->device));
  hipStream_t st = static_cast<hipStream_t>(stream);
  Carver c(workspace);
  LayerLayout L = layer_layout(c, num_qubits, batch);
  launch_fill_int(L.counter, 0, batch, st);
  launch_fill_int(L.error, 0, 1, st);
  QMX_CALL(run_generate_layer(h, num_qubits, batch, num_layers, layer, cutoff,
                              static_cast<const double2*>(psi_work), static_cast<double2*>(gates),
                              widths, bonds, L, nullptr, st));
  return device_error_status(h, L.error, st);
}

int qmx_apply_layer(qmx_handle* h, int num_qubits, int batch, int num_layers, int layer,
                    int inverse, void* psi, const void* gates, const int* widths, void* stream) {
  QMX_CALL(check_common(h, num_qubits, batch));
  if (num_layers < 1 || layer < 0 || layer >= num_layers) return fail(h, QMX_ERR_INVALID, "layer index");
  if (!psi || !gates || !widths) return fail(h, QMX_ERR_INVALID, "null pointer");
  QMX_HIP(h, hipSetDevice(h->device));
  hipStream_t st = static_cast<hipStream_t>(stream);
  const int n = num_qubits;
  if (inverse) {
    return run_passes<double2>(h, n, batch, num_layers, 0, layer, inverse_layer_passes(n),
                               static_cast<double2*>(psi), static_cast<const double2*>(gates),
                               widths, nullptr, nullptr, nullptr, nullptr, nullptr, nullptr, 1,
                               h->cus, st);
  }
  // forward: positions f = (L-1-layer)*n + i with layers = nullptr -> L
  std::vector<int> pos, sites;
  for (int i = 0; i < n; ++i) {
    pos.push_back((num_layers - 1 - layer) * n + i);
    sites.push_back(i);
  }
  return run_passes<double2>(h, n, batch, num_layers, 1, 0, plan_passes(n, pos, sites, kApplyMaxUnion),
                             static_cast<double2*>(psi), static_cast<const double2*>(gates), widths,
                             nullptr, nullptr, nullptr, nullptr, nullptr, nullptr, 1, h->cus, st);
}

size_t qmx_extract_layers_workspace(int num_qubits, int batch) {
  return qmx_layer_workspace(num_qubits, batch);
}

int qmx_extract_layers(qmx_handle* h, int num_qubits, int batch, int num_layers,
                       double fidelity_threshold, double cutoff, void* psi_work, void* gates,
                       int* widths, int* bonds, int* layers_used, void* amp0, void* workspace,
                       size_t workspace_bytes, void* stream) {
  QMX_CALL(check_common(h, num_qubits, batch));
  if (num_layers < 1) return fail(h, QMX_ERR_INVALID, "The number of layers must be a positive integer.");
  if (!psi_work || !gates || !widths || !layers_used || !workspace)
    return fail(h, QMX_ERR_INVALID, "null pointer");
  if (workspace_bytes < qmx_extract_layers_workspace(num_qubits, batch))
    return fail(h, QMX_ERR_WORKSPACE, "workspace too small");
  QMX_HIP(h, hipSetDevice(h->device));
  hipStream_t st = static_cast<hipStream_t>(stream);
  Carver c(workspace);
  LayerLayout L = layer_layout(c, num_qubits, batch);
  return run_extract(h, num_qubits, batch, num_layers, fidelity_threshold, cutoff,
                     static_cast<double2*>(psi_work), static_cast<double2*>(gates), widths, bonds,
                     layers_used, static_cast<double2*>(amp0), L, st);
}

size_t qmx_sweep_workspace(int num_qubits, int batch, int storage_dtype) {
  if (num_qubits < 2 || num_qubits > kMaxQubits || batch < 1) return 0;
  if (storage_dtype != QMX_C64 && storage_dtype != QMX_C128) return 0;
  return storage_dtype == QMX_C64 ? sweep_bytes<float2>(num_qubits, batch)
                                  : sweep_bytes<double2>(num_qubits, batch);
}

int qmx_sweep(qmx_handle* h, int num_qubits, int batch, int num_layers, const int* layers_used,
              int num_sweeps, const void* target, int storage_dtype, void* gates,
              const int* widths, const void* psi_ref, void* fidelity, double* overlap,
              const double* goal, int* sweeps_done, void* circuit_out, void* workspace,
              size_t workspace_bytes, void* stream) {
  return qmx_sweep_snapshots(h, num_qubits, batch, num_layers, layers_used, num_sweeps, target, storage_dtype,
                             gates, widths, psi_ref, fidelity, overlap, goal, sweeps_done, circuit_out, -1,
                             nullptr, workspace, workspace_bytes, stream);
}

int qmx_sweep_snapshots(qmx_handle* h, int num_qubits, int batch, int num_layers, const int* layers_used,
                        int num_sweeps, const void* target, int storage_dtype, void* gates,
                        const int* widths, const void* psi_ref, void* fidelity, double* overlap,
                        const double* goal, int* sweeps_done, void* circuit_out, int snap_first,
                        void* snapshots, void* workspace, size_t workspace_bytes, void* stream) {
  QMX_CALL(check_common(h, num_qubits, batch));
  if (num_layers < 1 || num_sweeps < 0) return fail(h, QMX_ERR_INVALID, "num_layers / num_sweeps");
  if (!layers_used || !target || !gates || !widths || !workspace)
    return fail(h, QMX_ERR_INVALID, "null pointer");
  if (storage_dtype != QMX_C64 && storage_dtype != QMX_C128) return fail(h, QMX_ERR_INVALID, "storage_dtype");
  if (snapshots && (snap_first < 0 || snap_first + 1 > num_sweeps))
    return fail(h, QMX_ERR_INVALID, "snapshots: 0 <= snap_first < num_sweeps");
  if (workspace_bytes < qmx_sweep_workspace(num_qubits, batch, storage_dtype))
    return fail(h, QMX_ERR_WORKSPACE, "workspace too small");
  QMX_HIP(h, hipSetDevice(h->device));
  hipStream_t st = static_cast<hipStream_t>(stream);
  Carver c(workspace);
  double2* snaps = static_cast<double2*>(snapshots);
  if (storage_dtype == QMX_C64)
    return sweep_dispatch<float2>(h, num_qubits, batch, num_layers, layers_used, num_sweeps,
                                  static_cast<const float2*>(target), static_cast<double2*>(gates), widths,
                                  static_cast<const double2*>(psi_ref), static_cast<double2*>(fidelity),
                                  overlap, goal, sweeps_done, static_cast<float2*>(circuit_out), c, st,
                                  snap_first, snaps);
  return sweep_dispatch<double2>(h, num_qubits, batch, num_layers, layers_used, num_sweeps,
                                 static_cast<const double2*>(target), static_cast<double2*>(gates), widths,
                                 static_cast<const double2*>(psi_ref), static_cast<double2*>(fidelity),
                                 overlap, goal, sweeps_done, static_cast<double2*>(circuit_out), c, st,
                                 snap_first, snaps);
}

size_t qmx_prepare_input_workspace(int num_qubits, int batch) {
  if (num_qubits < 2 || num_qubits > kMaxQubits || batch < 1 || batch > 65535) return 0;
  return prep_input_bytes(num_qubits, batch);
}

int qmx_prepare_input(qmx_handle* h, int num_qubits, int batch, const void* statevectors, int sv_dtype,
                      int index_type, double compression_percentage, void* out, void* workspace,
                      size_t workspace_bytes, void* stream) {
  QMX_CALL(check_common(h, num_qubits, batch));
  if (!statevectors || !out) return fail(h, QMX_ERR_INVALID, "null pointer");
  if (sv_dtype != QMX_C64 && sv_dtype != QMX_C128) return fail(h, QMX_ERR_INVALID, "dtype");
  if (index_type != QMX_INDEX_ROW && index_type != QMX_INDEX_SNAKE)
    return fail(h, QMX_ERR_INVALID, "Index type not supported.");
  if (!std::isfinite(compression_percentage)) return fail(h, QMX_ERR_INVALID, "compression_percentage");
  const size_t len = state_len(num_qubits);
  // int((pct / 100) * 2^N) amplitudes, as the reference computes it; only
  // for pct > 0 (base.py:101)
  size_t cutoff = 0;
  if (compression_percentage > 0.0) {
    const double x = (compression_percentage / 100.0) * static_cast<double>(len);
    if (x >= static_cast<double>(len))
      return fail(h, QMX_ERR_INVALID, "compression_percentage would set every amplitude to zero");
    cutoff = static_cast<size_t>(x);
  }
  if (cutoff > 0 && (!workspace || workspace_bytes < prep_input_bytes(num_qubits, batch)))
    return fail(h, QMX_ERR_WORKSPACE, "workspace too small");
  QMX_HIP(h, hipSetDevice(h->device));
  hipStream_t st = static_cast<hipStream_t>(stream);
  const int snake = index_type == QMX_INDEX_SNAKE ? 1 : 0;
  if (sv_dtype == QMX_C64)
    launch_prep_input(static_cast<const float2*>(statevectors), static_cast<float2*>(out), num_qubits, batch,
                      snake, cutoff, workspace, st);
  else
    launch_prep_input(static_cast<const double2*>(statevectors), static_cast<double2*>(out), num_qubits, batch,
                      snake, cutoff, workspace, st);
  QMX_LAUNCHED(h, "input preprocessing");
  return QMX_OK;
}

namespace {
bool mps_bonds_valid(int num_sites, const int* bond_dims) {
  if (num_sites < 2 || num_sites > kMaxQubits || !bond_dims) return false;
  if (bond_dims[0] != 1 || bond_dims[num_sites] != 1) return false;
  for (int i = 1; i < num_sites; ++i)
    if (bond_dims[i] < 1 || bond_dims[i] > (1 << 20)) return false;
  return true;
}
// largest intermediate X_i (2^i rows x bond_dims[i] columns), i = 1 .. N-1
size_t mps_chain_elems(int num_sites, const int* bond_dims) {
  size_t mx = 1;
  for (int i = 1; i < num_sites; ++i) mx = std::max(mx, ((size_t)1 << i) * (size_t)bond_dims[i]);
  return mx;
}
}  // namespace

size_t qmx_mps_to_state_workspace(int num_sites, const int* bond_dims) {
  if (!mps_bonds_valid(num_sites, bond_dims)) return 0;
  return 2 * mps_chain_elems(num_sites, bond_dims) * sizeof(double2) + 512;
}

int qmx_mps_to_state(qmx_handle* h, int num_sites, const int* bond_dims, const void* tensors, void* state_out,
                     void* workspace, size_t workspace_bytes, void* stream) {
  if (!h) return QMX_ERR_INVALID;
  if (!mps_bonds_valid(num_sites, bond_dims))
    return fail(h, QMX_ERR_INVALID, "bond dimensions: num_sites + 1 entries >= 1, first and last 1");
  if (!tensors || !state_out || !workspace) return fail(h, QMX_ERR_INVALID, "null pointer");
  if (workspace_bytes < qmx_mps_to_state_workspace(num_sites, bond_dims))
    return fail(h, QMX_ERR_WORKSPACE, "workspace too small");
  QMX_HIP(h, hipSetDevice(h->device));
  hipStream_t st = static_cast<hipStream_t>(stream);
  Carver c(workspace);
  const size_t mx = mps_chain_elems(num_sites, bond_dims);
  double2* buf[2] = {c.take<double2>(mx), c.take<double2>(mx)};
  const double2* site = static_cast<const double2*>(tensors);
  // X_1 = A_0 viewed as (2 x r_1): the site tensor itself
  const double2* x = site;
  site += 2 * (size_t)bond_dims[1];
  for (int i = 1; i < num_sites; ++i) {
    const int kd = bond_dims[i], nc = 2 * bond_dims[i + 1];
    double2* y = (i == num_sites - 1) ? static_cast<double2*>(state_out) : buf[i & 1];
    // X_{i+1} (2^i x 2 r_{i+1}) = X_i (2^i x r_i) A_i (r_i x 2 r_{i+1}), in row
    // chunks of at most 2^21 rows (the GEMM grid's row dimension)
    const int chunk_log2 = std::min(i, 21);
    const size_t rows = (size_t)1 << i, chunk = (size_t)1 << chunk_log2;
    for (size_t r0 = 0; r0 < rows; r0 += chunk) {
      GemmDims d{chunk_log2, nullptr, nullptr, 0, kd, nc, nc};
      launch_gemm(x + r0 * kd, 0, site, 0, y + r0 * nc, 0, d, 0, 1, st);
    }
    QMX_LAUNCHED(h, "MPS contraction step");
    site += (size_t)kd * nc;
    x = y;
  }
  return QMX_OK;
}

int qmx_sweep_noalloc(qmx_handle* h, int num_qubits, int batch, int num_layers, const int* layers_used,
                      int num_sweeps, const void* target, int storage_dtype, void* gates,
                      const int* widths, const void* psi_ref, void* fidelity, double* overlap,
                      const double* goal, int* sweeps_done, void* circuit_out, void* stream) {
  QMX_CALL(check_common(h, num_qubits, batch));
  if (storage_dtype != QMX_C64 && storage_dtype != QMX_C128) return fail(h, QMX_ERR_INVALID, "storage_dtype");
  const size_t need = qmx_sweep_workspace(num_qubits, batch, storage_dtype);
  if (need > kNoAllocScratchBytes)
    return fail(h, QMX_ERR_WORKSPACE,
                "qmx_sweep_noalloc: this problem needs " + std::to_string(need) +
                    " bytes of scratch (qmx_sweep_workspace); the library's scratch holds " +
                    std::to_string(kNoAllocScratchBytes) + " bytes per device");
  if (!h->scratch) return hip_fail(h, h->scratch_status, "qmx_sweep_noalloc: scratch address");
  // one stream, the caller's: no internal streams or events are created
  const bool split = h->split_sweeps;
  h->split_sweeps = false;
  const int rc = qmx_sweep(h, num_qubits, batch, num_layers, layers_used, num_sweeps, target, storage_dtype,
                           gates, widths, psi_ref, fidelity, overlap, goal, sweeps_done, circuit_out,
                           h->scratch, kNoAllocScratchBytes, stream);
  h->split_sweeps = split;
  return rc;
}

int qmx_circuit_state(qmx_handle* h, int num_qubits, int batch, int num_layers, const int* layers_used,
                      const void* gates, const int* widths, int storage_dtype, const void* psi_ref,
                      void* fidelity, void* state_out, void* workspace, size_t workspace_bytes,
                      void* stream) {
  if (!state_out) return h ? fail(h, QMX_ERR_INVALID, "state_out is null") : QMX_ERR_INVALID;
  // zero sweeps: only the final C|0..0> and its overlap are computed; the
  // target argument is never read
  return qmx_sweep(h, num_qubits, batch, num_layers, layers_used, 0, state_out, storage_dtype,
                   const_cast<void*>(gates), widths, psi_ref, fidelity, nullptr, nullptr, nullptr,
                   state_out, workspace, workspace_bytes, stream);
}

}  // extern "C"

namespace {
struct EncodeLayout {
  ChiLayout chi;
  LayerLayout layer;
  double2* psi;
  double2* target;
  double2* psi_w;
  void* target_s;
  void* sweep_ws;
  size_t sweep_bytes;
};

EncodeLayout encode_layout(Carver& c, const qmx_encode_config& g) {
  EncodeLayout E{};
  size_t len = state_len(g.num_qubits);
  E.chi = chi_layout(c, g.num_qubits, g.batch, g.bond_dimension, false);
  E.layer = layer_layout(c, g.num_qubits, g.batch);
  E.psi = c.take<double2>(len * g.batch);
  E.target = c.take<double2>(len * g.batch);
  E.psi_w = c.take<double2>(len * g.batch);
  E.target_s = (g.sweep_dtype == QMX_C64) ? static_cast<void*>(c.take<float2>(len * g.batch))
                                          : static_cast<void*>(E.target);
  E.sweep_bytes = qmx_sweep_workspace(g.num_qubits, g.batch, g.sweep_dtype);
  E.sweep_ws = c.take<char>(E.sweep_bytes);
  return E;
}
}  // namespace

extern "C" {

size_t qmx_encode_workspace(const qmx_encode_config* cfg) {
  if (!cfg || cfg->num_qubits < 2 || cfg->num_qubits > kMaxQubits || cfg->batch < 1 ||
      cfg->bond_dimension < 1 || cfg->num_layers < 1 || cfg->num_sweeps < 0 ||
      (cfg->sv_dtype != QMX_C64 && cfg->sv_dtype != QMX_C128) ||
      (cfg->sweep_dtype != QMX_C64 && cfg->sweep_dtype != QMX_C128))
    return 0;
  Carver c(nullptr);
  encode_layout(c, *cfg);
  return c.off + 256;
}

}  // extern "C"

namespace {
int encode_check(qmx_handle* h, const qmx_encode_config* cfg, size_t workspace_bytes) {
  if (!h || !cfg) return QMX_ERR_INVALID;
  const qmx_encode_config g = *cfg;
  QMX_CALL(check_common(h, g.num_qubits, g.batch));
  if (g.bond_dimension < 1) return fail(h, QMX_ERR_INVALID, "`bond_dimension` must be an integer greater than 0.");
  if (g.num_layers < 1) return fail(h, QMX_ERR_INVALID, "The number of layers must be a positive integer.");
  if (!chi_supported(g.num_qubits, g.bond_dimension))
    return fail(h, QMX_ERR_INVALID, "bond_dimension above 512 with truncation is not supported");
  if (g.num_sweeps < 0) return fail(h, QMX_ERR_INVALID, "num_sweeps must be >= 0");
  if ((g.sv_dtype != QMX_C64 && g.sv_dtype != QMX_C128) ||
      (g.sweep_dtype != QMX_C64 && g.sweep_dtype != QMX_C128))
    return fail(h, QMX_ERR_INVALID, "dtype");
  if (workspace_bytes < qmx_encode_workspace(cfg)) return fail(h, QMX_ERR_WORKSPACE, "workspace too small");
  return QMX_OK;
}

// chi-truncation, normalisation, layer extraction, conversion of the target
// to the sweep storage type: everything of prepare_state before the sweeps
int encode_prepare(qmx_handle* h, const qmx_encode_config& g, const void* statevectors, void* gates,
                   int* widths, int* layers_used, const EncodeLayout& E, hipStream_t st) {
  const int n = g.num_qubits, B = g.batch;
  QMX_CALL(run_chi(h, n, B, g.bond_dimension, g.cutoff, statevectors, g.sv_dtype, E.psi, E.target,
                   nullptr, E.chi, st));
  QMX_CALL(device_error_status(h, E.chi.error, st));
  // psi_w = target / |target|
  int nb = static_cast<int>(std::min<size_t>(E.chi.nblk_norm, std::max<size_t>(1, state_len(n) / 1024)));
  launch_fill_int(E.chi.ncnt, 0, B, st);
  launch_norm2(static_cast<const double2*>(E.target), n, B, E.chi.npart, E.chi.ncnt, E.chi.nrm, nb, st);
  launch_scale_convert(static_cast<const double2*>(E.target), E.psi_w, n, B, E.chi.nrm, st);
  QMX_LAUNCHED(h, "normalise target");
  QMX_CALL(run_extract(h, n, B, g.num_layers, g.fidelity_threshold, g.cutoff, E.psi_w,
                       static_cast<double2*>(gates), widths, nullptr, layers_used, nullptr,
                       E.layer, st));
  if (g.sweep_dtype == QMX_C64)
    launch_scale_convert(static_cast<const double2*>(E.target), static_cast<float2*>(E.target_s), n,
                         B, nullptr, st);
  QMX_LAUNCHED(h, "target conversion");
  return QMX_OK;
}
}  // namespace

extern "C" {

int qmx_encode(qmx_handle* h, const qmx_encode_config* cfg, const void* statevectors,
               void* gates, int* widths, int* layers_used, void* fidelity, const double* goal,
               int* sweeps_done, void* workspace, size_t workspace_bytes, void* stream) {
  QMX_CALL(encode_check(h, cfg, workspace_bytes));
  if (!statevectors || !gates || !widths || !layers_used || !fidelity || !workspace)
    return fail(h, QMX_ERR_INVALID, "null pointer");
  const qmx_encode_config g = *cfg;
  QMX_HIP(h, hipSetDevice(h->device));
  hipStream_t st = static_cast<hipStream_t>(stream);
  Carver c(workspace);
  EncodeLayout E = encode_layout(c, g);
  QMX_CALL(encode_prepare(h, g, statevectors, gates, widths, layers_used, E, st));
  return qmx_sweep(h, g.num_qubits, g.batch, g.num_layers, layers_used, g.num_sweeps, E.target_s,
                   g.sweep_dtype, gates, widths, E.psi, fidelity, nullptr, goal, sweeps_done, nullptr,
                   E.sweep_ws, E.sweep_bytes, stream);
}

int qmx_encode_prepare(qmx_handle* h, const qmx_encode_config* cfg, const void* statevectors, void* gates,
                       int* widths, int* layers_used, void* psi_out, void* target_out, void* workspace,
                       size_t workspace_bytes, void* stream) {
  QMX_CALL(encode_check(h, cfg, workspace_bytes));
  if (!statevectors || !gates || !widths || !layers_used || !psi_out || !target_out || !workspace)
    return fail(h, QMX_ERR_INVALID, "null pointer");
  const qmx_encode_config g = *cfg;
  QMX_HIP(h, hipSetDevice(h->device));
  hipStream_t st = static_cast<hipStream_t>(stream);
  Carver c(workspace);
  EncodeLayout E = encode_layout(c, g);
  QMX_CALL(encode_prepare(h, g, statevectors, gates, widths, layers_used, E, st));
  const size_t amps = state_len(g.num_qubits) * g.batch;
  QMX_HIP(h, hipMemcpyAsync(psi_out, E.psi, amps * sizeof(double2), hipMemcpyDeviceToDevice, st));
  QMX_HIP(h, hipMemcpyAsync(target_out, E.target_s,
                            amps * (g.sweep_dtype == QMX_C64 ? sizeof(float2) : sizeof(double2)),
                            hipMemcpyDeviceToDevice, st));
  return QMX_OK;
}

// ------------------------------------------------ grouped bond dimensions --
// One encode of a batch of equal-size states whose bond dimensions differ:
// states [first[k], first[k] + count[k]) share bond dimension chi[k] (the
// groups tile the batch in order).  Only the chi-truncation depends on the
// bond dimension, so it runs per group; normalisation, layer extraction and the
// sweeps run once over the whole batch.  Used by qmx_encode_ragged (one call
// per qubit count instead of one per (qubit count, bond dimension)).
}  // extern "C"

namespace {
struct GroupedLayout {
  std::vector<ChiLayout> chi;
  double* nrm;
  double* npart;
  int* ncnt;
  LayerLayout layer;
  double2* psi;
  double2* target;
  double2* psi_w;
  void* target_s;
  void* sweep_ws;
  size_t sweep_bytes;
};

constexpr int kGroupedNormBlocks = 256;

GroupedLayout grouped_layout(Carver& c, const qmx_encode_config& g, const int* count, const int* chi, int ng) {
  GroupedLayout E{};
  const size_t len = state_len(g.num_qubits);
  for (int k = 0; k < ng; ++k) E.chi.push_back(chi_layout(c, g.num_qubits, count[k], chi[k], false));
  E.nrm = c.take<double>(g.batch);
  E.npart = c.take<double>(static_cast<size_t>(g.batch) * kGroupedNormBlocks * 2);
  E.ncnt = c.take<int>(g.batch);
  E.layer = layer_layout(c, g.num_qubits, g.batch);
  E.psi = c.take<double2>(len * g.batch);
  E.target = c.take<double2>(len * g.batch);
  E.psi_w = c.take<double2>(len * g.batch);
  E.target_s = (g.sweep_dtype == QMX_C64) ? static_cast<void*>(c.take<float2>(len * g.batch))
                                          : static_cast<void*>(E.target);
  E.sweep_bytes = qmx_sweep_workspace(g.num_qubits, g.batch, g.sweep_dtype);
  E.sweep_ws = c.take<char>(E.sweep_bytes);
  return E;
}

bool groups_valid(const qmx_encode_config& g, const int* first, const int* count, const int* chi, int ng) {
  if (ng < 1 || !first || !count || !chi) return false;
  int next = 0;
  for (int k = 0; k < ng; ++k) {
    if (first[k] != next || count[k] < 1 || chi[k] < 1 || !chi_supported(g.num_qubits, chi[k])) return false;
    next += count[k];
  }
  return next == g.batch;
}
}  // namespace

extern "C" {

size_t qmx_encode_grouped_workspace(const qmx_encode_config* cfg, const int* first, const int* count,
                                    const int* chi, int ngroups) {
  if (qmx_encode_workspace(cfg) == 0 || !groups_valid(*cfg, first, count, chi, ngroups)) return 0;
  Carver c(nullptr);
  grouped_layout(c, *cfg, count, chi, ngroups);
  return c.off + 256;
}

int qmx_encode_grouped(qmx_handle* h, const qmx_encode_config* cfg, const int* first, const int* count,
                       const int* chi, int ngroups, const void* statevectors, void* gates, int* widths,
                       int* layers_used, void* fidelity, void* workspace, size_t workspace_bytes,
                       void* stream) {
  if (!h || !cfg) return QMX_ERR_INVALID;
  const qmx_encode_config g = *cfg;
  QMX_CALL(check_common(h, g.num_qubits, g.batch));
  if (g.num_layers < 1) return fail(h, QMX_ERR_INVALID, "The number of layers must be a positive integer.");
  if (g.num_sweeps < 0) return fail(h, QMX_ERR_INVALID, "num_sweeps must be >= 0");
  if ((g.sv_dtype != QMX_C64 && g.sv_dtype != QMX_C128) ||
      (g.sweep_dtype != QMX_C64 && g.sweep_dtype != QMX_C128))
    return fail(h, QMX_ERR_INVALID, "dtype");
  if (!groups_valid(g, first, count, chi, ngroups))
    return fail(h, QMX_ERR_INVALID, "bond-dimension groups must tile the batch (bond_dimension <= 512 when it truncates)");
  if (!statevectors || !gates || !widths || !layers_used || !fidelity || !workspace)
    return fail(h, QMX_ERR_INVALID, "null pointer");
  if (workspace_bytes < qmx_encode_grouped_workspace(cfg, first, count, chi, ngroups))
    return fail(h, QMX_ERR_WORKSPACE, "workspace too small");
  QMX_HIP(h, hipSetDevice(h->device));
  hipStream_t st = static_cast<hipStream_t>(stream);
  Carver c(workspace);
  GroupedLayout E = grouped_layout(c, g, count, chi, ngroups);
  const int n = g.num_qubits, B = g.batch;
  const size_t len = state_len(n);
  const size_t svb = len * (g.sv_dtype == QMX_C64 ? sizeof(float2) : sizeof(double2));
  for (int k = 0; k < ngroups; ++k) {
    const size_t o = static_cast<size_t>(first[k]);
    QMX_CALL(run_chi(h, n, count[k], chi[k], g.cutoff, static_cast<const char*>(statevectors) + o * svb,
                     g.sv_dtype, E.psi + o * len, E.target + o * len, nullptr, E.chi[k], st));
    QMX_CALL(device_error_status(h, E.chi[k].error, st));
  }
  // psi_w = target / |target|, then as encode_prepare
  int nb = static_cast<int>(std::min<size_t>(kGroupedNormBlocks, std::max<size_t>(1, len / 1024)));
  launch_fill_int(E.ncnt, 0, B, st);
  launch_norm2(static_cast<const double2*>(E.target), n, B, E.npart, E.ncnt, E.nrm, nb, st);
  launch_scale_convert(static_cast<const double2*>(E.target), E.psi_w, n, B, E.nrm, st);
  QMX_LAUNCHED(h, "normalise target");
  QMX_CALL(run_extract(h, n, B, g.num_layers, g.fidelity_threshold, g.cutoff, E.psi_w,
                       static_cast<double2*>(gates), widths, nullptr, layers_used, nullptr, E.layer, st));
  if (g.sweep_dtype == QMX_C64)
    launch_scale_convert(static_cast<const double2*>(E.target), static_cast<float2*>(E.target_s), n, B,
                         nullptr, st);
  QMX_LAUNCHED(h, "target conversion");
  return qmx_sweep(h, n, B, g.num_layers, layers_used, g.num_sweeps, E.target_s, g.sweep_dtype, gates, widths,
                   E.psi, fidelity, nullptr, nullptr, nullptr, nullptr, E.sweep_ws, E.sweep_bytes, stream);
}

// ---------------------------------------------------- learned warm start --
size_t qmx_mlp_param_count(int in_dim, int width, int out_dim) {
  if (!mlp_dims_ok(1, in_dim, width, out_dim)) return 0;
  return MlpDims{in_dim, width, out_dim}.count();
}

size_t qmx_mlp_workspace(int rows, int in_dim, int width, int out_dim) {
  if (!mlp_dims_ok(rows, in_dim, width, out_dim)) return 0;
  Carver c(nullptr);
  mlp_layout(c, rows, MlpDims{in_dim, width, out_dim});
  return c.off + 256;
}

int qmx_mlp_forward(qmx_handle* h, int rows, int in_dim, int width, int out_dim, const void* x_bf16,
                    const float* params, const void* params_bf16, float* y, void* workspace,
                    size_t workspace_bytes, void* stream) {
  if (!h) return QMX_ERR_INVALID;
  if (!mlp_dims_ok(rows, in_dim, width, out_dim) || !x_bf16 || !params || !params_bf16 || !y)
    return fail(h, QMX_ERR_INVALID, "mlp_forward: bad arguments");
  if (!workspace || workspace_bytes < qmx_mlp_workspace(rows, in_dim, width, out_dim))
    return fail(h, QMX_ERR_WORKSPACE, "mlp_forward: workspace too small");
  QMX_HIP(h, hipSetDevice(h->device));
  hipStream_t st = static_cast<hipStream_t>(stream);
  const MlpDims d{in_dim, width, out_dim};
  Carver c(workspace);
  MlpLayout L = mlp_layout(c, rows, d);
  const uint16_t* x = static_cast<const uint16_t*>(x_bf16);
  const uint16_t* pb = static_cast<const uint16_t*>(params_bf16);
  MlpGemmArgs g = gemm(rows, width, in_dim, x, in_dim, 0, pb + d.oW1(), width, 0);
  g.bias = params + d.ob1(); g.relu = 1; g.Cb = L.a1; g.ldcb = width;
  launch_mlp_gemm(g, st);
  g = gemm(rows, width, width, L.a1, width, 0, pb + d.oW2(), width, 0);
  g.bias = params + d.ob2(); g.relu = 1; g.Cb = L.a2; g.ldcb = width;
  launch_mlp_gemm(g, st);
  g = gemm(rows, out_dim, width, L.a2, width, 0, pb + d.oW3(), out_dim, 0);
  g.bias = params + d.ob3(); g.C = y; g.ldc = out_dim;
  launch_mlp_gemm(g, st);
  QMX_LAUNCHED(h, "mlp forward");
  return QMX_OK;
}

int qmx_mlp_backward(qmx_handle* h, int rows, int in_dim, int width, int out_dim, const void* x_bf16,
                     const float* params, const void* params_bf16, const float* dy, float* grads,
                     void* workspace, size_t workspace_bytes, void* stream) {
  (void)params;
  if (!h) return QMX_ERR_INVALID;
  if (!mlp_dims_ok(rows, in_dim, width, out_dim) || !x_bf16 || !params_bf16 || !dy || !grads)
    return fail(h, QMX_ERR_INVALID, "mlp_backward: bad arguments");
  if (!workspace || workspace_bytes < qmx_mlp_workspace(rows, in_dim, width, out_dim))
    return fail(h, QMX_ERR_WORKSPACE, "mlp_backward: workspace too small");
  QMX_HIP(h, hipSetDevice(h->device));
  hipStream_t st = static_cast<hipStream_t>(stream);
  const MlpDims d{in_dim, width, out_dim};
  Carver c(workspace);
  MlpLayout L = mlp_layout(c, rows, d);
  const uint16_t* x = static_cast<const uint16_t*>(x_bf16);
  const uint16_t* pb = static_cast<const uint16_t*>(params_bf16);
  launch_to_bf16(dy, L.dyb, (size_t)rows * out_dim, st);
  // layer 3
  MlpGemmArgs g = gemm(width, out_dim, rows, L.a2, width, 1, L.dyb, out_dim, 0);
  g.C = grads + d.oW3(); g.ldc = out_dim;
  launch_mlp_gemm(g, st);
  launch_colsum_bf16(L.dyb, rows, out_dim, out_dim, grads + d.ob3(), st);
  g = gemm(rows, width, out_dim, L.dyb, out_dim, 0, pb + d.oW3(), out_dim, 1);
  g.mask = L.a2; g.ldm = width; g.Cb = L.dz2; g.ldcb = width;
  launch_mlp_gemm(g, st);
  // layer 2
  g = gemm(width, width, rows, L.a1, width, 1, L.dz2, width, 0);
  g.C = grads + d.oW2(); g.ldc = width;
  launch_mlp_gemm(g, st);
  launch_colsum_bf16(L.dz2, rows, width, width, grads + d.ob2(), st);
  g = gemm(rows, width, width, L.dz2, width, 0, pb + d.oW2(), width, 1);
  g.mask = L.a1; g.ldm = width; g.Cb = L.dz1; g.ldcb = width;
  launch_mlp_gemm(g, st);
  // layer 1
  g = gemm(in_dim, width, rows, x, in_dim, 1, L.dz1, width, 0);
  g.C = grads + d.oW1(); g.ldc = width;
  launch_mlp_gemm(g, st);
  launch_colsum_bf16(L.dz1, rows, width, width, grads + d.ob1(), st);
  QMX_LAUNCHED(h, "mlp backward");
  return QMX_OK;
}

int qmx_mlp_adam(qmx_handle* h, size_t count, float* params, void* params_bf16, const float* grads, float* m,
                 float* v, float lr, float beta1, float beta2, float eps, int step, void* stream) {
  if (!h) return QMX_ERR_INVALID;
  if (count == 0 || !params || !grads || !m || !v || step < 1)
    return fail(h, QMX_ERR_INVALID, "mlp_adam: bad arguments");
  QMX_HIP(h, hipSetDevice(h->device));
  launch_adam(params, static_cast<uint16_t*>(params_bf16), grads, m, v, count, lr, beta1, beta2, eps, step,
              static_cast<hipStream_t>(stream));
  QMX_LAUNCHED(h, "mlp adam");
  return QMX_OK;
}

int qmx_mlp_to_bf16(qmx_handle* h, size_t count, const float* src, void* dst, void* stream) {
  if (!h) return QMX_ERR_INVALID;
  if (count == 0 || !src || !dst) return fail(h, QMX_ERR_INVALID, "to_bf16: bad arguments");
  QMX_HIP(h, hipSetDevice(h->device));
  launch_to_bf16(src, static_cast<uint16_t*>(dst), count, static_cast<hipStream_t>(stream));
  QMX_LAUNCHED(h, "to_bf16");
  return QMX_OK;
}

size_t qmx_mse_workspace(void) { return kMseBlocks * sizeof(float) + 256; }

int qmx_mse_loss(qmx_handle* h, int rows, int ld, int valid, const float* y, const float* t, const float* w,
                 float inv_norm, float* dy, float* loss, void* workspace, size_t workspace_bytes,
                 void* stream) {
  if (!h) return QMX_ERR_INVALID;
  if (rows < 1 || ld < 1 || valid < 1 || valid > ld || !y || !t || !dy || !loss)
    return fail(h, QMX_ERR_INVALID, "mse: bad arguments");
  if (!workspace || workspace_bytes < qmx_mse_workspace())
    return fail(h, QMX_ERR_WORKSPACE, "mse: workspace too small");
  QMX_HIP(h, hipSetDevice(h->device));
  launch_mse(y, t, w, rows, ld, valid, inv_norm, dy, static_cast<float*>(workspace), kMseBlocks, loss,
             static_cast<hipStream_t>(stream));
  QMX_LAUNCHED(h, "mse");
  return QMX_OK;
}

int qmx_gate_features(qmx_handle* h, int num_qubits, int batch, int num_layers, const int* layers_used,
                      const void* gates_a, const void* gates_b, const int* widths, void* features,
                      int ld, float* row_weight, void* stream) {
  if (!h) return QMX_ERR_INVALID;
  GateRowArgs r{};
  QMX_CALL(gate_rows_check(h, num_qubits, batch, num_layers, r, layers_used, widths));
  if (!gates_a || !gates_b || !features || ld < kGateFeatures)
    return fail(h, QMX_ERR_INVALID, "gate_features: bad arguments");
  QMX_HIP(h, hipSetDevice(h->device));
  launch_gate_features(r, static_cast<const double2*>(gates_a), static_cast<const double2*>(gates_b),
                       static_cast<uint16_t*>(features), ld, row_weight, static_cast<hipStream_t>(stream));
  QMX_LAUNCHED(h, "gate features");
  return QMX_OK;
}

int qmx_gate_log(qmx_handle* h, int num_qubits, int batch, int num_layers, const int* layers_used,
                 const void* gates_a, const void* gates_b, const int* widths, float* coeffs, int ld,
                 void* stream) {
  if (!h) return QMX_ERR_INVALID;
  GateRowArgs r{};
  QMX_CALL(gate_rows_check(h, num_qubits, batch, num_layers, r, layers_used, widths));
  if (!gates_a || !gates_b || !coeffs || ld < 15) return fail(h, QMX_ERR_INVALID, "gate_log: bad arguments");
  QMX_HIP(h, hipSetDevice(h->device));
  launch_gate_log(r, static_cast<const double2*>(gates_a), static_cast<const double2*>(gates_b), coeffs, ld,
                  static_cast<hipStream_t>(stream));
  QMX_LAUNCHED(h, "gate log");
  return QMX_OK;
}

int qmx_traj_features(qmx_handle* h, int num_qubits, int batch, int num_layers, const int* layers_used,
                      const void* gates_a, const void* gates_b, const void* gates_c, const int* widths,
                      float sweep_frac, void* features, int ld, float* row_weight, float* row_scale,
                      void* stream) {
  if (!h) return QMX_ERR_INVALID;
  GateRowArgs r{};
  QMX_CALL(gate_rows_check(h, num_qubits, batch, num_layers, r, layers_used, widths));
  if (!gates_a || !gates_b || !gates_c || !features || !row_scale || ld < kTrajFeatures)
    return fail(h, QMX_ERR_INVALID, "traj_features: bad arguments");
  QMX_HIP(h, hipSetDevice(h->device));
  launch_traj_features(r, static_cast<const double2*>(gates_a), static_cast<const double2*>(gates_b),
                       static_cast<const double2*>(gates_c), static_cast<uint16_t*>(features), ld, row_weight,
                       row_scale, sweep_frac, static_cast<hipStream_t>(stream));
  QMX_LAUNCHED(h, "trajectory features");
  return QMX_OK;
}
int qmx_traj_targets(qmx_handle* h, int num_qubits, int batch, int num_layers, const int* layers_used,
                     const void* gates_a, const void* gates_b, const void* gates_c, const void* gates_d,
                     const int* widths, float* targets, int ld, void* stream) {
  if (!h) return QMX_ERR_INVALID;
  GateRowArgs r{};
  QMX_CALL(gate_rows_check(h, num_qubits, batch, num_layers, r, layers_used, widths));
  if (!gates_a || !gates_b || !gates_c || !gates_d || !targets || ld < 2)
    return fail(h, QMX_ERR_INVALID, "traj_targets: bad arguments");
  QMX_HIP(h, hipSetDevice(h->device));
  launch_traj_targets(r, static_cast<const double2*>(gates_a), static_cast<const double2*>(gates_b),
                      static_cast<const double2*>(gates_c), static_cast<const double2*>(gates_d), targets, ld,
                      static_cast<hipStream_t>(stream));
  QMX_LAUNCHED(h, "trajectory targets");
  return QMX_OK;
}
int qmx_traj_jump(qmx_handle* h, int num_qubits, int batch, int num_layers, const int* layers_used,
                  const void* gates_a, const void* gates_b, const float* coeffs, int ld, void* gates,
                  const int* widths, const int* skip, void* stream) {
  if (!h) return QMX_ERR_INVALID;
  GateRowArgs r{};
  QMX_CALL(gate_rows_check(h, num_qubits, batch, num_layers, r, layers_used, widths));
  if (!gates_a || !gates_b || !coeffs || !gates || ld < 2) return fail(h, QMX_ERR_INVALID, "traj_jump: bad arguments");
  QMX_HIP(h, hipSetDevice(h->device));
  launch_traj_jump(r, static_cast<const double2*>(gates_a), static_cast<const double2*>(gates_b), coeffs, ld,
                   static_cast<double2*>(gates), skip, static_cast<hipStream_t>(stream));
  QMX_LAUNCHED(h, "trajectory jump");
  return QMX_OK;
}
int qmx_gate_expmap(qmx_handle* h, int num_qubits, int batch, int num_layers, const int* layers_used,
                    const float* coeffs, int ld, void* gates, const int* widths, const int* skip,
                    void* stream) {
  if (!h) return QMX_ERR_INVALID;
  GateRowArgs r{};
  QMX_CALL(gate_rows_check(h, num_qubits, batch, num_layers, r, layers_used, widths));
  if (!coeffs || !gates || ld < 15) return fail(h, QMX_ERR_INVALID, "gate_expmap: bad arguments");
  QMX_HIP(h, hipSetDevice(h->device));
  launch_gate_expmap(r, coeffs, ld, static_cast<double2*>(gates), skip, static_cast<hipStream_t>(stream));
  QMX_LAUNCHED(h, "gate expmap");
  return QMX_OK;
}

}  // extern "C"

