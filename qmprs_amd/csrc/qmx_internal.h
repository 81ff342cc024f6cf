// Internal (host + device) declarations of the qmx library: kernel argument
// blocks, limits and launchers.  Not part of the C-ABI.
#pragma once
#include <hip/hip_runtime.h>
#include <stddef.h>
#include <stdint.h>

namespace qmx {

constexpr int kMaxQubits = 30;
constexpr int kMaxPassGates = 16;
constexpr int kMaxHighBits = 8;
constexpr int kApplyTileBits = 11;   // 2^11 complex128 = 32 KiB LDS
constexpr int kSweepTileBits = 10;   // two tiles of 2^10 complex128 = 32 KiB
constexpr int kApplyMaxUnion = 8;    // gate bits per pass (>= 3 coalescing bits; 8 high bits max)
constexpr double kGsTol = 1e-7;      // canonical-gauge Gram-Schmidt threshold
constexpr double kPhaseTol = 1e-12;  // Ran two-site phase fix skipped below this

constexpr int kWinTileBits = 11;     // window pass tile: 2^11 amplitudes
constexpr int kMaxWindow = 8;        // gates per window
constexpr int kWinMinQubits = 10;    // windowed sweep: 4 waves x k-steps of 4 rest indices

// allocation-free entry point: library-owned scratch per device (qmx_scratch.hip)
constexpr size_t kNoAllocScratchBytes = size_t(64) << 20;
constexpr int kMaxDevices = 64;
hipError_t scratch_address(int device, void** out);

// error bits accumulated on the device
constexpr int kErrNotUnitary = 1;
constexpr int kErrBasis = 2;
constexpr int kErrEigNoConv = 4;

template <typename S>
struct PassArgs {
  int n, lmax, tbits;
  int mode;            // 0: inverse layer `layer` (pos = sites); 1: forward positions
  int layer;
  int count;
  int pos[kMaxPassGates];
  S* psi;
  const double2* gates;
  const int* widths;
  const int* layers;   // per-state layer count (mode 1); nullptr -> lmax
  const int* skip;     // per-state skip flag, may be nullptr
  const double2* ref;  // optional overlap reference (complex128)
  double* dot_part;
  int* dot_counter;
  double2* dot_out;    // dot_out[b * dot_stride]
  int dot_stride;
};

template <typename S>
struct SweepArgs {
  int n, lmax, tbits, step;
  const S* target;
  S* tprime;
  S* phi;
  double2* gates;
  const int* widths;
  const int* layers;
  const int* skip;
  double* part;
  int* counter;
  double* overlap;     // overlap[b * overlap_stride], may be nullptr
  int overlap_stride;
  double omega_m1;     // over-relaxation of the gate update minus one (0: plain)
};

// Host-computed tile geometry.  A tile is a set P of t bits (gate bits plus
// the lowest free bits); tiles partition the state and tile id k sits at the
// deposit of k into the bits outside P (qmask).
//
// A 4-bit group set U inside a tile: element x of group g (bit m of x <->
// m-th lowest bit of U) sits at global offset base + dep(g, restg) + xg(x)
// and at tile-local index dep(g, restl) + xl(x), where restg / restl are the
// tile bits outside U as global / tile-local masks and xg / xl deposit x
// into ubit / ux.
struct GroupMap {
  uint64_t mask;       // the 4 global bits
  uint64_t restg;      // tile bits outside U (global positions)
  uint32_t restl;      // the same bits as tile-local positions
  int ubit[4];         // global positions of U's bits (ascending)
  int ux[4];           // their tile-local positions
};

// Lane roles of a stage in the matrix-core forward pass (2^12-amplitude
// tiles: 8 rest bits, 4 gate bits).  Lane jl + 16 kq of wave w handles, for
// column block cb, the group whose rest bits are jl at jl*, cb and w at h*
// (cb in h*[0..1], w in h*[2..3]), and amplitudes whose gate bits are kq at
// k* and the k-step s (inputs) or the item r (outputs) at s*.  *l: tile-local
// positions, *g: global positions, kx / sx: indices into the stage's
// ux / ubit.  The host picks them so that the 32 lanes of each LDS access
// phase fall on distinct banks.  Each word holds four positions, one per byte
// (32-bit words: the kernel reads them with scalar loads).
struct MfmaMap {
  uint32_t jl, hl, ksl;  // jl[0..3]; hl[0..3]; kl[0], kl[1], sl[0], sl[1]
  uint32_t jg, hg, ksg;  // the same as global positions
  uint32_t kxsx;         // kx[0], kx[1], sx[0], sx[1]
};

// Windowed sweep (qmx_window.hip): window = sweep steps [s0, s1) on the
// 4-bit set umask; the previous window [p0, p1) (p1 == s0, empty for the
// first) has its updated gates applied to T' in the same pass.
template <typename S>
struct WinPassArgs {
  int n, lmax, t;
  int s0, s1, p0, p1;
  uint64_t qmask;      // bits outside the tile
  GroupMap gu;         // this window's set (phi groups, R)
  GroupMap gp;         // previous window's set (T' groups; == gu for the first)
  const S* target;
  S* tprime;
  S* phi;
  const double2* gates;
  const int* widths;
  const int* layers;
  const int* skip;
  const int* list;   // grid row -> state (live states only), or null: row = state
  double* part;        // [B][nblk][512]
  int staged;          // copy tiles through LDS in tile order (see window_pass_kernel)
  int store;           // write T' and phi back (0 for a sweep's last window: neither is read again)
};

// Forward circuit pass (qmx_window.hip): up to kMaxStages stages of gates,
// each stage acting inside its own 4-bit group set, exchanged through LDS.
// Stage k applies the gates at forward positions fpos[gstart[k] .. gstart[k+1]).
constexpr int kMaxStages = 12;
constexpr int kMaxFwdGates = 64;
constexpr int kFwdTileBits = 12;
// The forward pass runs on the matrix cores (fwd_mfma_kernel for complex64,
// fwd_mfma64_kernel for complex128) when enabled, on tiles of 2^kFwdTileBits
// amplitudes, while 32-bit byte offsets reach every amplitude.
inline bool fwd_mfma_applies(int enabled, int t, int n, size_t amp_bytes) {
  return enabled && t == kFwdTileBits && n <= (amp_bytes == 8 ? 28 : 27);
}
constexpr int kFwdMaxUnion = 8;      // gate bits per pass (>= 4 coalescing bits)
template <typename S>
struct FwdArgs {
  int n, lmax, t, nstage;
  int from_basis0;     // the input is |0..0> (psi is not read)
  int mfma;            // tiles of 2^kFwdTileBits: stage matrices on the matrix cores (fwd_mfma_applies)
  int full;            // write every amplitude (else only those inside live_out)
  uint64_t tmask;      // tile-index bits visited (state bits outside the tile)
  uint64_t live_in;    // amplitudes with a bit outside live_in are zero before the pass
  uint64_t live_out;   // ... and outside live_out after it
  GroupMap gs[kMaxStages];
  MfmaMap mm[kMaxStages];  // lane roles of the matrix-core pass (mfma only)
  int gstart[kMaxStages + 1];
  int fpos[kMaxFwdGates];
  S* psi;
  const double2* gates;
  const int* widths;
  const int* layers;
  const int* skip;
  const int* list;   // grid row -> state (live states only), or null: row = state
};

struct WinFinArgs {
  int n, lmax, nblk;
  int pstride;         // partial slots summed: 0, pstride, 2 pstride, ... (< nblk)
  int s0, s1;
  uint64_t umask;
  double2* gates;
  const int* widths;
  const int* layers;
  const int* skip;
  const int* list;   // grid row -> state (live states only), or null: row = state
  const double* part;
  double* overlap;
  int overlap_stride;
  double omega_m1;     // over-relaxation of the gate update minus one (0: plain)
};

struct Chain2Args {
  int n, step;
  double cutoff;
  const double2* psi;  // X_N (normalised working state)
  double2* xa;
  double2* xb;
  int* r;              // [B][n+1]
  double2* w;          // [B][n+1][16]
  double* spectra;     // [B][n+1][4] or nullptr
  const int* skip;
  double* part;
  int* counter;
  int* error;
};

struct RanArgs {
  int n, lmax, layer;
  const double2* psi;
  const double2* xa;
  const double2* xb;
  const int* r;
  const double2* w;
  double2* gates;
  int* widths;
  int* bonds;          // optional [B][lmax][n]
  const int* skip;
  int* error;
  double unitary_atol;  // allclose atol of the unitarity check (1e-8; < 0 fails every gate)
};

struct Amp0Args {
  int n, batch, lmax, layer;
  double atol;
  const double2* psi;
  int* skip;
  int* layers;
  double2* amp0;       // optional [B][lmax]
};

// Buffer holding X_j (2^j x r_j) in the bond-2 chain (j <= N-1).
inline __host__ __device__ int chain_buf_is_a(int n, int j) { return ((n - j) & 1) == 1; }

// launchers (qmx_kernels.hip)
// BF16 GEMM with fused epilogue (qmx_mlp.hip)
struct MlpGemmArgs {
  int M, N, K;
  const uint16_t* A;
  int lda, transA;
  const uint16_t* B;
  int ldb, transB;
  const float* bias;       // [N] or nullptr
  int relu;
  const uint16_t* mask;    // keep where mask > 0 (BF16 [M x ldm]) or nullptr
  int ldm;
  float* C;                // FP32 out or nullptr
  int ldc;
  uint16_t* Cb;            // BF16 out or nullptr
  int ldcb;
};

// gate slots (b, layer, site < n - 1) as rows of a feature matrix; a row
// holds kGateFeatures values (see gate_features_kernel)
constexpr int kGateFeatures = 50;
struct GateRowArgs {
  int n, lmax, rows;
  const int* layers;
  const int* widths;
};

void launch_mlp_gemm(const MlpGemmArgs&, hipStream_t);
void launch_colsum_bf16(const uint16_t* dz, int M, int N, int ld, float* db, hipStream_t);
void launch_to_bf16(const float* x, uint16_t* y, size_t n, hipStream_t);
void launch_adam(float* p, uint16_t* pb, const float* g, float* m, float* v, size_t n, float lr, float b1,
                 float b2, float eps, int step, hipStream_t);
void launch_mse(const float* y, const float* t, const float* w, int rows, int ld, int valid, float inv_norm,
                float* dy, float* part, int nblk, float* loss, hipStream_t);
void launch_gate_features(const GateRowArgs&, const double2* ga, const double2* gb, uint16_t* feat, int ld,
                          float* weight, hipStream_t);
void launch_gate_log(const GateRowArgs&, const double2* ga, const double2* gb, float* out, int ld, hipStream_t);
void launch_gate_expmap(const GateRowArgs&, const float* coef, int ld, double2* g, const int* skip, hipStream_t);
constexpr int kTrajFeatures = 35;   // used columns of a trajectory feature row
void launch_traj_features(const GateRowArgs&, const double2* ga, const double2* gb, const double2* gc,
                          uint16_t* feat, int ld, float* weight, float* scale, float sweep_frac, hipStream_t);
void launch_traj_targets(const GateRowArgs&, const double2* ga, const double2* gb, const double2* gc,
                         const double2* gd, float* out, int ld, hipStream_t);
void launch_traj_jump(const GateRowArgs&, const double2* ga, const double2* gb, const float* coef, int ld,
                      double2* g, const int* skip, hipStream_t);

template <typename S> void launch_apply_pass(const PassArgs<S>&, int nblk, int batch, hipStream_t);
template <typename S> void launch_sweep_step(const SweepArgs<S>&, int nblk, int batch, hipStream_t);
template <typename S>
void launch_norm2(const S* x, int n, int batch, double* part, int* counter, double* out, int nblk,
                  hipStream_t);
template <typename S, typename D>
void launch_scale_convert(const S* x, D* y, int n, int batch, const double* norm2, hipStream_t);
void launch_fill_int(int* p, int v, int count, hipStream_t);
template <typename S>
void launch_overlap(const S* x, const double2* ref, int n, int batch, double* part, int* counter,
                    double2* out, int out_stride, const int* skip, hipStream_t, const int* list = nullptr,
                    int rows = 0);
template <typename S> void launch_basis0(S* psi, int n, int batch, const int* skip, hipStream_t);
void launch_chain2_step(const Chain2Args&, int nblk, int batch, hipStream_t);
template <typename S> void launch_window_pass(const WinPassArgs<S>&, int nblk, int batch, hipStream_t);
template <typename S> void launch_fwd_pass(const FwdArgs<S>&, int nblk, int batch, hipStream_t);
void launch_window_finalize(const WinFinArgs&, int batch, hipStream_t);
// pre-sums groups of `group` window-pass partial blocks into the first slot of
// each group (large grids, before the one-block finalize)
void launch_window_reduce(double* part, int nblk, int group, int batch, const int* list, hipStream_t st);
void launch_ran_gates(const RanArgs&, int batch, hipStream_t);
void launch_amp0(const Amp0Args&, hipStream_t);

// chi-truncation chain (qmx_linalg.hip)
void launch_gram(const double2* y, size_t y_stride, const int* ncols_dev, int rows_log2,
                 int ncols_max, double2* part, int nchunks, int batch, hipStream_t);
void launch_gram_reduce(const double2* part, int nchunks, const int* ncols_dev, int ncols_max,
                        double2* a, size_t a_stride, int batch, hipStream_t);
// block Jacobi for Gram matrices above 512 x 512 (subproblems of 32 in LDS,
// block rotations applied as small GEMMs); returns -1 on a HIP failure;
// synchronises the stream once per sweep
size_t jacobi_block_bytes(int ncols_max, int batch);
int launch_jacobi_block(double2* a, double2* v, size_t stride, const int* ncols_dev, int ncols_max, int batch,
                        void* ws, int* conv_host_buf, int max_sweeps, int* error, hipStream_t st);
// LDS-resident Jacobi for matrices up to 128 x 128 (qmx_linalg.hip)
size_t jacobi_lds_log_doubles(int ncols_max);
bool jacobi_lds_fits(int ncols_max);
hipError_t jacobi_lds_setup(int ncols_max);
void launch_jacobi_lds(double2* a, double2* v, size_t stride, const int* ncols_dev, int ncols_max,
                       int batch, double* log, int* state, int max_sweeps, int* error, hipStream_t st);
void launch_jacobi(double2* a, double2* v, size_t stride, const int* ncols_dev, int ncols_max,
                   int batch, int max_sweeps, int* error, hipStream_t);
void launch_select(const double2* a, const double2* v, double2* w, size_t stride,
                   const int* ncols_dev, int ncols_max, int chi, double cutoff, int* k_out,
                   int kstride, double* spectra, size_t spec_stride, int batch, hipStream_t);
void launch_chain_dims(int* nn, int* ncur, const int* kk, int n, int step, int first_cols,
                       int batch, hipStream_t);
void launch_goal_check(const double2* dot, int dot_stride, const double* goal, int* skip,
                       int* active, int* sweeps_done, int batch, hipStream_t);
// list[0 .. live) = the states b < batch with skip[b] == 0, in increasing order
void launch_live_list(const int* skip, int batch, int* list, hipStream_t);
// C[b] (m x ncols) = A[b] (m x kdim) * op(B[b]); conj_b: op(B)[j][c] = conj(B[c][j])
// with leading dimension ldb; dims read from device arrays (per state) when given.
struct GemmDims {
  int rows_log2;          // m = 2^rows_log2
  const int* kdim;        // inner dimension per state (kdim[b * dim_stride]) or nullptr
  const int* ncols;       // output columns per state or nullptr
  int dim_stride;
  int kmax, nmax;         // host upper bounds (used when the pointers are null)
  int ldb;                // leading dimension of B
};
void launch_gemm(const double2* a, size_t a_stride, const double2* bm, size_t b_stride,
                 double2* c, size_t c_stride, const GemmDims& d, int conj_b, int batch,
                 hipStream_t);

// input preprocessing (qmx_prep.hip): re-index (snake) and zero the `cutoff`
// smallest-magnitude amplitudes of each state
size_t prep_input_bytes(int n, int batch);
template <typename S>
void launch_prep_input(const S* in, S* out, int n, int batch, int snake, size_t cutoff, void* ws, hipStream_t);

}  // namespace qmx
