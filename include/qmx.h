/*
 * qmx -- C-ABI of the MI355X Sequential MPS encoder (libqmx.so).
 *
 * The reference (Qualition/qmprs) is pure Python over quimb/numpy and has no
 * FFI; each entry point below replaces the reference routine named in its
 * comment, and is bound from Python with ctypes (qmprs_amd/_lib.py).
 *
 * Conventions
 *   - Every pointer marked "device" is device memory owned by the caller
 *     (torch tensors); the library never allocates device memory.
 *   - Scratch memory comes from a caller-provided workspace whose size is
 *     returned by the matching *_workspace() query (256-byte aligned inside).
 *   - Complex numbers are interleaved (re, im): complex128 = 2 doubles,
 *     complex64 = 2 floats.  Dense states of N qubits hold 2^N amplitudes,
 *     site 0 is the most significant bit of the amplitude index, and a batch
 *     of B states is stored state after state.
 *   - Gates are 4x4 complex128 matrices (out x in, row-major; a one-qubit gate
 *     occupies the top-left 2x2) stored at gates[((b*L + layer)*N + site)*16],
 *     with widths[(b*L + layer)*N + site] in {1, 2}.  Layer index = extraction
 *     order (the circuit applies layer L_b-1 first, as sequential.py:396).
 *   - Calls are asynchronous on `stream` (a hipStream_t, NULL = default)
 *     unless stated; they return a QMX_* status.
 *   - One handle per host thread.
 */
#ifndef QMX_H_
#define QMX_H_

#include <stddef.h>

#ifdef __cplusplus
extern "C" {
#endif

#define QMX_OK 0
#define QMX_ERR_INVALID 1        /* bad argument          -> ValueError        */
#define QMX_ERR_NOT_UNITARY 2    /* mps.py:837-839        -> ValueError        */
#define QMX_ERR_NO_CONVERGENCE 3 /* eigensolver           -> RuntimeError      */
#define QMX_ERR_HIP 4            /* HIP runtime failure   -> RuntimeError      */
#define QMX_ERR_WORKSPACE 5      /* workspace too small   -> ValueError        */
#define QMX_ERR_BASIS 6          /* degenerate projector  -> RuntimeError      */

#define QMX_C64 0
#define QMX_C128 1

typedef struct qmx_handle qmx_handle;

int qmx_version(void);
int qmx_create(int device, qmx_handle** out);
int qmx_destroy(qmx_handle* h);
const char* qmx_last_error(const qmx_handle* h);

/* Tuning options (no reference counterpart).
 * QMX_OPT_SWEEP_MODE: QMX_SWEEP_AUTO (windowed passes for N >= 10, else one
 * pass per gate), QMX_SWEEP_STEP (one pass per gate) or QMX_SWEEP_WINDOW
 * (windowed passes; below N = 10, where a tile holds too few rest indices for
 * the pass kernel's four waves, it also runs one pass per gate). */
#define QMX_OPT_SWEEP_MODE 1
#define QMX_SWEEP_AUTO 0
#define QMX_SWEEP_STEP 1
#define QMX_SWEEP_WINDOW 2
/* QMX_OPT_SPLIT_SWEEPS: 1 (default) runs a windowed sweep of a batch of two
 * or more states as two half-batches on two internal streams whose window
 * passes alternate, so that each half's window finalize overlaps the other
 * half's pass; 0 runs the whole batch on the caller's stream.  Results agree
 * to rounding (states are independent).
 * QMX_OPT_WINDOW_BLOCKS_PER_CU: grid of the windowed pass, in blocks per
 * compute unit over the whole batch (1..16, default 16). */
#define QMX_OPT_SPLIT_SWEEPS 2
#define QMX_OPT_WINDOW_BLOCKS_PER_CU 3
/* QMX_OPT_FWD_BLOCKS_PER_CU: grid of the forward-circuit rebuild passes, in
 * blocks per compute unit over the whole batch (1..16, default 4). */
#define QMX_OPT_FWD_BLOCKS_PER_CU 4
/* QMX_OPT_FWD_MFMA: 1 (default) runs the forward-circuit rebuild on the
 * matrix cores, with 2^12-amplitude tiles: complex64 states with
 * 12 <= N <= 28 (FP32 MFMA products) and complex128 states with
 * 12 <= N <= 27 (FP64 MFMA products); each stage's gates are fused into one
 * 16 x 16 complex matrix.  0 applies the gates one by one on the vector
 * units.  Results agree to the rounding of the state's type. */
#define QMX_OPT_FWD_MFMA 9
/* QMX_OPT_STAGED_WINDOWS: windows whose register groups stride the state
 * (gate bits at the lowest state bits) can copy their tiles between HBM and
 * LDS in tile order.  1 (default) does so when the tile's contiguous low run
 * is at least 2^7 amplitudes, 2 from 2^3, 0 never.  Results are identical. */
#define QMX_OPT_STAGED_WINDOWS 5
/* QMX_OPT_SWEEP_RELAXATION: omega * 1000 for the sweeps' gate update,
 * 1000..1950.  1000 (default) is the reference's update G' = conj(polar(E))
 * (sequential.py:473-478); omega > 1 replaces it by the over-relaxed
 * polar(G' + (omega - 1)(G' - G)), G the gate before the update, which
 * reaches a given fidelity in fewer sweeps but no longer follows the
 * reference's gate trajectory. */
#define QMX_OPT_SWEEP_RELAXATION 6
/* Failure injection (host-forced error paths for tests; every kernel still
 * runs to its normal end):
 * QMX_OPT_JACOBI_SWEEP_CAP: at most this many sweeps in the chi-truncation's
 * Jacobi eigensolvers (0, the default: their own limits of 30 / 40 / 80); a
 * state that has not converged then fails with QMX_ERR_NO_CONVERGENCE.
 * QMX_OPT_FAIL_UNITARITY: 1 makes every generated gate fail the unitarity
 * check of mps.py:837-839 (QMX_ERR_NOT_UNITARY); 0 (default) restores it. */
#define QMX_OPT_JACOBI_SWEEP_CAP 7
#define QMX_OPT_FAIL_UNITARITY 8
int qmx_set_option(qmx_handle* h, int option, int value);

/* Phase timers (no reference counterpart: instrumentation for benchmarks).
 * When enabled, every later call records HIP event pairs on its stream around
 * each phase; qmx_profile_read synchronises on them and returns, per phase,
 * the summed milliseconds and the number of kernel launches inside, then
 * clears the records.  ms and launches hold QMX_PHASE_COUNT entries. */
#define QMX_PHASE_CHI 0      /* statevector -> chi-MPS chain            */
#define QMX_PHASE_LAYERS 1   /* layer extraction loop                    */
#define QMX_PHASE_REBUILD 2  /* circuit state C|0..0> before each sweep  */
#define QMX_PHASE_SWEEP 3    /* sweep gate steps (all sweep kernels)     */
#define QMX_PHASE_WINDOW_PASS 4  /* the windowed sweep's streaming pass
                                    kernel alone, timed launch by launch
                                    (also counted in QMX_PHASE_SWEEP)      */
#define QMX_PHASE_COUNT 5
int qmx_profile_enable(qmx_handle* h, int enable);
int qmx_profile_read(qmx_handle* h, double* ms, long long* launches);

/* (a)+(b) statevector -> chi-truncated MPS state.
 * Replaces MPS.from_statevector (qmprs/primitives/mps.py:218-249): quick Ket
 * normalisation (base.py:96-97), MatrixProductState.from_dense and
 * tensor_network_1d_compress(max_bond=chi), evaluated as the right-to-left
 * truncated projection chain with FP64 Jacobi eigensolves.
 *   statevectors  device, B x 2^N of sv_dtype (QMX_C64 / QMX_C128)
 *   psi_out       device, B x 2^N complex128: normalised input (may be NULL)
 *   target_out    device, B x 2^N complex128: state of the chi-MPS (its norm
 *                 is below 1 by the discarded weight, as quimb leaves it)
 *   spectra       device or NULL, B x (N+1) x (2*chi) doubles: eigenvalues
 *                 (squared singular values) of each truncated bond, decreasing */
size_t qmx_chi_truncate_workspace(int num_qubits, int batch, int bond_dimension);
int qmx_chi_truncate(qmx_handle* h, int num_qubits, int batch, int bond_dimension, double cutoff,
                     const void* statevectors, int sv_dtype, void* psi_out, void* target_out,
                     double* spectra, void* workspace, size_t workspace_bytes, void* stream);

/* Input preprocessing of prepare_state (base.py:99-102), the quick Ket
 * operations the reference applies before building the MPS (quick is not
 * available here; the rules below are the ones assumed, see the oracle's
 * change_indexing / compress_amplitudes):
 *   index_type QMX_INDEX_ROW keeps the order; QMX_INDEX_SNAKE (N >= 3)
 *   reverses the second half of the amplitudes (the amplitude vector viewed
 *   as two rows, the odd row reversed);
 *   compression_percentage > 0 then sets the int(pct / 100 * 2^N) amplitudes
 *   of smallest magnitude to zero (equal magnitudes in index order);
 *   percentages that would zero every amplitude are QMX_ERR_INVALID.
 *   statevectors, out  device, B x 2^N of sv_dtype (out may equal the input
 *                      only when index_type is QMX_INDEX_ROW)
 * No normalisation: qmx_chi_truncate normalises, as Ket does. */
#define QMX_INDEX_ROW 0
#define QMX_INDEX_SNAKE 1
size_t qmx_prepare_input_workspace(int num_qubits, int batch);
int qmx_prepare_input(qmx_handle* h, int num_qubits, int batch, const void* statevectors, int sv_dtype,
                      int index_type, double compression_percentage, void* out, void* workspace,
                      size_t workspace_bytes, void* stream);

/* Dense state of an MPS given by its site tensors: the MPS(mps=...) entry
 * (mps.py:183-197, with MPS.to_statevector mps.py:251-270).
 *   bond_dims  host, num_sites + 1 ints, bond_dims[0] = bond_dims[N] = 1
 *   tensors    device, complex128 'lpr' site tensors A_i (bond_dims[i] x 2 x
 *              bond_dims[i+1]) stored one after the other
 *   state_out  device, 2^N complex128 (site 0 = most significant bit) */
size_t qmx_mps_to_state_workspace(int num_sites, const int* bond_dims);
int qmx_mps_to_state(qmx_handle* h, int num_sites, const int* bond_dims, const void* tensors, void* state_out,
                     void* workspace, size_t workspace_bytes, void* stream);

/* (b)+(c) one Ran layer from a normalised dense state.
 * Replaces MPS.generate_bond_D_unitary_layer (mps.py:849-891) with
 * generate_unitary_layer (mps.py:746-847): bond-2 compression in right form,
 * block split (quick _get_submps_indices), first/two/last-site unitaries with
 * Householder completion (scipy.linalg.null_space, mps.py:597,657,730) and
 * the unitarity check.  Writes layer `layer` of gates/widths (L = num_layers).
 *   psi_work device, B x 2^N complex128 (unit norm)
 *   bonds    device or NULL, B x L x N ints: bond dims of the bond-2 MPS
 * Synchronises the stream; returns QMX_ERR_NOT_UNITARY like mps.py:837-839. */
size_t qmx_layer_workspace(int num_qubits, int batch);
int qmx_generate_layer(qmx_handle* h, int num_qubits, int batch, int num_layers, int layer,
                       double cutoff, const void* psi_work, void* gates, int* widths, int* bonds,
                       void* workspace, size_t workspace_bytes, void* stream);

/* (d) apply layer `layer` (inverse = 1: U^dagger, gates i = end..start as
 * mps.py:933-971; inverse = 0: U, as mps.py:893-931) to dense states in place.
 *   psi device, B x 2^N complex128 */
int qmx_apply_layer(qmx_handle* h, int num_qubits, int batch, int num_layers, int layer,
                    int inverse, void* psi, const void* gates, const int* widths, void* stream);

/* (b)-(d) the whole layer loop of Sequential._get_unitary_layers
 * (sequential.py:330-398): per layer, generate + apply the inverse layer and
 * stop a state once |conj(psi[0]) - 1| <= (1 - fidelity_threshold) + 1e-5
 * (sequential.py:390).  layers_used[b] = layers kept for state b.
 *   psi_work  device, B x 2^N complex128, in: normalised chi-state, out:
 *             disentangled state
 *   amp0      device or NULL, B x L complex128: conj(psi[0]) after each layer
 * Synchronises the stream. */
size_t qmx_extract_layers_workspace(int num_qubits, int batch);
int qmx_extract_layers(qmx_handle* h, int num_qubits, int batch, int num_layers,
                       double fidelity_threshold, double cutoff, void* psi_work, void* gates,
                       int* widths, int* bonds, int* layers_used, void* amp0, void* workspace,
                       size_t workspace_bytes, void* stream);

/* (e) environment sweeps (sequential.py:400-541, Sequential._optimize_unitary_layers):
 * num_sweeps passes over the gates in reverse circuit order; per gate
 * phi <- G^dagger phi, E = <T'|phi>_rest, G' = conj(polar(E)), T' <- G'^dagger T'.
 *   target      device, B x 2^N of storage_dtype: the chi-MPS state
 *   gates       device, in/out (updated in place)
 *   psi_ref     device or NULL, B x 2^N complex128: fidelity reference
 *   fidelity    device or NULL, B x (num_sweeps+1) complex128: <psi_ref|C|0>
 *               before sweep s (s < num_sweeps) and after the last sweep
 *   overlap     device or NULL, B x num_sweeps doubles: <T|C'|0> at the end of
 *               each sweep (sum of singular values of the last environment)
 *   goal        device or NULL, B doubles: a state stops sweeping once
 *               |<psi_ref|C|0>| >= goal[b] (requires psi_ref; synchronises
 *               once per sweep); sweeps_done (device or NULL, B ints) then
 *               receives the number of sweeps each state ran.
 *   circuit_out device or NULL, B x 2^N of storage_dtype: final C|0..0> */
size_t qmx_sweep_workspace(int num_qubits, int batch, int storage_dtype);
int qmx_sweep(qmx_handle* h, int num_qubits, int batch, int num_layers, const int* layers_used,
              int num_sweeps, const void* target, int storage_dtype, void* gates,
              const int* widths, const void* psi_ref, void* fidelity, double* overlap,
              const double* goal, int* sweeps_done, void* circuit_out, void* workspace,
              size_t workspace_bytes, void* stream);

/* qmx_sweep that also copies the gates after snap_first and snap_first + 1
 * completed sweeps of this call (0: the gates it starts from; 0 <= snap_first
 * < num_sweeps) into snapshots (device, 2 x B x L x N x 16 complex128: slot
 * 0, then slot 1).  No counterpart in the reference: the learned warm start
 * reads a gate's last two updates from them (one call per stretch of sweeps
 * instead of three).  A snapshot whose sweep is not reached (every state at
 * its goal earlier) is not written. */
int qmx_sweep_snapshots(qmx_handle* h, int num_qubits, int batch, int num_layers, const int* layers_used,
                        int num_sweeps, const void* target, int storage_dtype, void* gates,
                        const int* widths, const void* psi_ref, void* fidelity, double* overlap,
                        const double* goal, int* sweeps_done, void* circuit_out, int snap_first,
                        void* snapshots, void* workspace, size_t workspace_bytes, void* stream);

/* qmx_sweep without a workspace argument.  No counterpart in the reference:
 * for callers that cannot allocate device memory (code under HIP graph
 * capture, tight serving loops, harnesses that forbid hipMalloc).  The
 * scratch qmx_sweep takes from its workspace comes from a library-owned
 * array of QMX_NOALLOC_SCRATCH_BYTES (64 MiB) per device, held from library
 * load to process end; problems whose qmx_sweep_workspace() exceeds it fail
 * with QMX_ERR_WORKSPACE.  The sweep runs on `stream` alone (no internal
 * streams or events).  Nothing is allocated; every call on one GPU shares
 * that GPU's scratch, so calls that overlap on one GPU race.  Results equal
 * qmx_sweep's with QMX_OPT_SPLIT_SWEEPS = 0 bit for bit. */
#define QMX_NOALLOC_SCRATCH_BYTES (64u << 20)
int qmx_sweep_noalloc(qmx_handle* h, int num_qubits, int batch, int num_layers, const int* layers_used,
                      int num_sweeps, const void* target, int storage_dtype, void* gates,
                      const int* widths, const void* psi_ref, void* fidelity, double* overlap,
                      const double* goal, int* sweeps_done, void* circuit_out, void* stream);

/* C|0..0> of a batch of gate lists on the GPU: the state a circuit backend's
 * get_statevector() returns for the emitted circuit (used by the reference's
 * tests, test_sequential_encoding.py:64-67), and the forward layer
 * application of mps.py:893-931 from |0..0>.  Workspace: qmx_sweep_workspace.
 *   state_out  device, B x 2^N of storage_dtype
 *   psi_ref / fidelity  optional (both or neither): fidelity[b] = <psi_ref|C|0> */
int qmx_circuit_state(qmx_handle* h, int num_qubits, int batch, int num_layers, const int* layers_used,
                      const void* gates, const int* widths, int storage_dtype, const void* psi_ref,
                      void* fidelity, void* state_out, void* workspace, size_t workspace_bytes,
                      void* stream);

/* Schedule of the forward circuit passes that qmx_circuit_state and the
 * sweeps' circuit rebuilds launch (no reference counterpart: host-only
 * inspection of the schedule; no GPU needed).  Forward position f is gate
 * (layer L - 1 - f / N, site f % N).  records gets 3 ints per gate in the
 * order the gates are applied: pass index, stage index (over all passes),
 * f; stage_masks[stage] = the stage's 4 state bits.  num_records /
 * num_stages are always set; returns QMX_ERR_WORKSPACE when a capacity is
 * too small. */
int qmx_forward_plan(int num_qubits, int num_layers, int* records, size_t record_capacity,
                     unsigned long long* stage_masks, size_t mask_capacity, int* num_records,
                     int* num_stages);

/* Whole Sequential.prepare_state (base.py:58-106 -> sequential.py:543-600)
 * for a batch of same-size states: chi-truncation, layer extraction, sweeps
 * and the final fidelity |<psi|C|0..0>|.  Gate emission into a circuit
 * object stays on the host (sequential.py:155-213).
 *   fidelity   device, B x (num_sweeps+1) complex128 (see qmx_sweep); when
 *              num_sweeps == 0 a single entry per state
 *   goal       device or NULL (see qmx_sweep) */
typedef struct qmx_encode_config {
  int num_qubits;
  int batch;
  int bond_dimension;
  int num_layers;
  int num_sweeps;
  int sv_dtype;        /* QMX_C64 / QMX_C128 input statevectors */
  int sweep_dtype;     /* storage of the sweep's dense states */
  double fidelity_threshold;
  double cutoff;
} qmx_encode_config;

size_t qmx_encode_workspace(const qmx_encode_config* cfg);
int qmx_encode(qmx_handle* h, const qmx_encode_config* cfg, const void* statevectors,
               void* gates, int* widths, int* layers_used, void* fidelity, const double* goal,
               int* sweeps_done, void* workspace, size_t workspace_bytes, void* stream);

/* The part of qmx_encode before the sweeps (chi-truncation, normalisation,
 * layer extraction), for callers that drive the sweeps themselves with
 * qmx_sweep: qmx_encode equals qmx_encode_prepare followed by
 * qmx_sweep(target_out, psi_ref = psi_out).  Workspace: qmx_encode_workspace.
 *   psi_out     device, B x 2^N complex128: the normalised input
 *   target_out  device, B x 2^N of cfg->sweep_dtype: the chi-MPS state */
int qmx_encode_prepare(qmx_handle* h, const qmx_encode_config* cfg, const void* statevectors,
                       void* gates, int* widths, int* layers_used, void* psi_out, void* target_out,
                       void* workspace, size_t workspace_bytes, void* stream);

/* Grouped bond dimensions (no reference counterpart: each reference call
 * has one bond_dimension, base.py:58-65).  qmx_encode of a batch of equal-size
 * states whose bond dimensions differ: states [first[k], first[k] + count[k])
 * use bond dimension chi[k]; the groups tile the batch in order (first[0] = 0,
 * counts summing to cfg->batch).  Only the chi-truncation depends on the bond
 * dimension and runs per group; the layer extraction and the sweeps run once
 * over the whole batch.  cfg->bond_dimension is ignored; no fidelity goal.
 * With one group the results equal qmx_encode's bit for bit. */
size_t qmx_encode_grouped_workspace(const qmx_encode_config* cfg, const int* first, const int* count,
                                    const int* chi, int ngroups);
int qmx_encode_grouped(qmx_handle* h, const qmx_encode_config* cfg, const int* first, const int* count,
                       const int* chi, int ngroups, const void* statevectors, void* gates, int* widths,
                       int* layers_used, void* fidelity, void* workspace, size_t workspace_bytes,
                       void* stream);

/* Ragged batches (no reference counterpart: the reference encodes one state
 * per prepare_state call, base.py:58-106).  Each state brings its own qubit
 * count, bond dimension and device buffers; states with equal N are encoded
 * together as one batch (qmx_encode_grouped, ordered by bond dimension), and
 * the buckets, largest estimated cost first, run on up to max_concurrency
 * worker threads (6 when max_concurrency <= 0) with their own handles and
 * streams, so they overlap on the GPU.  All states share num_layers,
 * num_sweeps and the dtypes.  Synchronises (returns when every state is
 * done).
 *   gates        device, num_layers x N x 16 complex128 (qmx gate layout)
 *   widths       device, num_layers x N ints
 *   layers_used  device, 1 int
 *   fidelity     device, num_sweeps + 1 complex128 (see qmx_sweep) */
typedef struct qmx_ragged_state {
  int num_qubits;
  int bond_dimension;
  const void* statevector;
  void* gates;
  int* widths;
  int* layers_used;
  void* fidelity;
} qmx_ragged_state;

size_t qmx_encode_ragged_workspace(int count, const qmx_ragged_state* states, int num_layers,
                                   int num_sweeps, int sv_dtype, int sweep_dtype);
int qmx_encode_ragged(qmx_handle* h, int count, const qmx_ragged_state* states, int num_layers,
                      int num_sweeps, int sv_dtype, int sweep_dtype, double fidelity_threshold,
                      double cutoff, int max_concurrency, void* workspace, size_t workspace_bytes,
                      void* stream);

/* ---- Learned warm start (no reference counterpart) ------------------------
 * A 3-layer ReLU MLP (BF16 weights, FP32 accumulation on MFMA) maps a feature
 * row per two-qubit gate slot to 15 su(4) coefficients c; the proposal is
 * G exp(i sum_k c_k l_k) (Gell-Mann basis, tr(l_a l_b) = 2 delta_ab).  It
 * proposes starting gates for the sweeps of Sequential._optimize_unitary_layers
 * (sequential.py:509-541); the plain sweeps remain the fallback and decide the final fidelity.
 *
 * Parameters (FP32 master, BF16 shadow with the same offsets):
 *   W1[in][width], b1[width], W2[width][width], b2[width], W3[width][out], b3[out]
 * Matrices are row-major; x is BF16 [rows][in], y FP32 [rows][out].
 * qmx_mlp_backward must follow qmx_mlp_forward on the same x, parameters and
 * workspace (it reads the activations the forward pass stored there) and
 * writes the gradient of sum(dy * y) for every parameter into grads. */
size_t qmx_mlp_param_count(int in_dim, int width, int out_dim);
size_t qmx_mlp_workspace(int rows, int in_dim, int width, int out_dim);
int qmx_mlp_forward(qmx_handle* h, int rows, int in_dim, int width, int out_dim, const void* x_bf16,
                    const float* params, const void* params_bf16, float* y, void* workspace,
                    size_t workspace_bytes, void* stream);
int qmx_mlp_backward(qmx_handle* h, int rows, int in_dim, int width, int out_dim, const void* x_bf16,
                     const float* params, const void* params_bf16, const float* dy, float* grads,
                     void* workspace, size_t workspace_bytes, void* stream);
/* Adam with bias correction (step >= 1); refreshes the BF16 shadow if given. */
int qmx_mlp_adam(qmx_handle* h, size_t count, float* params, void* params_bf16, const float* grads,
                 float* m, float* v, float lr, float beta1, float beta2, float eps, int step,
                 void* stream);
int qmx_mlp_to_bf16(qmx_handle* h, size_t count, const float* src, void* dst, void* stream);
/* Weighted MSE over the first `valid` of `ld` columns: loss (device float) =
 * inv_norm * sum_r w_r sum_c (y - t)^2, dy = d loss / d y (w may be NULL). */
size_t qmx_mse_workspace(void);
int qmx_mse_loss(qmx_handle* h, int rows, int ld, int valid, const float* y, const float* t,
                 const float* w, float inv_norm, float* dy, float* loss, void* workspace,
                 size_t workspace_bytes, void* stream);

/* Gate slot rows: row = (b * num_layers + layer) * (N - 1) + site for every
 * site < N - 1; a row is live when layer < layers_used[b] and the slot holds a
 * two-qubit gate (widths == 2).  Gates use the qmx gate layout above.
 *   qmx_gate_features: BF16 rows of width ld >= 50:
 *       [c(Ga^H Gb) (15), Re/Im of Gb row-major (32), layer / L, site / N, 1, 0 ...]
 *       row_weight (device or NULL) = 1 for live rows, 0 otherwise
 *   qmx_gate_log:      FP32 rows of width ld >= 15: c(Ga^H Gb), the traceless
 *                      principal logarithm (eigenphases in (-pi, pi])
 *   qmx_gate_expmap:   G <- G exp(i sum_k c_k l_k) on live rows of states with
 *                      skip[b] == 0 (skip may be NULL) */
int qmx_gate_features(qmx_handle* h, int num_qubits, int batch, int num_layers, const int* layers_used,
                      const void* gates_a, const void* gates_b, const int* widths, void* features,
                      int ld, float* row_weight, void* stream);
int qmx_gate_log(qmx_handle* h, int num_qubits, int batch, int num_layers, const int* layers_used,
                 const void* gates_a, const void* gates_b, const int* widths, float* coeffs, int ld,
                 void* stream);
int qmx_gate_expmap(qmx_handle* h, int num_qubits, int batch, int num_layers, const int* layers_used,
                    const float* coeffs, int ld, void* gates, const int* widths, const int* skip,
                    void* stream);

/* Trajectory extrapolation (the learned warm start's jumps; no reference
 * counterpart).  Rows as above.  For the gates after sweeps s-2 (a), s-1 (b)
 * and s (c): d1 = c(Ga^H Gb), d2 = c(Gb^H Gc), row_scale = |d2| (floor
 * 1e-12); e2 = d2 / |d2| and e1 = the part of d1 orthogonal to e2,
 * normalised (zero when d1 is parallel to d2 to 1e-6).
 *   qmx_traj_features: BF16 rows (ld >= 35) = [d1 / scale (15), d2 / scale
 *                      (15), log10(scale) / 8, sweep_frac, layer / L, site / N,
 *                      1, 0 ...]; row_weight as above; row_scale
 *   qmx_traj_targets:  FP32 rows (ld >= 2) = [t.e2, t.e1, 0 ...] with t =
 *                      c(Gc^H Gd) / scale, Gd the gates K sweeps after s
 *   qmx_traj_jump:     G <- G exp(i scale sum_k (c_0 e2_k + c_1 e1_k) l_k),
 *                      the basis from (Ga, Gb, G), on live rows of states with
 *                      skip[b] == 0 (skip may be NULL) */
int qmx_traj_features(qmx_handle* h, int num_qubits, int batch, int num_layers, const int* layers_used,
                      const void* gates_a, const void* gates_b, const void* gates_c, const int* widths,
                      float sweep_frac, void* features, int ld, float* row_weight, float* row_scale,
                      void* stream);
int qmx_traj_targets(qmx_handle* h, int num_qubits, int batch, int num_layers, const int* layers_used,
                     const void* gates_a, const void* gates_b, const void* gates_c, const void* gates_d,
                     const int* widths, float* targets, int ld, void* stream);
int qmx_traj_jump(qmx_handle* h, int num_qubits, int batch, int num_layers, const int* layers_used,
                  const void* gates_a, const void* gates_b, const float* coeffs, int ld, void* gates,
                  const int* widths, const int* skip, void* stream);

#ifdef __cplusplus
}
#endif

#endif /* QMX_H_ */
