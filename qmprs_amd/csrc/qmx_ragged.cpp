// Ragged batches: many states of different qubit counts and bond dimensions
// in one call (no reference counterpart: the reference encodes one state per
// Sequential.prepare_state call, base.py:58-106).
//
// States are grouped into buckets of equal num_qubits, ordered inside a
// bucket by bond dimension; buckets are ordered by estimated cost (2^N * N *
// layers * sweeps, largest first) and handed to a pool of worker threads.
// Each worker owns a qmx handle and a non-blocking stream; per bucket it
// gathers the inputs into a contiguous batch (device-to-device copies), runs
// one encode in which only the chi-truncation is split by bond dimension
// (qmx_encode_grouped), and scatters gates, widths, layer counts and
// fidelities back to each state's buffers.
// The host synchronisations inside an encode (chi-chain error check, layer
// extraction) then block only their own worker, so buckets overlap on the
// GPU.  The caller's stream is joined at both ends.
#include <hip/hip_runtime.h>

#include <algorithm>
#include <atomic>
#include <map>
#include <mutex>
#include <string>
#include <thread>
#include <utility>
#include <vector>

#include "qmx.h"

// handle internals shared with qmx_api.cpp
void qmx_set_error_text(qmx_handle* h, const char* msg);
int qmx_handle_device(const qmx_handle* h);
void qmx_copy_options(const qmx_handle* from, qmx_handle* to);

namespace {

struct Bucket {
  int n;
  std::vector<int> idx;                           // states, by (bond dimension, index)
  std::vector<int> gfirst, gcount, gchi;          // bond-dimension groups of idx
  double cost;
  size_t off_in, off_gates, off_widths, off_used, off_fid, off_ws, ws_bytes;
};

size_t align256(size_t x) { return (x + 255) & ~static_cast<size_t>(255); }

size_t sv_bytes(int dtype) { return dtype == QMX_C64 ? 8 : 16; }

// bucket plan and workspace offsets; returns total bytes (0 on bad input)
size_t plan(int count, const qmx_ragged_state* st, int L, int S, int sv_dtype, int sweep_dtype,
            std::vector<Bucket>& out) {
  if (count < 1 || !st || L < 1 || S < 0) return 0;
  std::map<int, std::vector<int>> groups;
  for (int i = 0; i < count; ++i) {
    if (st[i].num_qubits < 2 || st[i].num_qubits > 30 || st[i].bond_dimension < 1) return 0;
    groups[st[i].num_qubits].push_back(i);
  }
  out.clear();
  for (auto& kv : groups) {
    Bucket b{};
    b.n = kv.first;
    b.idx = kv.second;
    std::stable_sort(b.idx.begin(), b.idx.end(),
                     [&](int x, int y) { return st[x].bond_dimension < st[y].bond_dimension; });
    for (int k = 0; k < static_cast<int>(b.idx.size()); ++k) {
      const int chi = st[b.idx[k]].bond_dimension;
      if (b.gchi.empty() || b.gchi.back() != chi) {
        b.gfirst.push_back(k);
        b.gcount.push_back(0);
        b.gchi.push_back(chi);
      }
      ++b.gcount.back();
    }
    b.cost = static_cast<double>(b.idx.size()) * static_cast<double>(1ull << b.n) * b.n * L * (S + 1);
    out.push_back(b);
  }
  std::sort(out.begin(), out.end(), [](const Bucket& a, const Bucket& b) { return a.cost > b.cost; });
  size_t off = 0;
  for (auto& b : out) {
    const size_t B = b.idx.size(), len = static_cast<size_t>(1) << b.n;
    b.off_in = off = align256(off);
    off += B * len * sv_bytes(sv_dtype);
    b.off_gates = off = align256(off);
    off += B * L * b.n * 16 * 16;
    b.off_widths = off = align256(off);
    off += B * L * b.n * sizeof(int);
    b.off_used = off = align256(off);
    off += B * sizeof(int);
    b.off_fid = off = align256(off);
    off += B * (S + 1) * 16;
    qmx_encode_config cfg{b.n, static_cast<int>(B), b.gchi.back(), L, S, sv_dtype, sweep_dtype, 1 - 1e-6, 1e-10};
    b.ws_bytes = qmx_encode_grouped_workspace(&cfg, b.gfirst.data(), b.gcount.data(), b.gchi.data(),
                                              static_cast<int>(b.gchi.size()));
    if (b.ws_bytes == 0) return 0;
    b.off_ws = off = align256(off);
    off += b.ws_bytes;
  }
  return off + 256;
}

// One bucket on stream ws: gather the inputs into the bucket's contiguous
// batch (device-to-device copies), one grouped encode, scatter gates, widths,
// layer counts and fidelities back to each state's buffers (only the encode
// when scatter is false).  Returns a QMX status; msg names the failure.
int issue_bucket(qmx_handle* wh, const Bucket& b, const qmx_ragged_state* states, char* base, int L, int S,
                 int sv_dtype, int sweep_dtype, double thr, double cutoff, hipStream_t ws, bool scatter,
                 std::string& msg) {
  const int B = static_cast<int>(b.idx.size());
  const size_t len = static_cast<size_t>(1) << b.n;
  const size_t svb = len * sv_bytes(sv_dtype);
  const size_t gb = static_cast<size_t>(L) * b.n * 16 * 16, wb = static_cast<size_t>(L) * b.n * sizeof(int);
  const size_t fb = static_cast<size_t>(S + 1) * 16;
  for (int k = 0; k < B; ++k)
    if (hipMemcpyAsync(base + b.off_in + k * svb, states[b.idx[k]].statevector, svb, hipMemcpyDeviceToDevice,
                       ws) != hipSuccess) {
      msg = "gather";
      return QMX_ERR_HIP;
    }
  qmx_encode_config cfg{b.n, B, b.gchi.back(), L, S, sv_dtype, sweep_dtype, thr, cutoff};
  const int rc = qmx_encode_grouped(wh, &cfg, b.gfirst.data(), b.gcount.data(), b.gchi.data(),
                                    static_cast<int>(b.gchi.size()), base + b.off_in, base + b.off_gates,
                                    reinterpret_cast<int*>(base + b.off_widths),
                                    reinterpret_cast<int*>(base + b.off_used), base + b.off_fid, base + b.off_ws,
                                    b.ws_bytes, ws);
  if (rc != QMX_OK) {
    msg = std::string("bucket N=") + std::to_string(b.n) + ": " + qmx_last_error(wh);
    return rc;
  }
  if (!scatter) return QMX_OK;
  for (int k = 0; k < B; ++k) {
    const qmx_ragged_state& s = states[b.idx[k]];
    const bool ok =
        hipMemcpyAsync(s.gates, base + b.off_gates + k * gb, gb, hipMemcpyDeviceToDevice, ws) == hipSuccess &&
        hipMemcpyAsync(s.widths, base + b.off_widths + k * wb, wb, hipMemcpyDeviceToDevice, ws) == hipSuccess &&
        hipMemcpyAsync(s.layers_used, base + b.off_used + k * sizeof(int), sizeof(int), hipMemcpyDeviceToDevice,
                       ws) == hipSuccess &&
        hipMemcpyAsync(s.fidelity, base + b.off_fid + k * fb, fb, hipMemcpyDeviceToDevice, ws) == hipSuccess;
    if (!ok) {
      msg = "scatter";
      return QMX_ERR_HIP;
    }
  }
  return QMX_OK;
}

bool states_valid(int count, const qmx_ragged_state* states) {
  for (int i = 0; i < count; ++i)
    if (!states[i].statevector || !states[i].gates || !states[i].widths || !states[i].layers_used ||
        !states[i].fidelity)
      return false;
  return true;
}

}  // namespace


extern "C" {

size_t qmx_encode_ragged_workspace(int count, const qmx_ragged_state* states, int num_layers,
                                   int num_sweeps, int sv_dtype, int sweep_dtype) {
  std::vector<Bucket> b;
  return plan(count, states, num_layers, num_sweeps, sv_dtype, sweep_dtype, b);
}

int qmx_encode_ragged(qmx_handle* h, int count, const qmx_ragged_state* states, int num_layers,
                      int num_sweeps, int sv_dtype, int sweep_dtype, double fidelity_threshold,
                      double cutoff, int max_concurrency, void* workspace, size_t workspace_bytes,
                      void* stream) {
  if (!h) return QMX_ERR_INVALID;
  if ((sv_dtype != QMX_C64 && sv_dtype != QMX_C128) || (sweep_dtype != QMX_C64 && sweep_dtype != QMX_C128))
    return QMX_ERR_INVALID;
  std::vector<Bucket> buckets;
  const size_t need = plan(count, states, num_layers, num_sweeps, sv_dtype, sweep_dtype, buckets);
  if (need == 0) return QMX_ERR_INVALID;
  if (!states_valid(count, states)) return QMX_ERR_INVALID;
  if (!workspace || workspace_bytes < need) return QMX_ERR_WORKSPACE;
  const int device = qmx_handle_device(h);
  if (hipSetDevice(device) != hipSuccess) return QMX_ERR_HIP;
  hipStream_t caller = static_cast<hipStream_t>(stream);
  hipEvent_t start = nullptr;
  if (hipEventCreateWithFlags(&start, hipEventDisableTiming) != hipSuccess) return QMX_ERR_HIP;
  if (hipEventRecord(start, caller) != hipSuccess) {
    (void)hipEventDestroy(start);
    return QMX_ERR_HIP;
  }
  const int workers = std::max(1, std::min<int>(max_concurrency > 0 ? max_concurrency : 6,
                                               static_cast<int>(buckets.size())));
  std::atomic<size_t> next{0};
  std::atomic<int> status{QMX_OK};
  std::mutex err_mu;
  std::string err_msg;
  std::vector<hipEvent_t> done(workers, nullptr);
  std::vector<hipStream_t> streams(workers, nullptr);
  char* base = static_cast<char*>(workspace);
  const int L = num_layers, S = num_sweeps;
  auto work = [&](int w) {
    auto fail = [&](int code, const std::string& msg) {
      int expected = QMX_OK;
      if (status.compare_exchange_strong(expected, code)) {
        std::lock_guard<std::mutex> g(err_mu);
        err_msg = msg;
      }
    };
    if (hipSetDevice(device) != hipSuccess) return fail(QMX_ERR_HIP, "hipSetDevice");
    qmx_handle* wh = nullptr;
    if (qmx_create(device, &wh) != QMX_OK) return fail(QMX_ERR_HIP, "worker handle");
    qmx_copy_options(h, wh);   // the caller's sweep settings apply to every bucket
    hipStream_t ws = nullptr;
    if (hipStreamCreateWithFlags(&ws, hipStreamNonBlocking) != hipSuccess ||
        hipStreamWaitEvent(ws, start, 0) != hipSuccess) {
      qmx_destroy(wh);
      return fail(QMX_ERR_HIP, "worker stream");
    }
    streams[w] = ws;
    for (size_t bi = next.fetch_add(1); bi < buckets.size() && status.load() == QMX_OK;
         bi = next.fetch_add(1)) {
      std::string msg;
      const int rc = issue_bucket(wh, buckets[bi], states, base, L, S, sv_dtype, sweep_dtype, fidelity_threshold,
                                  cutoff, ws, true, msg);
      if (rc != QMX_OK) {
        fail(rc, msg);
        break;
      }
    }
    hipEvent_t ev = nullptr;
    if (hipEventCreateWithFlags(&ev, hipEventDisableTiming) != hipSuccess || hipEventRecord(ev, ws) != hipSuccess)
      fail(QMX_ERR_HIP, "worker event");
    done[w] = ev;
    // the handle's internal streams must drain before it is destroyed
    (void)hipStreamSynchronize(ws);
    qmx_destroy(wh);
  };
  std::vector<std::thread> pool;
  for (int w = 0; w < workers; ++w) pool.emplace_back(work, w);
  for (auto& t : pool) t.join();
  int rc = status.load();
  for (int w = 0; w < workers; ++w) {
    if (done[w] && hipStreamWaitEvent(caller, done[w], 0) != hipSuccess && rc == QMX_OK) rc = QMX_ERR_HIP;
  }
  if (rc != QMX_OK) (void)hipStreamSynchronize(caller);
  for (int w = 0; w < workers; ++w) {
    if (streams[w]) (void)hipStreamSynchronize(streams[w]);
    if (done[w]) (void)hipEventDestroy(done[w]);
    if (streams[w]) (void)hipStreamDestroy(streams[w]);
  }
  (void)hipEventDestroy(start);
  if (rc != QMX_OK) qmx_set_error_text(h, err_msg.c_str());
  return rc;
}


}  // extern "C"
