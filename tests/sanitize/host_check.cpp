// Host-only entry points of libqmx under AddressSanitizer and
// UndefinedBehaviorSanitizer (built by tests/test_host_sanitizers.py): the
// workspace queries, the forward-pass schedule, ragged planning, the MLP
// sizes and the argument checks that return before any HIP call.  Needs no
// GPU; qmx_create is expected to fail cleanly where none is present.
#include <cstdio>
#include <cstdlib>
#include <vector>

#include "qmx.h"

static int failures = 0;
#define EXPECT(cond)                                                        \
  do {                                                                      \
    if (!(cond)) {                                                          \
      std::fprintf(stderr, "%s:%d: expected %s\n", __FILE__, __LINE__, #cond); \
      ++failures;                                                           \
    }                                                                       \
  } while (0)

int main() {
  EXPECT(qmx_version() >= 1);

  // workspace queries over every size the library takes, and invalid ones
  for (int n = -1; n <= 34; ++n) {
    for (int b : {0, 1, 3, 64}) {
      for (int chi : {0, 1, 2, 64, 512, 1024}) (void)qmx_chi_truncate_workspace(n, b, chi);
      (void)qmx_layer_workspace(n, b);
      (void)qmx_extract_layers_workspace(n, b);
      (void)qmx_prepare_input_workspace(n, b);
      for (int dt : {-1, QMX_C64, QMX_C128, 7}) (void)qmx_sweep_workspace(n, b, dt);
    }
  }
  EXPECT(qmx_chi_truncate_workspace(1, 1, 4) == 0);
  EXPECT(qmx_chi_truncate_workspace(10, 0, 4) == 0);
  EXPECT(qmx_sweep_workspace(12, 2, 5) == 0);
  EXPECT(qmx_sweep_workspace(12, 2, QMX_C64) > 0);

  // MPS site-tensor chains: valid, ragged, and malformed bond lists
  {
    std::vector<int> ok = {1, 2, 4, 3, 1};
    EXPECT(qmx_mps_to_state_workspace(4, ok.data()) > 0);
    std::vector<int> bad_end = {2, 2, 1};
    EXPECT(qmx_mps_to_state_workspace(2, bad_end.data()) == 0);
    std::vector<int> zero = {1, 0, 1};
    EXPECT(qmx_mps_to_state_workspace(2, zero.data()) == 0);
    EXPECT(qmx_mps_to_state_workspace(0, ok.data()) == 0);
    EXPECT(qmx_mps_to_state_workspace(3, nullptr) == 0);
  }

  // forward-pass schedule: sizes first, then capacities too small and exact
  for (int n = 3; n <= 30; ++n) {
    for (int L : {0, 1, 2, 10, 20}) {
      int nr = -1, ns = -1;
      const int rc = qmx_forward_plan(n, L, nullptr, 0, nullptr, 0, &nr, &ns);
      if (n < 5 || L < 1) {
        EXPECT(rc == QMX_ERR_INVALID);
        continue;
      }
      EXPECT(rc == QMX_ERR_WORKSPACE || (rc == QMX_OK && nr == 0));
      EXPECT(nr == L * n && ns > 0);
      std::vector<int> rec(3 * (size_t)nr);
      std::vector<unsigned long long> masks(ns);
      EXPECT(qmx_forward_plan(n, L, rec.data(), rec.size() - 1, masks.data(), masks.size(), &nr, &ns) ==
             QMX_ERR_WORKSPACE);
      EXPECT(qmx_forward_plan(n, L, rec.data(), rec.size(), masks.data(), masks.size(), &nr, &ns) == QMX_OK);
      for (int i = 0; i < nr; ++i) EXPECT(rec[3 * i + 2] >= 0 && rec[3 * i + 2] < L * n);
    }
  }
  EXPECT(qmx_forward_plan(8, 2, nullptr, 0, nullptr, 0, nullptr, nullptr) == QMX_ERR_INVALID);

  // whole-encode and grouped workspaces
  qmx_encode_config cfg = {12, 4, 16, 3, 2, QMX_C64, QMX_C64, 0.9, 1e-12};
  EXPECT(qmx_encode_workspace(&cfg) > 0);
  EXPECT(qmx_encode_workspace(nullptr) == 0);
  {
    const int first[2] = {0, 1}, count[2] = {1, 3}, chi[2] = {8, 16};
    EXPECT(qmx_encode_grouped_workspace(&cfg, first, count, chi, 2) > 0);
    const int bad_count[2] = {1, 2};
    EXPECT(qmx_encode_grouped_workspace(&cfg, first, bad_count, chi, 2) == 0);
    EXPECT(qmx_encode_grouped_workspace(&cfg, first, count, chi, 0) == 0);
  }

  // ragged planning: buckets of equal N ordered by bond dimension
  {
    std::vector<qmx_ragged_state> st;
    const int ns[] = {12, 20, 12, 15, 20, 9, 12};
    const int chis[] = {16, 64, 32, 16, 32, 8, 16};
    for (int i = 0; i < 7; ++i) st.push_back({ns[i], chis[i], nullptr, nullptr, nullptr, nullptr, nullptr});
    EXPECT(qmx_encode_ragged_workspace((int)st.size(), st.data(), 8, 30, QMX_C64, QMX_C64) > 0);
    EXPECT(qmx_encode_ragged_workspace(0, st.data(), 8, 30, QMX_C64, QMX_C64) == 0);
    st[3].num_qubits = 1;
    EXPECT(qmx_encode_ragged_workspace((int)st.size(), st.data(), 8, 30, QMX_C64, QMX_C64) == 0);
  }

  // MLP sizes
  EXPECT(qmx_mlp_param_count(64, 512, 2) == (size_t)64 * 512 + 512 + 512 * 512 + 512 + 512 * 2 + 2);
  (void)qmx_mlp_workspace(1000, 64, 512, 2);
  (void)qmx_mlp_workspace(0, 64, 512, 2);
  (void)qmx_mse_workspace();

  // argument checks that return before any HIP call
  EXPECT(qmx_last_error(nullptr) != nullptr);
  EXPECT(qmx_destroy(nullptr) == QMX_ERR_INVALID || qmx_destroy(nullptr) == QMX_OK);
  EXPECT(qmx_set_option(nullptr, 0, 0) == QMX_ERR_INVALID);
  EXPECT(qmx_create(0, nullptr) == QMX_ERR_INVALID);
  {
    qmx_handle* h = nullptr;
    const int rc = qmx_create(0, &h);   // QMX_ERR_HIP without a GPU
    EXPECT(rc == QMX_OK || (rc == QMX_ERR_HIP && h == nullptr));
    EXPECT(qmx_create(-1, &h) == QMX_ERR_HIP && h == nullptr);
    if (rc == QMX_OK) qmx_destroy(h);
  }
  EXPECT(qmx_sweep(nullptr, 12, 1, 2, nullptr, 1, nullptr, QMX_C64, nullptr, nullptr, nullptr, nullptr, nullptr,
                   nullptr, nullptr, nullptr, nullptr, 0, nullptr) == QMX_ERR_INVALID);
  EXPECT(qmx_encode(nullptr, &cfg, nullptr, nullptr, nullptr, nullptr, nullptr, nullptr, nullptr, nullptr, 0,
                    nullptr) == QMX_ERR_INVALID);
  EXPECT(qmx_encode_ragged(nullptr, 0, nullptr, 1, 1, QMX_C64, QMX_C64, 0.9, 0.0, 1, nullptr, 0, nullptr) ==
         QMX_ERR_INVALID);

  if (failures) {
    std::fprintf(stderr, "%d host checks failed\n", failures);
    return 1;
  }
  std::printf("host checks passed\n");
  return 0;
}
