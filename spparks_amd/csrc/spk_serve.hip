// spk_serve.hip -- libspk_hip.so, serving: continuous batching of a stream of
// solves over the batched-replica API (see spk_hip_serve in spk_hip.h).
// Host code only: the device work is the batch's step kernels.
#include <algorithm>
#include <chrono>
#include <deque>
#include <thread>
#include <vector>

#include "spk_host.h"

using namespace spk;

namespace {

double now_s(std::chrono::steady_clock::time_point t0) {
  return std::chrono::duration<double>(std::chrono::steady_clock::now() - t0).count();
}

struct BatchGuard {
  spk_hip_batch *b = nullptr;
  ~BatchGuard() { spk_hip_batch_destroy(b); }
};

}  // namespace

extern "C" int spk_hip_serve(const spk_hip_request *req, int n, int64_t pool_bytes, int max_live,
                             int block_sweeps, int block, int grid, spk_hip_result *out,
                             spk_hip_serve_stats *stats) {
  if (!req || !out || n < 0 || block_sweeps <= 0) return set_error(SPK_HIP_EINVAL, "bad serve arguments");
  for (int i = 1; i < n; i++)
    if (req[i].arrive < req[i - 1].arrive) return set_error(SPK_HIP_EINVAL, "requests not sorted by arrival");
  spk_hip_serve_stats st{};
  BatchGuard g;
  int rc = spk_hip_batch_create(pool_bytes, max_live, &g.b);
  if (rc) return rc;
  if ((rc = spk_hip_batch_set_launch(g.b, block, grid))) return rc;

  std::vector<int> owner(max_live, -1);   // replica id -> request index
  std::deque<int> waiting;                // arrived, not yet admitted (FIFO)
  int next = 0, completed = 0, live = 0;
  const auto t0 = std::chrono::steady_clock::now();
  while (completed < n) {
    double t = now_s(t0);
    while (next < n && req[next].arrive <= t) waiting.push_back(next++);
    // admit in arrival order while the pool and the slot table have room
    while (!waiting.empty()) {
      const int r = waiting.front();
      int id = -1;
      rc = spk_hip_batch_add(g.b, &req[r].replica, &id);
      if (rc == SPK_HIP_ENOMEM) {
        if (live == 0) {   // larger than the empty pool: can never run
          out[r] = spk_hip_result{t, t, 0, 0.0, 0, SPK_HIP_ENOMEM};
          completed++;
          waiting.pop_front();
          continue;
        }
        break;             // wait for retirements
      }
      if (rc == SPK_HIP_EINVAL || rc == SPK_HIP_EVALUE) {   // not a valid solve: refused alone
        out[r] = spk_hip_result{t, t, 0, 0.0, 0, rc};
        completed++;
        waiting.pop_front();
        continue;
      }
      if (rc) return rc;
      owner[id] = r;
      out[r] = spk_hip_result{now_s(t0), 0.0, 0, 0.0, 0, 0};
      live++;
      waiting.pop_front();
    }
    st.max_live = std::max(st.max_live, live);
    int64_t used = 0;
    spk_hip_batch_pool(g.b, &used, nullptr, nullptr);
    st.max_pool_used = std::max(st.max_pool_used, used);
    if (live == 0) {
      if (next >= n) break;
      const double wait = req[next].arrive - now_s(t0);
      if (wait > 0) std::this_thread::sleep_for(std::chrono::duration<double>(wait));
      continue;
    }
    const double tb = now_s(t0);
    if ((rc = spk_hip_batch_run(g.b, block_sweeps, 1))) return rc;
    st.busy += now_s(t0) - tb;
    st.blocks++;
    // retire finished requests at once
    for (int id = 0; id < max_live; id++) {
      const int r = owner[id];
      if (r < 0) continue;
      int done = 0, left = 0;
      int64_t nacc = 0;
      double e = 0.0;
      if ((rc = spk_hip_batch_status(g.b, id, &done, &left, &nacc, &e))) return rc;
      if (left > 0) continue;
      uint64_t sum = 0;
      if ((rc = spk_hip_batch_checksum(g.b, id, &sum))) return rc;
      out[r].finish = now_s(t0);
      out[r].naccept = nacc;
      out[r].energy = e;
      out[r].checksum = sum;
      const spk_hip_desc &d = req[r].replica.lattice;
      st.site_updates += (int64_t)d.nx * d.ny * d.nz * req[r].replica.nsweeps;
      if ((rc = spk_hip_batch_remove(g.b, id))) return rc;
      owner[id] = -1;
      live--;
      completed++;
    }
  }
  st.wall = now_s(t0);
  if (stats) *stats = st;
  return SPK_HIP_OK;
}
