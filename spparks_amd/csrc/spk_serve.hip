// spk_serve.hip -- libspk_hip.so, serving: continuous batching of a stream of
// solves over the batched-replica path (see spk_hip_serve in spk_hip.h).
// Host code only: the device work is the batch's step kernels.
//
// One block is in flight at a time.  While it runs, the host admits the
// requests that have arrived (their set-up kernels queue behind the block)
// and queues the checksums of the requests the block finishes; it then waits
// once per block, reports and retires those requests, and launches the next
// block.  The host waits for nothing else.
#include <algorithm>
#include <chrono>
#include <deque>
#include <thread>
#include <vector>

#include "spk_batch_async.h"
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
  bool tuned = block != 0;   // a zero shape is tuned on the first block, which waits

  std::vector<int> owner(max_live, -1);   // replica id -> request index
  std::deque<int> waiting;                // arrived, not yet admitted (FIFO)
  int next = 0, completed = 0, live = 0;
  const auto t0 = std::chrono::steady_clock::now();

  // admits, in arrival order, every arrived request while the pool and the
  // slot table have room; requests that can never run are refused alone
  auto admit = [&]() -> int {
    const double t = now_s(t0);
    while (next < n && req[next].arrive <= t) waiting.push_back(next++);
    while (!waiting.empty()) {
      const int r = waiting.front();
      int id = -1;
      const int rc = batch_add_async(g.b, &req[r].replica, &id);
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
    return SPK_HIP_OK;
  };

  std::vector<int> finishing;
  while (completed < n) {
    if ((rc = admit())) return rc;
    if (live == 0) {
      if (next >= n) break;
      const double wait = req[next].arrive - now_s(t0);
      if (wait > 0) std::this_thread::sleep_for(std::chrono::duration<double>(wait));
      continue;
    }
    // the requests this block finishes, known before it runs
    finishing.clear();
    for (int id = 0; id < max_live; id++)
      if (owner[id] >= 0 && batch_left(g.b, id) <= block_sweeps) finishing.push_back(id);
    if (!tuned) {
      // the first block tunes the launch shape on its own steps and waits
      if ((rc = spk_hip_batch_run(g.b, block_sweeps, 1))) return rc;
      tuned = true;
    } else if ((rc = batch_launch(g.b, block_sweeps, true))) {
      return rc;
    }
    for (int id : finishing)
      if ((rc = batch_queue_checksum(g.b, id))) return rc;
    // while the block runs: admit what has arrived (queued behind it)
    if ((rc = admit())) return rc;
    if ((rc = batch_complete(g.b))) return rc;
    double ms = 0.0;
    spk_hip_batch_last_timing(g.b, &ms, nullptr, nullptr, nullptr, nullptr);
    st.busy += ms / 1e3;
    st.blocks++;
    // report and retire the finished requests
    const double tf = now_s(t0);
    for (int id : finishing) {
      const int r = owner[id];
      int64_t nacc = 0;
      double e = 0.0;
      if ((rc = spk_hip_batch_status(g.b, id, nullptr, nullptr, &nacc, &e))) return rc;
      out[r].finish = tf;
      out[r].naccept = nacc;
      out[r].energy = e;
      out[r].checksum = batch_checksum_value(g.b, id);
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
