// spk_batch.hip -- libspk_hip.so, batched replicas: a device pool holding many
// independent lattices, one persistent kernel per colour step over a tile
// work-list spanning all replicas due a pass at that step, HIP-graph replay of
// blocks, and an autotuner for the launch shape.  ABI in include/spk_hip.h.
#include <hip/hip_runtime.h>

#include <algorithm>
#include <cstring>
#include <map>
#include <new>
#include <tuple>
#include <vector>

#include "spk_batch_async.h"
#include "spk_device.h"
#include "spk_host.h"
#include "spk_kernels.h"

using namespace spk;

namespace {

// Device-side description of one replica for the current block.
struct BatchRep {
  Geom g;
  Rule rule;
  int stype;       // SpinType
  int P;           // colour passes this block (K * sweeps)
  int acc_row;     // first accumulator row (one row per sweep of the block)
  int tiles;       // tiles per pass
  int64_t spin_off;  // byte offsets into the pool
  int64_t seed_off;
};

// Active replicas of one step: table indices act[first .. first+count) with
// tile prefix sums pre[first + count + first_pre ...] (see build_schedule).
struct StepDesc {
  int first;       // into act[]
  int count;       // active replicas
  int pre_first;   // into pre[]: count+1 prefix sums of tiles
  int total;       // tiles in the step
};

constexpr int kVec = 4;
constexpr int kBatchSlots = 8;                         // accumulator slots per replica sweep row
constexpr int kBatchRow = kBatchSlots * kSlotStride;   // uint64 words per row

constexpr int kTile = 64;   // threads per tile: one wavefront, 4 sites each

// One colour step of a block for the replicas of one (lattice, storage)
// variant.  A tile is one wavefront's work: 64 packs (4 consecutive sites of
// one sub-grid row each) of one replica's colour pass.  Every wavefront of
// the persistent grid takes a contiguous chunk of the step's tile list and
// walks it alone, replica by replica: a replica's table entry is read once
// (scalar loads) and its thresholds go to the wavefront's own LDS table;
// the counts of all its tiles in the chunk stay in registers and leave by
// one wavefront reduction and one atomic per counter.  No wavefront ever
// waits for another.  Consecutive wavefronts (XCD-remapped) take
// consecutive chunks, so neighbouring tiles share an XCD's L2.
template <int LAT, typename T>
__global__ __launch_bounds__(1024) void k_batch_step(char *__restrict__ pool, const BatchRep *__restrict__ reps,
                                                     const StepDesc *__restrict__ steps, const int *__restrict__ act,
                                                     const int *__restrict__ pre, int j, int L,
                                                     unsigned long long *__restrict__ acc) {
  __shared__ int s_thr[16][32];   // one threshold table per wavefront
  const StepDesc sd = steps[j];
  // the wavefront index is the same on all 64 lanes; saying so puts it, the
  // chunk, the replica and its geometry in scalar registers and their loads
  // on the scalar unit (the compiler treats threadIdx-derived values as
  // per-lane otherwise)
  const int wpb = blockDim.x >> 6, lane = threadIdx.x & 63;
  const int wid = __builtin_amdgcn_readfirstlane((int)(threadIdx.x >> 6));
  const int W = gridDim.x * wpb;
  const int gw = xcd_tile(blockIdx.x, gridDim.x) * wpb + wid;
  const int chunk = (sd.total + W - 1) / W;
  int t = gw * chunk;
  const int t1 = min(sd.total, t + chunk);
  if (t >= t1) return;   // nothing below waits for another wavefront
  // i, the replica and the tile are uniform across the wavefront: the
  // replica table and step lists are read with scalar loads
  const int *pp = pre + sd.pre_first;
  int lo = 0, hi = sd.count - 1;   // the replica holding tile t: last i with pp[i] <= t
  while (lo < hi) {
    const int mid = (lo + hi + 1) >> 1;
    if (pp[mid] <= t) lo = mid;
    else hi = mid - 1;
  }
  int *thr = s_thr[wid];
  for (int i = lo; t < t1; i++) {
    const int first = pp[i], tend = min(t1, pp[i + 1]);
    const BatchRep &rep = reps[act[sd.first + i]];
    const Geom g = rep.g;
    const Rule &rule = rep.rule;
    if (lane < 28) thr[lane] = rule.thr[lane];
    wave_lds_sync();
    const int p = (int)(((int64_t)j * rep.P) / L);
    const int c = p % g.K;
    T *sp = reinterpret_cast<T *>(pool + rep.spin_off);
    int32_t *sdp = reinterpret_cast<int32_t *>(pool + rep.seed_off);
    const uint32_t npack = (uint32_t)(g.Mp / kVec);
    bool two = false;
    if constexpr (sizeof(T) == 1) two = rule.nspins == 2;   // two states are stored as u8
    int nacc = 0, de = 0;
    // the replica's tiles of this chunk; XD is the x-shift side when the
    // colour fixes it (sq/8n, sc/26n: the colour's x parity), a template
    // argument so that the shifts are not selected per lane
    auto tiles = [&]<bool TWO, int XD>() {
      for (; t < tend; t++) {
        const uint32_t ti = (uint32_t)(t - first) * kTile + (uint32_t)lane;
        if (ti < npack) {
          int a0, b, e;
          pack_site(g, ti, a0, b, e);
          PackWork<T, TWO, uint32_t> w = pack_count<LAT, T, TWO, uint32_t, XD>(sp, sdp, g, rule, c, a0, b, e);
          const Counts cnt = pack_decide<LAT, T, TWO, uint32_t>(w, sp, sdp, g, rule, thr, c);
          nacc += cnt.nacc;
          de += cnt.de;
        }
      }
    };
    auto with_two = [&]<int XD>() {
      if constexpr (sizeof(T) == 1) {
        if (two) {
          tiles.template operator()<true, XD>();
          return;
        }
      }
      tiles.template operator()<false, XD>();
    };
    if constexpr (LAT == SPK_HIP_SQ_8N || LAT == SPK_HIP_SC_26N) {
      if (c & 1) with_two.template operator()<1>();
      else with_two.template operator()<0>();
    } else {
      with_two.template operator()<-1>();
    }
    // the chunk's tiles of this replica pass: one sum per counter, into one
    // of the row's slots (one cache line each)
    wave_accumulate(nacc, de, acc + (int64_t)(rep.acc_row + p / g.K) * kBatchRow, gw & (kBatchSlots - 1));
  }
}

constexpr int kNumVariants = 12;   // 4 lattice styles x 3 spin storage types

__global__ void k_fill_u64(unsigned long long *__restrict__ v, int64_t n) {
  const int64_t t = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
  if (t < n) v[t] = 0ull;
}

unsigned grid_for(int64_t n, int bs) { return (unsigned)((n + bs - 1) / bs); }

int64_t round_up(int64_t x, int64_t a) { return (x + a - 1) / a * a; }

struct Replica {
  bool live = false;
  spk_hip_replica_desc desc;
  Geom g;
  Rule rule;
  int stype = kI32;
  int64_t off = 0, bytes = 0;       // pool extent
  int64_t spin_off = 0, seed_off = 0;
  int done = 0;                     // sweeps run
  int64_t naccept = 0;
  int64_t energy = 0;
  std::vector<int64_t> trace_acc;   // cumulative per sweep
  std::vector<int64_t> trace_eng;
  int block_sweeps = 0;             // sweeps in the current block
  int acc_row = 0;
  bool energy_pending = false;      // initial energy on its way to pinned memory
  int variant = 0;                  // lat*3 + stype: which step kernel runs it
};

// first-fit extent allocator with coalescing
class PoolAlloc {
 public:
  explicit PoolAlloc(int64_t size) { free_[0] = size; }
  bool alloc(int64_t n, int64_t *off) {
    for (auto it = free_.begin(); it != free_.end(); ++it)
      if (it->second >= n) {
        *off = it->first;
        const int64_t rest = it->second - n, at = it->first + n;
        free_.erase(it);
        if (rest) free_[at] = rest;
        used_ += n;
        return true;
      }
    return false;
  }
  void release(int64_t off, int64_t n) {
    used_ -= n;
    auto it = free_.emplace(off, n).first;
    auto nx = std::next(it);
    if (nx != free_.end() && it->first + it->second == nx->first) {
      it->second += nx->second;
      free_.erase(nx);
    }
    if (it != free_.begin()) {
      auto pv = std::prev(it);
      if (pv->first + pv->second == it->first) {
        pv->second += it->second;
        free_.erase(it);
      }
    }
  }
  int64_t used() const { return used_; }
  int64_t largest() const {
    int64_t m = 0;
    for (auto &f : free_) m = std::max(m, f.second);
    return m;
  }

 private:
  std::map<int64_t, int64_t> free_;   // offset -> size
  int64_t used_ = 0;
};

struct Shape {
  int block, grid;
};

}  // namespace

struct spk_hip_batch {
  int device = 0;
  int ncu = 256;
  hipStream_t stream = nullptr;
  char *pool = nullptr;
  int64_t pool_bytes = 0;
  PoolAlloc *alloc = nullptr;
  std::vector<Replica> reps;        // slot table
  // staging for transfers and per-replica set-up
  int32_t *stage = nullptr;
  int64_t stage_n = 0;
  // block tables (device) and their host images
  BatchRep *d_reps = nullptr;
  StepDesc *d_steps = nullptr;
  int *d_act = nullptr, *d_pre = nullptr;
  unsigned long long *d_acc = nullptr;    // [rows][kBatchRow] slots
  unsigned long long *d_accc = nullptr;   // [rows][2] folded (Naccept, dE)
  int cap_reps = 0, cap_steps = 0, cap_act = 0, cap_pre = 0, cap_acc = 0, cap_accc = 0;
  std::vector<BatchRep> h_reps;
  std::vector<StepDesc> h_steps;
  std::vector<int> h_act, h_pre;
  std::vector<unsigned long long> h_acc;
  // variants present in the current block, one stream each (fork/join)
  std::vector<int> variants;
  hipStream_t vstream[kNumVariants] = {};
  hipEvent_t fork_ev = nullptr, join_ev[kNumVariants] = {};
  // launch shape
  int block = 0, grid = 0;          // 0 = autotune on the next run
  std::map<std::tuple<int, int, int, int>, hipGraphExec_t> graphs;   // (L, block, grid, variant mask)
  std::vector<hipEvent_t> events;
  double last_ms = 0.0;
  int64_t last_updates = 0, last_nlaunch = 0;
  // a block launched and not yet completed (batch_launch / batch_complete)
  bool inflight = false;
  int inflight_rows = 0;
  unsigned long long *h_rows = nullptr;   // pinned: folded rows of the block in flight
  int cap_h_rows = 0;
  // per replica slot: an energy accumulator row, the folded initial energy
  // and the checksum, copied to pinned memory in stream order
  unsigned long long *d_erow = nullptr;   // [slots][kAccRow]
  unsigned long long *d_slot = nullptr;   // [slots][4]: energy, (unused), checksum, (unused)
  unsigned long long *h_slot = nullptr;   // pinned copy of d_slot
};

namespace {

// Grows a device table; returns 1 in *moved when it was reallocated (cached
// graphs then hold a stale pointer and must be dropped).
template <typename P>
int grow(P **ptr, int *cap, int need, const char *what, bool *moved) {
  if (need <= *cap) return SPK_HIP_OK;
  *moved = true;
  int c = std::max(need, *cap * 2);
  P *p = nullptr;
  if (hipMalloc(&p, (size_t)c * sizeof(P)) != hipSuccess)
    return set_error(SPK_HIP_ENOMEM, "device allocation of %s failed", what);
  if (*ptr) {
    (void)hipDeviceSynchronize();
    (void)hipFree(*ptr);
  }
  *ptr = p;
  *cap = c;
  return SPK_HIP_OK;
}

int check_id(spk_hip_batch *b, int id) {
  if (!b) return set_error(SPK_HIP_EINVAL, "null batch");
  if (id < 0 || id >= (int)b->reps.size() || !b->reps[id].live)
    return set_error(SPK_HIP_EINVAL, "no live replica %d", id);
  return SPK_HIP_OK;
}

int ensure_stage(spk_hip_batch *b, int64_t n) {
  if (n <= b->stage_n) return SPK_HIP_OK;
  int32_t *p = nullptr;
  if (hipMalloc(&p, (size_t)n * sizeof(int32_t)) != hipSuccess)
    return set_error(SPK_HIP_ENOMEM, "staging allocation failed");
  if (b->stage) {
    (void)hipStreamSynchronize(b->stream);
    (void)hipFree(b->stage);
  }
  b->stage = p;
  b->stage_n = n;
  return SPK_HIP_OK;
}

// Total energy of replica slot `id` (exact integer) on the batch stream,
// into the slot's device row and then its pinned copy; nothing waits.  The
// value is read by resolve_energies once the stream has passed it.
int replica_energy_async(spk_hip_batch *b, int id, Replica &R) {
  unsigned long long *row = b->d_erow + (size_t)id * kAccRow;
  hipLaunchKernelGGL(k_fill_u64, dim3(grid_for(kAccRow, 256)), dim3(256), 0, b->stream, row, (int64_t)kAccRow);
  const int64_t n = (int64_t)R.g.M * R.g.K;
  with_lattice(R.g.lat, [&]<int LAT>() {
    with_type(R.stype, [&]<typename T>() {
      hipLaunchKernelGGL((k_energy<LAT, T>), dim3(grid_for(n, 256)), dim3(256), 0, b->stream,
                         reinterpret_cast<const T *>(b->pool + R.spin_off), R.g, row);
    });
  });
  hipLaunchKernelGGL(k_fold_rows, dim3(1), dim3(64), 0, b->stream, row, kAccSlots, kAccRow, b->d_slot + 4 * (size_t)id);
  int rc = check_hip(hipGetLastError(), "energy launch");
  if (!rc)
    rc = check_hip(hipMemcpyAsync(b->h_slot + 4 * (size_t)id, b->d_slot + 4 * (size_t)id, sizeof(unsigned long long),
                                  hipMemcpyDeviceToHost, b->stream),
                   "energy readback");
  if (!rc) R.energy_pending = true;
  return rc;
}

// Takes the initial energies that have arrived in pinned memory; the
// caller has synchronised the batch stream.
void resolve_energies(spk_hip_batch *b) {
  for (int id = 0; id < (int)b->reps.size(); id++) {
    Replica &R = b->reps[id];
    if (!R.energy_pending) continue;
    R.energy = (int64_t)b->h_slot[4 * (size_t)id];
    R.energy_pending = false;
  }
}

// Queues the checksum of replica slot `id` (after everything already on the
// batch stream) into its device slot and pinned copy.
int replica_checksum_async(spk_hip_batch *b, int id) {
  const Replica &R = b->reps[id];
  unsigned long long *out = b->d_slot + 4 * (size_t)id + 2;
  hipLaunchKernelGGL(k_fill_u64, dim3(1), dim3(64), 0, b->stream, out, (int64_t)1);
  const int64_t n = (int64_t)R.g.M * R.g.K;
  with_type(R.stype, [&]<typename T>() {
    hipLaunchKernelGGL((k_checksum<T>), dim3(grid_for(n, 256)), dim3(256), 0, b->stream,
                       reinterpret_cast<const T *>(b->pool + R.spin_off),
                       reinterpret_cast<const int32_t *>(b->pool + R.seed_off), R.g, out);
  });
  int rc = check_hip(hipGetLastError(), "checksum launch");
  if (!rc)
    rc = check_hip(hipMemcpyAsync(b->h_slot + 4 * (size_t)id + 2, out, sizeof(unsigned long long),
                                  hipMemcpyDeviceToHost, b->stream),
                   "checksum readback");
  return rc;
}

// Builds the block's replica table and, per variant present, the step
// schedule.  Replica r with P_r passes is due at step j when
// floor((j+1)P_r/L) > floor(j P_r/L); its pass index is then floor(j P_r/L),
// so its passes run once each, in order.  Steps of variant k live at
// h_steps[k*L + j].
int build_schedule(spk_hip_batch *b, int max_sweeps, int *L_out, int *rows_out, int64_t *updates_out) {
  b->h_reps.clear();
  std::vector<int> vrep;   // variant of each table entry
  int L = 0, rows = 0;
  int64_t updates = 0;
  bool present[kNumVariants] = {};
  for (int id = 0; id < (int)b->reps.size(); id++) {
    Replica &R = b->reps[id];
    R.block_sweeps = 0;
    if (!R.live) continue;
    const int left = R.desc.nsweeps - R.done;
    if (left <= 0) continue;
    R.block_sweeps = std::min(left, max_sweeps);
    R.acc_row = rows;
    rows += R.block_sweeps;
    BatchRep br;
    br.g = R.g;
    br.rule = R.rule;
    br.stype = R.stype;
    br.P = R.g.K * R.block_sweeps;
    br.acc_row = R.acc_row;
    br.tiles = (int)((R.g.Mp / kVec + kTile - 1) / kTile);
    br.spin_off = R.spin_off;
    br.seed_off = R.seed_off;
    b->h_reps.push_back(br);
    vrep.push_back(R.variant);
    present[R.variant] = true;
    L = std::max(L, br.P);
    updates += (int64_t)R.g.M * R.g.K * R.block_sweeps;
  }
  b->variants.clear();
  for (int v = 0; v < kNumVariants; v++)
    if (present[v]) b->variants.push_back(v);
  b->h_steps.assign((size_t)L * b->variants.size(), StepDesc{0, 0, 0, 0});
  b->h_act.clear();
  b->h_pre.clear();
  for (size_t k = 0; k < b->variants.size(); k++)
    for (int j = 0; j < L; j++) {
      StepDesc &sd = b->h_steps[k * L + j];
      sd.first = (int)b->h_act.size();
      sd.pre_first = (int)b->h_pre.size();
      int tot = 0;
      b->h_pre.push_back(0);
      for (int r = 0; r < (int)b->h_reps.size(); r++) {
        if (vrep[r] != b->variants[k]) continue;
        const int64_t P = b->h_reps[r].P;
        if (((int64_t)(j + 1) * P) / L > ((int64_t)j * P) / L) {
          b->h_act.push_back(r);
          tot += b->h_reps[r].tiles;
          b->h_pre.push_back(tot);
        }
      }
      sd.count = (int)b->h_act.size() - sd.first;
      sd.total = tot;
    }
  *L_out = L;
  *rows_out = rows;
  *updates_out = updates;
  return SPK_HIP_OK;
}

void drop_graphs(spk_hip_batch *b) {
  for (auto &kv : b->graphs) (void)hipGraphExecDestroy(kv.second);
  b->graphs.clear();
}

int upload_schedule(spk_hip_batch *b, int rows, bool clear_acc) {
  int rc;
  bool moved = false;
  if ((rc = grow(&b->d_reps, &b->cap_reps, std::max<int>(1, b->h_reps.size()), "replica table", &moved)) ||
      (rc = grow(&b->d_steps, &b->cap_steps, std::max<int>(1, b->h_steps.size()), "step table", &moved)) ||
      (rc = grow(&b->d_act, &b->cap_act, std::max<int>(1, b->h_act.size()), "step lists", &moved)) ||
      (rc = grow(&b->d_pre, &b->cap_pre, std::max<int>(1, b->h_pre.size()), "step lists", &moved)) ||
      (rc = grow(&b->d_acc, &b->cap_acc, std::max(1, rows) * kBatchRow, "accumulators", &moved)) ||
      (rc = grow(&b->d_accc, &b->cap_accc, std::max(2, 2 * rows), "accumulators", &moved))) {
    if (moved) drop_graphs(b);   // an earlier table moved before a later one failed
    return rc;
  }
  if (moved) drop_graphs(b);   // captured launches hold the old table addresses
  auto put = [&](void *dst, const void *src, size_t n) {
    return n ? check_hip(hipMemcpyAsync(dst, src, n, hipMemcpyHostToDevice, b->stream), "table upload") : 0;
  };
  if ((rc = put(b->d_reps, b->h_reps.data(), b->h_reps.size() * sizeof(BatchRep))) ||
      (rc = put(b->d_steps, b->h_steps.data(), b->h_steps.size() * sizeof(StepDesc))) ||
      (rc = put(b->d_act, b->h_act.data(), b->h_act.size() * sizeof(int))) ||
      (rc = put(b->d_pre, b->h_pre.data(), b->h_pre.size() * sizeof(int))))
    return rc;
  if (!clear_acc) return SPK_HIP_OK;
  // accumulators cleared by a kernel (no memset node ever enters a graph)
  hipLaunchKernelGGL(k_fill_u64, dim3(grid_for((int64_t)rows * kBatchRow, 256)), dim3(256), 0, b->stream, b->d_acc,
                     (int64_t)rows * kBatchRow);
  return check_hip(hipGetLastError(), "accumulator clear");
}

int ensure_streams(spk_hip_batch *b) {
  int rc;
  if (!b->fork_ev && (rc = check_hip(hipEventCreateWithFlags(&b->fork_ev, hipEventDisableTiming), "event")))
    return rc;
  for (int v : b->variants) {
    if (!b->vstream[v] && (rc = check_hip(hipStreamCreate(&b->vstream[v]), "hipStreamCreate"))) return rc;
    if (!b->join_ev[v] && (rc = check_hip(hipEventCreateWithFlags(&b->join_ev[v], hipEventDisableTiming), "event")))
      return rc;
  }
  return SPK_HIP_OK;
}

// Steps j0..j1-1 of the block: each variant's chain on its own stream,
// forked from and joined back into the batch stream (so a graph captured on
// the batch stream holds one independent branch per variant).
int launch_steps(spk_hip_batch *b, int j0, int j1, int L, int block, int grid) {
  int rc;
  if ((rc = check_hip(hipEventRecord(b->fork_ev, b->stream), "fork"))) return rc;
  for (int v : b->variants)
    if ((rc = check_hip(hipStreamWaitEvent(b->vstream[v], b->fork_ev, 0), "fork wait"))) return rc;
  for (int j = j0; j < j1; j++)
    for (size_t k = 0; k < b->variants.size(); k++) {
      const int v = b->variants[k];
      const StepDesc *steps = b->d_steps + k * L;
      with_lattice(v / 3, [&]<int LAT>() {
        with_type(v % 3, [&]<typename T>() {
          hipLaunchKernelGGL((k_batch_step<LAT, T>), dim3(grid), dim3(block), 0, b->vstream[v], b->pool, b->d_reps,
                             steps, b->d_act, b->d_pre, j, L, b->d_acc);
        });
      });
      if ((rc = check_hip(hipGetLastError(), "step kernel launch"))) return rc;
    }
  for (int v : b->variants) {
    if ((rc = check_hip(hipEventRecord(b->join_ev[v], b->vstream[v]), "join"))) return rc;
    if ((rc = check_hip(hipStreamWaitEvent(b->stream, b->join_ev[v], 0), "join wait"))) return rc;
  }
  return SPK_HIP_OK;
}

int ensure_events(spk_hip_batch *b, size_t n) {
  while (b->events.size() < n) {
    hipEvent_t e;
    int rc = check_hip(hipEventCreate(&e), "hipEventCreate");
    if (rc) return rc;
    b->events.push_back(e);
  }
  return SPK_HIP_OK;
}

}  // namespace

extern "C" {

int spk_hip_replica_bytes(const spk_hip_desc *d, int64_t *bytes) {
  if (!bytes) return set_error(SPK_HIP_EINVAL, "null argument");
  Geom g;
  int rc = make_geom(d, &g);
  if (rc) return rc;
  const int64_t slots = (int64_t)g.K * g.Mp;
  *bytes = round_up(slots * (int64_t)spin_bytes(spin_type_for(d->app, d->nspins)), 256) +
           round_up(slots * 4, 256);
  return SPK_HIP_OK;
}

int spk_hip_batch_create(int64_t pool_bytes, int max_replicas, spk_hip_batch **out) {
  if (!out || pool_bytes <= 0 || max_replicas <= 0) return set_error(SPK_HIP_EINVAL, "bad batch arguments");
  *out = nullptr;
  spk_hip_batch *b = new (std::nothrow) spk_hip_batch();
  if (!b) return set_error(SPK_HIP_ENOMEM, "host allocation failed");
  int rc;
  hipDeviceProp_t prop;
  if ((rc = check_hip(hipGetDevice(&b->device), "hipGetDevice")) ||
      (rc = check_hip(hipGetDeviceProperties(&prop, b->device), "hipGetDeviceProperties")) ||
      (rc = check_hip(hipStreamCreate(&b->stream), "hipStreamCreate"))) {
    spk_hip_batch_destroy(b);
    return rc;
  }
  b->ncu = prop.multiProcessorCount;
  b->pool_bytes = round_up(pool_bytes, 256);
  if (hipMalloc(&b->pool, (size_t)b->pool_bytes) != hipSuccess) {
    spk_hip_batch_destroy(b);
    return set_error(SPK_HIP_ENOMEM, "device pool of %lld bytes not available", (long long)pool_bytes);
  }
  const size_t slots = (size_t)max_replicas;
  if (hipMalloc(&b->d_erow, slots * kAccRow * sizeof(unsigned long long)) != hipSuccess ||
      hipMalloc(&b->d_slot, slots * 4 * sizeof(unsigned long long)) != hipSuccess ||
      hipHostMalloc(&b->h_slot, slots * 4 * sizeof(unsigned long long), hipHostMallocDefault) != hipSuccess) {
    spk_hip_batch_destroy(b);
    return set_error(SPK_HIP_ENOMEM, "per-replica rows for %d replicas not available", max_replicas);
  }
  b->alloc = new PoolAlloc(b->pool_bytes);
  b->reps.resize(max_replicas);
  *out = b;
  return SPK_HIP_OK;
}

void spk_hip_batch_destroy(spk_hip_batch *b) {
  if (!b) return;
  if (b->stream) (void)hipStreamSynchronize(b->stream);
  drop_graphs(b);
  for (auto e : b->events) (void)hipEventDestroy(e);
  for (int v = 0; v < kNumVariants; v++) {
    if (b->vstream[v]) {
      (void)hipStreamSynchronize(b->vstream[v]);
      (void)hipStreamDestroy(b->vstream[v]);
    }
    if (b->join_ev[v]) (void)hipEventDestroy(b->join_ev[v]);
  }
  if (b->fork_ev) (void)hipEventDestroy(b->fork_ev);
  if (b->pool) (void)hipFree(b->pool);
  if (b->d_erow) (void)hipFree(b->d_erow);
  if (b->d_slot) (void)hipFree(b->d_slot);
  if (b->h_slot) (void)hipHostFree(b->h_slot);
  if (b->h_rows) (void)hipHostFree(b->h_rows);
  if (b->stage) (void)hipFree(b->stage);
  if (b->d_reps) (void)hipFree(b->d_reps);
  if (b->d_steps) (void)hipFree(b->d_steps);
  if (b->d_act) (void)hipFree(b->d_act);
  if (b->d_pre) (void)hipFree(b->d_pre);
  if (b->d_acc) (void)hipFree(b->d_acc);
  if (b->d_accc) (void)hipFree(b->d_accc);
  if (b->stream) (void)hipStreamDestroy(b->stream);
  delete b->alloc;
  delete b;
}

}  // extern "C"

// Admits a replica and queues its set-up and initial energy on the batch
// stream without waiting (after a block in flight, if any); the energy
// arrives in pinned memory and is taken when the stream is next waited for.
int spk::batch_add_async(spk_hip_batch *b, const spk_hip_replica_desc *r, int *id) {
  if (!b || !r || !id) return set_error(SPK_HIP_EINVAL, "null argument");
  Replica R;
  int rc = make_geom(&r->lattice, &R.g);
  if (rc) return rc;
  if ((rc = make_rule(r->lattice.app, r->lattice.nspins, r->lattice.lattice, r->lattice.temperature, &R.rule)))
    return rc;
  if (r->seed <= 0) return set_error(SPK_HIP_EINVAL, "Illegal seed command");
  if (r->nsweeps < 0) return set_error(SPK_HIP_EINVAL, "nsweeps < 0");
  const int qmax = app_qmax(r->lattice.app, r->lattice.nspins);
  if (r->lo < 1 || r->hi > qmax || r->hi < r->lo)
    return set_error(SPK_HIP_EVALUE, "One or more sites have invalid values (set %d %d, Q = %d)", r->lo, r->hi,
                     qmax);
  int slot = -1;
  for (int k = 0; k < (int)b->reps.size(); k++)
    if (!b->reps[k].live) {
      slot = k;
      break;
    }
  if (slot < 0) return set_error(SPK_HIP_ENOMEM, "all %d replica slots are in use", (int)b->reps.size());
  if (variant_app(r->lattice.app))
    return set_error(SPK_HIP_EINVAL, "batched replicas run app styles potts and ising (got app %d)", r->lattice.app);
  if (!offsets_fit_32(R.g))
    return set_error(SPK_HIP_EINVAL,
                     "a batched replica holds fewer than 2^30 sites and rows shorter than 2^25 sites "
                     "(got %lld sites); run it alone",
                     (long long)R.g.K * R.g.M);
  R.stype = spin_type_for(r->lattice.app, r->lattice.nspins);
  R.variant = R.g.lat * 3 + R.stype;
  const int64_t slots = (int64_t)R.g.K * R.g.Mp;
  const int64_t sbytes = round_up(slots * (int64_t)spin_bytes(R.stype), 256);
  R.bytes = sbytes + round_up(slots * 4, 256);
  if (!b->alloc->alloc(R.bytes, &R.off))
    return set_error(SPK_HIP_ENOMEM, "pool has no free extent of %lld bytes", (long long)R.bytes);
  R.spin_off = R.off;
  R.seed_off = R.off + sbytes;
  R.desc = *r;
  // the replica script's master draws: `set` takes draw 0; the first run's
  // init takes draws 1-2 (ranapp) and 3 (ranstrict); draw 4 is rseed
  MasterRng m(r->seed);
  double draws[5];
  for (double &x : draws) x = m.uniform();
  const int seed0 = park_seed_from_draw(draws[0]);
  const int64_t n = (int64_t)R.g.M * R.g.K;
  const unsigned gs = grid_for(n, 256);
  int32_t *seed = reinterpret_cast<int32_t *>(b->pool + R.seed_off);
  with_type(R.stype, [&]<typename T>() {
    T *sp = reinterpret_cast<T *>(b->pool + R.spin_off);
    if (r->lo == r->hi)
      hipLaunchKernelGGL((k_fill<T>), dim3(grid_for(slots, 256)), dim3(256), 0, b->stream, sp, slots, r->lo);
    else
      hipLaunchKernelGGL((k_set_range<T>), dim3(gs), dim3(256), 0, b->stream, sp, R.g, seed0, r->lo,
                         r->hi - r->lo + 1, park_powers());
  });
  hipLaunchKernelGGL(k_init_seeds, dim3(gs), dim3(256), 0, b->stream, seed, R.g, draws[4] * (double)kIM,
                     park_warmup100());
  if ((rc = check_hip(hipGetLastError(), "replica set-up launch")) ||
      (rc = replica_energy_async(b, slot, R))) {
    b->alloc->release(R.off, R.bytes);
    return rc;
  }
  R.live = true;
  b->reps[slot] = std::move(R);
  *id = slot;
  return SPK_HIP_OK;
}

extern "C" {

int spk_hip_batch_add(spk_hip_batch *b, const spk_hip_replica_desc *r, int *id) {
  if (!b || !r || !id) return set_error(SPK_HIP_EINVAL, "null argument");
  int rc;
  if ((rc = batch_complete(b)) || (rc = batch_add_async(b, r, id))) return rc;
  if ((rc = check_hip(hipStreamSynchronize(b->stream), "replica set-up sync"))) return rc;
  resolve_energies(b);
  return SPK_HIP_OK;
}

int spk_hip_batch_remove(spk_hip_batch *b, int id) {
  int rc = check_id(b, id);
  if (rc || (rc = batch_complete(b))) return rc;
  Replica &R = b->reps[id];
  b->alloc->release(R.off, R.bytes);
  R = Replica();
  return SPK_HIP_OK;
}

int spk_hip_batch_upload(spk_hip_batch *b, int id, const int32_t *spin, const int32_t *seed) {
  int rc = check_id(b, id);
  if (rc || (rc = batch_complete(b))) return rc;
  Replica &R = b->reps[id];
  const int64_t n = (int64_t)R.g.M * R.g.K;
  if ((rc = ensure_stage(b, n))) return rc;
  const int qmax = app_qmax(R.desc.lattice.app, R.desc.lattice.nspins);
  for (int w = 0; w < 2; w++) {
    const int32_t *src = w == 0 ? spin : seed;
    if (!src) continue;
    const int lo = 1, hi = w == 0 ? qmax : kIM - 1;
    for (int64_t k = 0; k < n; k++)
      if (src[k] < lo || src[k] > hi)
        return set_error(SPK_HIP_EVALUE, w == 0 ? "One or more sites have invalid values" : "seed out of range");
    if ((rc = check_hip(hipMemcpyAsync(b->stage, src, n * 4, hipMemcpyHostToDevice, b->stream), "upload")))
      return rc;
    if (w == 0) {
      with_type(R.stype, [&]<typename T>() {
        hipLaunchKernelGGL((k_to_blocked<T>), dim3(grid_for(n, 256)), dim3(256), 0, b->stream, b->stage,
                           reinterpret_cast<T *>(b->pool + R.spin_off), R.g);
      });
    } else {
      hipLaunchKernelGGL(k_to_blocked<int32_t>, dim3(grid_for(n, 256)), dim3(256), 0, b->stream, b->stage,
                         reinterpret_cast<int32_t *>(b->pool + R.seed_off), R.g);
    }
    if ((rc = check_hip(hipGetLastError(), "layout kernel"))) return rc;
  }
  if (spin && (rc = replica_energy_async(b, id, R))) return rc;
  if ((rc = check_hip(hipStreamSynchronize(b->stream), "upload sync"))) return rc;
  resolve_energies(b);
  return SPK_HIP_OK;
}

int spk_hip_batch_download(spk_hip_batch *b, int id, int32_t *spin, int32_t *seed) {
  int rc = check_id(b, id);
  if (rc || (rc = batch_complete(b))) return rc;
  Replica &R = b->reps[id];
  const int64_t n = (int64_t)R.g.M * R.g.K;
  if ((rc = ensure_stage(b, n))) return rc;
  for (int w = 0; w < 2; w++) {
    int32_t *dst = w == 0 ? spin : seed;
    if (!dst) continue;
    if (w == 0) {
      with_type(R.stype, [&]<typename T>() {
        hipLaunchKernelGGL((k_to_natural<T>), dim3(grid_for(n, 256)), dim3(256), 0, b->stream,
                           reinterpret_cast<const T *>(b->pool + R.spin_off), b->stage, R.g);
      });
    } else {
      hipLaunchKernelGGL(k_to_natural<int32_t>, dim3(grid_for(n, 256)), dim3(256), 0, b->stream,
                         reinterpret_cast<const int32_t *>(b->pool + R.seed_off), b->stage, R.g);
    }
    if ((rc = check_hip(hipGetLastError(), "layout kernel"))) return rc;
    if ((rc = check_hip(hipMemcpyAsync(dst, b->stage, n * 4, hipMemcpyDeviceToHost, b->stream), "download")))
      return rc;
    if ((rc = check_hip(hipStreamSynchronize(b->stream), "download sync"))) return rc;
  }
  return SPK_HIP_OK;
}

int spk_hip_batch_set_launch(spk_hip_batch *b, int block, int grid) {
  if (!b) return set_error(SPK_HIP_EINVAL, "null batch");
  if (block == 0 && grid == 0) {
    b->block = b->grid = 0;
    return SPK_HIP_OK;
  }
  if (block < 64 || block > 1024 || block % 64 || grid < 1)
    return set_error(SPK_HIP_EINVAL, "block must be a multiple of 64 in [64,1024] and grid >= 1");
  b->block = block;
  b->grid = (grid + kNumXcd - 1) / kNumXcd * kNumXcd;
  return SPK_HIP_OK;
}

}  // extern "C"

namespace spk {

bool batch_inflight(const spk_hip_batch *b) { return b && b->inflight; }

// Folds the block's accumulator rows and queues their copy to pinned memory.
static int queue_rows(spk_hip_batch *b, int rows) {
  if (rows > b->cap_h_rows) {   // only between blocks: nothing reads h_rows now
    unsigned long long *p = nullptr;
    const int cap = std::max(rows, 2 * b->cap_h_rows);
    if (hipHostMalloc(&p, (size_t)cap * 2 * sizeof(unsigned long long), hipHostMallocDefault) != hipSuccess)
      return set_error(SPK_HIP_ENOMEM, "pinned rows for %d sweeps not available", cap);
    if (b->h_rows) (void)hipHostFree(b->h_rows);
    b->h_rows = p;
    b->cap_h_rows = cap;
  }
  hipLaunchKernelGGL(k_fold_rows, dim3(rows), dim3(64), 0, b->stream, b->d_acc, kBatchSlots, kBatchRow, b->d_accc);
  int rc = check_hip(hipGetLastError(), "fold launch");
  if (!rc)
    rc = check_hip(hipMemcpyAsync(b->h_rows, b->d_accc, 2 * (size_t)rows * sizeof(unsigned long long),
                                  hipMemcpyDeviceToHost, b->stream),
                   "accumulator readback");
  return rc;
}

// After the stream has passed the block: each replica's per-sweep Naccept
// and energy from the folded rows.
static void apply_rows(spk_hip_batch *b) {
  resolve_energies(b);
  for (auto &R : b->reps) {
    if (!R.live || R.block_sweeps == 0) continue;
    for (int s = 0; s < R.block_sweeps; s++) {
      R.naccept += (int64_t)b->h_rows[2 * (size_t)(R.acc_row + s)];
      R.energy += 2 * (int64_t)b->h_rows[2 * (size_t)(R.acc_row + s) + 1];
      R.trace_acc.push_back(R.naccept);
      R.trace_eng.push_back(R.energy);
    }
    R.done += R.block_sweeps;
    R.block_sweeps = 0;
  }
}

// Queues one block (every replica with sweeps left advances by up to
// max_sweeps) with the launch shape already set, and returns without
// waiting.  Work queued on the batch stream afterwards (admissions,
// checksums) runs after the block.
int batch_launch(spk_hip_batch *b, int max_sweeps, bool use_graph) {
  if (b->inflight) return set_error(SPK_HIP_ESTATE, "a block is already in flight");
  if (b->block == 0) return set_error(SPK_HIP_ESTATE, "no launch shape set");
  int rc;
  int L = 0, rows = 0;
  int64_t updates = 0;
  if ((rc = build_schedule(b, max_sweeps, &L, &rows, &updates))) return rc;
  b->last_updates = updates;
  b->last_nlaunch = (int64_t)L * (int64_t)std::max<size_t>(1, b->variants.size());
  b->last_ms = 0.0;
  if (L == 0) return SPK_HIP_OK;
  if ((rc = ensure_events(b, 2 * (size_t)L + 2)) || (rc = ensure_streams(b)) ||
      (rc = upload_schedule(b, rows, true)))
    return rc;
  const Shape sh{b->block, b->grid};
  int mask = 0;
  for (int v : b->variants) mask |= 1 << v;
  if (use_graph) {
    auto key = std::make_tuple(L, sh.block, sh.grid, mask);
    auto it = b->graphs.find(key);
    if (it == b->graphs.end()) {
      hipGraph_t graph = nullptr;
      if ((rc = check_hip(hipStreamBeginCapture(b->stream, hipStreamCaptureModeThreadLocal), "begin capture")))
        return rc;
      rc = launch_steps(b, 0, L, L, sh.block, sh.grid);
      const hipError_t e = hipStreamEndCapture(b->stream, &graph);
      if (!rc) rc = check_hip(e, "end capture");
      if (rc) {
        if (graph) (void)hipGraphDestroy(graph);
        return rc;
      }
      hipGraphExec_t exec;
      rc = check_hip(hipGraphInstantiate(&exec, graph, nullptr, nullptr, 0), "graph instantiate");
      (void)hipGraphDestroy(graph);
      if (rc) return rc;
      it = b->graphs.emplace(key, exec).first;
    }
    if ((rc = check_hip(hipEventRecord(b->events[0], b->stream), "event")) ||
        (rc = check_hip(hipGraphLaunch(it->second, b->stream), "graph launch")) ||
        (rc = check_hip(hipEventRecord(b->events[1], b->stream), "event")))
      return rc;
  } else {
    if ((rc = check_hip(hipEventRecord(b->events[0], b->stream), "event")) ||
        (rc = launch_steps(b, 0, L, L, sh.block, sh.grid)) ||
        (rc = check_hip(hipEventRecord(b->events[1], b->stream), "event")))
      return rc;
  }
  if ((rc = queue_rows(b, rows))) return rc;
  b->inflight = true;
  b->inflight_rows = rows;
  return SPK_HIP_OK;
}

// Waits for the batch stream; applies the rows of the block in flight (if
// any) and the initial energies that have arrived.
int batch_complete(spk_hip_batch *b) {
  int rc;
  if ((rc = check_hip(hipStreamSynchronize(b->stream), "run sync"))) return rc;
  if (!b->inflight) {
    resolve_energies(b);
    return SPK_HIP_OK;
  }
  b->inflight = false;
  float ms = 0.f;
  if ((rc = check_hip(hipEventElapsedTime(&ms, b->events[0], b->events[1]), "event time"))) return rc;
  b->last_ms = ms;
  apply_rows(b);
  return SPK_HIP_OK;
}

// Queues the checksum of a replica after everything on the batch stream;
// batch_checksum_value reads it once the stream has passed it.
int batch_queue_checksum(spk_hip_batch *b, int id) { return replica_checksum_async(b, id); }
uint64_t batch_checksum_value(const spk_hip_batch *b, int id) { return b->h_slot[4 * (size_t)id + 2]; }

int batch_left(const spk_hip_batch *b, int id) {
  const Replica &R = b->reps[id];
  return R.live ? R.desc.nsweeps - R.done : 0;
}

}  // namespace spk

extern "C" {

int spk_hip_batch_run(spk_hip_batch *b, int max_sweeps, int use_graph) {
  if (!b) return set_error(SPK_HIP_EINVAL, "null batch");
  if (max_sweeps <= 0) return set_error(SPK_HIP_EINVAL, "max_sweeps must be positive");
  int rc;
  if ((rc = batch_complete(b))) return rc;
  if (b->block != 0) {
    if ((rc = batch_launch(b, max_sweeps, use_graph != 0))) return rc;
    return batch_complete(b);
  }
  // autotune on this block's own steps, then its rows as a launched block's
  int L = 0, rows = 0;
  int64_t updates = 0;
  if ((rc = build_schedule(b, max_sweeps, &L, &rows, &updates))) return rc;
  b->last_updates = updates;
  b->last_nlaunch = (int64_t)L * (int64_t)std::max<size_t>(1, b->variants.size());
  if (L == 0) {
    b->last_ms = 0.0;
    return SPK_HIP_OK;
  }
  if ((rc = ensure_events(b, 2 * (size_t)L + 2)) || (rc = ensure_streams(b)) ||
      (rc = upload_schedule(b, rows, true)))
    return rc;
  if (b->block == 0) {
    // autotune: candidate shapes (workgroup size x wavefronts per CU) run
    // the block's own steps, as a replay runs them (no wait between steps);
    // results do not depend on the shape.  The tile list is the same for
    // every shape (a tile is one wavefront's work).
    std::vector<Shape> cand;
    for (int bs : {64, 128, 256})
      for (int wpc : {8, 16, 32}) {
        const int grid = std::max(kNumXcd, b->ncu * wpc * 64 / bs / kNumXcd * kNumXcd);
        cand.push_back(Shape{bs, grid});
      }
    auto step_work = [&](int j) {   // sites updated in step j
      double w = 0.0;
      for (size_t k = 0; k < b->variants.size(); k++) {
        const StepDesc &sd = b->h_steps[k * L + j];
        for (int a = 0; a < sd.count; a++) w += (double)b->h_reps[b->h_act[sd.first + a]].g.M;
      }
      return w;
    };
    const int C = (int)cand.size();
    // each candidate gets whole sweeps of an 8-colour replica when the
    // block has room, else single steps in turn
    const int per = L / C >= 8 ? L / C / 8 * 8 : 1;
    std::vector<double> t_ms(C, 0.0), work(C, 0.0);
    double tot = 0.0;
    int j = 0;
    auto timed = [&](int j0, int j1, const Shape &sh, double *ms_out) -> int {
      int rc;
      if ((rc = check_hip(hipEventRecord(b->events[0], b->stream), "event")) ||
          (rc = launch_steps(b, j0, j1, L, sh.block, sh.grid)) ||
          (rc = check_hip(hipEventRecord(b->events[1], b->stream), "event")) ||
          (rc = check_hip(hipStreamSynchronize(b->stream), "tune sync")))
        return rc;
      float ms = 0.f;
      if ((rc = check_hip(hipEventElapsedTime(&ms, b->events[0], b->events[1]), "event time"))) return rc;
      *ms_out = ms;
      return SPK_HIP_OK;
    };
    auto trial = [&](int c) -> int {   // candidate c takes the next `per` steps
      const int j1 = std::min(L, j + per);
      double ms = 0.0;
      const int rc = timed(j, j1, cand[c], &ms);
      if (rc) return rc;
      for (int q = j; q < j1; q++) work[c] += step_work(q);
      t_ms[c] += ms;
      tot += ms;
      j = j1;
      return SPK_HIP_OK;
    };
    if (per > 1) {
      for (int c = 0; c < C; c++)
        if ((rc = trial(c))) return rc;
    } else {
      while (j < L)
        for (int c = 0; c < C && j < L; c++)
          if ((rc = trial(c))) return rc;
    }
    int best = 0;
    for (int c = 1; c < C; c++)
      if (work[c] / std::max(t_ms[c], 1e-9) > work[best] / std::max(t_ms[best], 1e-9)) best = c;
    b->block = cand[best].block;
    b->grid = cand[best].grid;
    if (j < L) {   // the steps after the trial chunks run with the winner
      double ms = 0.0;
      if ((rc = timed(j, L, cand[best], &ms))) return rc;
      tot += ms;
    }
    b->last_ms = tot;
  }
  if ((rc = queue_rows(b, rows)) || (rc = check_hip(hipStreamSynchronize(b->stream), "readback sync")))
    return rc;
  apply_rows(b);
  return SPK_HIP_OK;
}

int spk_hip_batch_status(spk_hip_batch *b, int id, int *done, int *left, int64_t *naccept, double *energy) {
  int rc = check_id(b, id);
  if (rc || (rc = batch_complete(b))) return rc;
  const Replica &R = b->reps[id];
  if (done) *done = R.done;
  if (left) *left = R.desc.nsweeps - R.done;
  if (naccept) *naccept = R.naccept;
  if (energy) *energy = (double)R.energy;
  return SPK_HIP_OK;
}

int spk_hip_batch_trace(spk_hip_batch *b, int id, int n, int64_t *naccept, double *energy) {
  int rc = check_id(b, id);
  if (rc || (rc = batch_complete(b))) return rc;
  const Replica &R = b->reps[id];
  if (n < 0 || n > (int)R.trace_acc.size()) return set_error(SPK_HIP_EINVAL, "trace has %d sweeps", (int)R.trace_acc.size());
  for (int s = 0; s < n; s++) {
    if (naccept) naccept[s] = R.trace_acc[s];
    if (energy) energy[s] = (double)R.trace_eng[s];
  }
  return SPK_HIP_OK;
}

int spk_hip_batch_last_timing(spk_hip_batch *b, double *ms, int64_t *updates, int64_t *nlaunch, int *block,
                              int *grid) {
  if (!b) return set_error(SPK_HIP_EINVAL, "null batch");
  if (ms) *ms = b->last_ms;
  if (updates) *updates = b->last_updates;
  if (nlaunch) *nlaunch = b->last_nlaunch;
  if (block) *block = b->block;
  if (grid) *grid = b->grid;
  return SPK_HIP_OK;
}

int spk_hip_batch_pool(spk_hip_batch *b, int64_t *used, int64_t *largest_free, int *live) {
  if (!b) return set_error(SPK_HIP_EINVAL, "null batch");
  if (used) *used = b->alloc->used();
  if (largest_free) *largest_free = b->alloc->largest();
  if (live) {
    int n = 0;
    for (auto &R : b->reps) n += R.live;
    *live = n;
  }
  return SPK_HIP_OK;
}

int spk_hip_batch_checksum(spk_hip_batch *b, int id, uint64_t *sum) {
  int rc = check_id(b, id);
  if (rc) return rc;
  if (!sum) return set_error(SPK_HIP_EINVAL, "null argument");
  if ((rc = batch_complete(b)) || (rc = replica_checksum_async(b, id)) ||
      (rc = check_hip(hipStreamSynchronize(b->stream), "sync")))
    return rc;
  *sum = b->h_slot[4 * (size_t)id + 2];
  return SPK_HIP_OK;
}

}  // extern "C"
