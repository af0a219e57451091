// spk_dump_writer.h -- formatting and writing of dump snapshots on a
// background thread, so that a run's next sweeps go on on the GPU while the
// host writes the previous snapshot (the reference writes each snapshot
// inline in Output::compute, src/output.cpp:145-202, which leaves the GPU
// idle for the whole write here).
//
// A snapshot is copied from the device into its own job when the dump is
// due, so what the file holds is the state at that time; the jobs are
// written one at a time in submission order, so files and snapshots within
// a file keep the reference's order.
#pragma once

#include <condition_variable>
#include <cstdint>
#include <cstdio>
#include <deque>
#include <mutex>
#include <string>
#include <thread>
#include <vector>

namespace spkfront {

enum class DumpField : uint8_t { Id, Site, Seed, X, Y, Z, Energy };

// Everything one snapshot's lines need, owned by the job.
struct DumpJob {
  FILE *fp = nullptr;
  bool close = false;               // one file per snapshot: closed after writing
  std::string header;               // the style's header lines
  std::vector<DumpField> fields;
  std::vector<int32_t> spin, seed;  // id order; seed is empty without a seed field
  int64_t nx = 0, ny = 0, nz = 0;
  int dimension = 3;
  bool nearest_only = false;        // sq/4n, sc/6n: the energy field counts the nearest shell
  double latconst = 0.0;
  double boxlo[3] = {0, 0, 0}, boxhi[3] = {0, 0, 0};
};

// One line per site in id order, each value followed by a space
// (DumpText::write_text, src/dump_text.cpp:405-421): integers as %d, the
// coordinate and energy fields as %g.  Returns false on a write error.
bool write_snapshot(const DumpJob &job);

class DumpWriter {
 public:
  DumpWriter() = default;
  ~DumpWriter();
  DumpWriter(const DumpWriter &) = delete;
  DumpWriter &operator=(const DumpWriter &) = delete;

  // Queues a snapshot; waits while kMaxQueued are already waiting.
  void submit(DumpJob &&job);
  // Returns once every submitted snapshot is written.  Returns false if a
  // write failed since the last drain.
  bool drain();

 private:
  static constexpr size_t kMaxQueued = 1;
  void loop();

  std::thread worker_;
  std::mutex mu_;
  std::condition_variable cv_;
  std::deque<DumpJob> queue_;
  bool busy_ = false, stop_ = false, failed_ = false;
};

}  // namespace spkfront
