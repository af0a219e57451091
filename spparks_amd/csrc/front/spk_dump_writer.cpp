// spk_dump_writer.cpp -- see spk_dump_writer.h.
#include "spk_dump_writer.h"

#include <algorithm>
#include <charconv>
#include <cstdlib>

namespace spkfront {

namespace {

// #unlike neighbours of site (i, j, k) in its periodic shell (the energy
// dump field: AppPotts / AppIsing::site_energy over the lattice's stencil)
int site_energy(const DumpJob &d, int64_t i, int64_t j, int64_t k, int32_t me) {
  int e = 0;
  for (int dk = -1; dk <= 1; dk++) {
    if (d.dimension == 2 && dk) continue;
    for (int dj = -1; dj <= 1; dj++)
      for (int di = -1; di <= 1; di++) {
        const int nb = std::abs(di) + std::abs(dj) + std::abs(dk);
        if (nb == 0 || (d.nearest_only && nb != 1)) continue;
        const int64_t a = (i + di + d.nx) % d.nx, b = (j + dj + d.ny) % d.ny, c = (k + dk + d.nz) % d.nz;
        e += d.spin[a + d.nx * (b + d.ny * c)] != me;
      }
  }
  return e;
}

}  // namespace

bool write_snapshot(const DumpJob &d) {
  bool ok = std::fwrite(d.header.data(), 1, d.header.size(), d.fp) == d.header.size();
  // a value is at most 24 characters ("%g" of a double, a 64-bit integer)
  const size_t line_max = 32 * d.fields.size() + 2;
  std::vector<char> buf(std::max<size_t>(1 << 20, 4 * line_max));
  char *const base = buf.data(), *const end = base + buf.size();
  char *p = base;
  const int64_t n = d.nx * d.ny * d.nz;
  for (int64_t g = 0; g < n; g++) {
    if ((size_t)(end - p) < line_max) {
      ok &= std::fwrite(base, 1, p - base, d.fp) == (size_t)(p - base);
      p = base;
    }
    const int64_t i = g % d.nx, j = (g / d.nx) % d.ny, k = g / (d.nx * d.ny);
    for (DumpField f : d.fields) {
      switch (f) {
        case DumpField::Id: p = std::to_chars(p, end, (long long)(g + 1)).ptr; break;
        case DumpField::Site: p = std::to_chars(p, end, d.spin[g]).ptr; break;
        case DumpField::Seed: p = std::to_chars(p, end, d.seed[g]).ptr; break;
        case DumpField::X: p += std::snprintf(p, end - p, "%g", i * d.latconst + d.boxlo[0]); break;
        case DumpField::Y: p += std::snprintf(p, end - p, "%g", j * d.latconst + d.boxlo[1]); break;
        case DumpField::Z:
          p += std::snprintf(p, end - p, "%g",
                             d.dimension == 3 ? k * d.latconst + d.boxlo[2] : 0.5 * (d.boxlo[2] + d.boxhi[2]));
          break;
        case DumpField::Energy:
          p += std::snprintf(p, end - p, "%g", (double)site_energy(d, i, j, k, d.spin[g]));
          break;
      }
      *p++ = ' ';
    }
    *p++ = '\n';
  }
  ok &= std::fwrite(base, 1, p - base, d.fp) == (size_t)(p - base);
  if (d.close) ok &= std::fclose(d.fp) == 0;
  else ok &= std::fflush(d.fp) == 0;
  return ok;
}

DumpWriter::~DumpWriter() {
  {
    std::lock_guard<std::mutex> lk(mu_);
    stop_ = true;
  }
  cv_.notify_all();
  if (worker_.joinable()) worker_.join();
}

void DumpWriter::submit(DumpJob &&job) {
  std::unique_lock<std::mutex> lk(mu_);
  if (!worker_.joinable()) worker_ = std::thread(&DumpWriter::loop, this);
  cv_.wait(lk, [&] { return queue_.size() < kMaxQueued; });
  queue_.push_back(std::move(job));
  cv_.notify_all();
}

bool DumpWriter::drain() {
  std::unique_lock<std::mutex> lk(mu_);
  cv_.wait(lk, [&] { return queue_.empty() && !busy_; });
  const bool ok = !failed_;
  failed_ = false;
  return ok;
}

// Takes jobs in order until stopped with nothing left to write.
void DumpWriter::loop() {
  std::unique_lock<std::mutex> lk(mu_);
  for (;;) {
    cv_.wait(lk, [&] { return stop_ || !queue_.empty(); });
    if (queue_.empty()) return;
    DumpJob job = std::move(queue_.front());
    queue_.pop_front();
    busy_ = true;
    cv_.notify_all();   // room in the queue
    lk.unlock();
    const bool ok = write_snapshot(job);
    job = DumpJob();    // release the snapshot before taking the lock
    lk.lock();
    busy_ = false;
    failed_ |= !ok;
    cv_.notify_all();   // drain() may be waiting
  }
}

}  // namespace spkfront
