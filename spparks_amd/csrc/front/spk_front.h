// spk_front.h -- input-script front end of the spk_hip program: the subset of
// SPPARKS' command language that drives on-lattice Potts / Ising models with
// the strict colour sweep, executed on the GPU through libspk_hip.so.
//
// Mirrors the reference's Input / App / AppLattice / Output objects for this
// path (src/input.cpp:426-516, src/app.cpp:86-148, src/app_lattice.cpp:133-463,
// :632-711, src/output.cpp:82-202, :340-437, src/finish.cpp:30-120): same
// commands, argument meaning, master-RNG draw order, stats timing and line
// formats.  Errors throw spkfront::Error carrying the reference's message.
#pragma once

#include <cstdint>
#include <cstdio>
#include <map>
#include <memory>
#include <stdexcept>
#include <string>
#include <vector>

#include "spk_dump_writer.h"
#include "spk_hip.h"
#include "spk_master_rng.h"

namespace spkfront {

class Error : public std::runtime_error {
 public:
  explicit Error(const std::string &m) : std::runtime_error(m) {}
};

struct Region {
  double lo[3], hi[3];
};

struct Dump {
  std::string id, file;
  std::string style;                 // "text" or "sites"
  bool multifile = false;            // '*' in the file name: one file per snapshot
  double delta = 0.0;
  std::vector<std::string> fields;   // id site i1 x y z energy seed
  double next_time = 0.0;
  int idump = 0;
  FILE *fp = nullptr;
  // dump_modify (src/dump.cpp:644-790)
  double delay = 0.0, tol = 0.0, scale = 0.0;
  int logfreq = 0, nrepeat = 0, pad = 0;
  bool written = false;              // a snapshot has been written (the reference's !firstflag)
};

class Spparks {
 public:
  Spparks(FILE *screen, FILE *logfile, bool banner = true);
  ~Spparks();
  Spparks(const Spparks &) = delete;
  Spparks &operator=(const Spparks &) = delete;

  // executes every command of an input script (Input::file)
  void file(const char *path);
  void file(FILE *fp);
  void file(FILE *fp, bool owned);
  // executes one command line (Input::one); returns the command word
  std::string one(const std::string &line);
  // -var name value (Input's variable table, index style)
  void set_variable(const std::string &name, const std::string &value);

  // library access (src/library.cpp:95-147)
  double energy();                       // total energy of current spins
  // host arrays of the spins / strict seeds in id order, kept in step with
  // the device by library_enter / library_leave once handed out
  int32_t *site_values();
  int32_t *site_seeds();
  double **site_xyz();                   // xyz[i][0..2] of site i (App::xyz)
  // around every library call: values written through a handed-out spin
  // or seed array go to the device first (enter); after the call the
  // arrays are refreshed from the device (leave)
  void library_enter();
  void library_leave();
  int64_t nglobal() const { return nsites_; }
  double time() const { return time_; }
  int64_t naccept() const { return naccept_; }
  int64_t nattempt() const { return nattempt_; }
  int nsweeps() const { return nsweeps_; }
  int dimension() const { return dimension_; }
  double box_bound(int b) const { return b % 2 ? boxhi_[b / 2] : boxlo_[b / 2]; }
  void set_quiet(bool q) { quiet_ = q; }

 private:
  // commands
  void cmd_seed(const std::vector<std::string> &a);
  void cmd_app_style(const std::vector<std::string> &a);
  void cmd_dimension(const std::vector<std::string> &a);
  void cmd_lattice(const std::vector<std::string> &a);
  void cmd_region(const std::vector<std::string> &a);
  void cmd_boundary(const std::vector<std::string> &a);
  void cmd_create_box(const std::vector<std::string> &a);
  void cmd_create_sites(const std::vector<std::string> &a);
  void cmd_set(const std::vector<std::string> &a);
  void cmd_sweep(const std::vector<std::string> &a);
  void cmd_temperature(const std::vector<std::string> &a);
  void cmd_diag_style(const std::vector<std::string> &a);
  void cmd_stats(const std::vector<std::string> &a);
  void cmd_dump(const std::vector<std::string> &a);
  void cmd_read_sites(const std::vector<std::string> &a);
  void cmd_undump(const std::vector<std::string> &a);
  void cmd_dump_modify(const std::vector<std::string> &a);
  void cmd_clear(const std::vector<std::string> &a);
  void cmd_next(const std::vector<std::string> &a);
  void cmd_jump(const std::vector<std::string> &a);
  void cmd_run(const std::vector<std::string> &a);
  void cmd_variable(const std::vector<std::string> &a);
  void cmd_print(const std::vector<std::string> &a);
  void cmd_log(const std::vector<std::string> &a);

  // run machinery
  void init_run();
  double output_setup();
  void stats_compute(bool done, const double *energy_now);
  void dumps_compute();
  void stats_line(bool timeflag);
  void stats_header();
  void write_dump(Dump &d);
  void drain_dumps();
  double next_time(double t, int logfreq, double delta, int nrepeat, double scale, double delay) const;
  double stats_next(double t) const;
  double output_next(double stats_time) const;

  // helpers
  void message(const char *fmt, ...) __attribute__((format(printf, 2, 3)));
  void hip(int rc);
  std::string substitute(const std::string &line) const;
  void need_app(const char *cmd) const;
  void download();

  FILE *screen_, *logfile_;
  bool quiet_ = false;
  // the variable table (Variable, src/variable.cpp): index, loop and
  // equal (a number) styles; string is a one-value index
  struct Var {
    bool loop = false;
    std::vector<std::string> data;   // index values; one formatted number for equal
    int num = 0, index = 0;
    bool equal = false;
  };
  std::map<std::string, Var> vars_;
  // Input's jump / label / next state (src/input.cpp:190-205, :621-662)
  bool label_active_ = false, jump_skip_ = false;
  int file_depth_ = 0;   // input files being read (a jump needs one)
  std::string label_, jump_to_;

  std::unique_ptr<spk::MasterRng> ranmaster_;

  // app
  std::string app_;          // "potts", "ising", a proposal variant, or ""
  int nspins_ = 0;
  double dt_sweep_ = 0.0;
  int dimension_ = 3;
  std::string lattice_;
  double latconst_ = 0.0;
  std::map<std::string, Region> regions_;
  bool box_exist_ = false;
  double boxlo_[3] = {0, 0, 0}, boxhi_[3] = {0, 0, 0};
  bool periodic_[3] = {true, true, true};
  int nx_ = 0, ny_ = 0, nz_ = 0;
  int64_t nsites_ = 0;
  spk_hip_lattice *h_ = nullptr;
  bool values_set_ = false;  // every site holds a value from `set` or read_sites
  bool seeds_loaded_ = false;  // strict seeds came from a read_sites seed column

  // rejection KMC state (AppLattice)
  int sweepflag_ = 0;        // 0 none, 1 color/strict on the GPU
  bool mask_ = false;
  double temperature_ = 0.0;
  bool ranapp_made_ = false, ranstrict_made_ = false;
  double time_ = 0.0, stoptime_ = 0.0;
  int64_t naccept_ = 0, nattempt_ = 0;
  int nsweeps_ = 0;
  bool first_run_ = true;

  // output (Output, DiagEnergy, Timer)
  bool diag_energy_ = false;
  bool diag_stats_ = true;
  double energy_ = 0.0;
  double stats_delta_ = 0.0, stats_delay_ = 0.0, stats_tol_ = 0.0, stats_scale_ = 0.0, stats_time_ = 0.0;
  int stats_logfreq_ = 0, stats_nrepeat_ = 0;
  std::vector<Dump> dumps_;
  double loop_t0_ = 0.0;
  double t_solve_ = 0.0, t_output_ = 0.0, t_loop_ = 0.0;

  std::vector<int32_t> host_spin_, host_seed_;
  // arrays handed out through the library, and their contents as handed out
  bool spin_out_ = false, seed_out_ = false;
  std::vector<int32_t> spin_ref_, seed_ref_;
  std::vector<double> xyz_data_;
  std::vector<double *> xyz_rows_;
  DumpWriter writer_;   // snapshots being written; drained at the end of a run
};

}  // namespace spkfront
