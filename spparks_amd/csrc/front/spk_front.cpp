// spk_front.cpp -- command execution of the spk_hip front end (see
// spk_front.h).  Each command cites the reference routine whose behaviour it
// reproduces; everything that touches lattice data goes through the C ABI of
// libspk_hip.so.
#include "spk_front.h"

#include <chrono>
#include <cmath>
#include <cstdarg>
#include <cstdlib>
#include <cstring>
#include <new>
#include <sstream>

namespace spkfront {

namespace {

constexpr double kEpsilon = 1.0e-6;   // src/create_sites.cpp EPSILON
constexpr int kCallSweeps = 4096;     // most sweeps per device call of a run

double wall() {
  using namespace std::chrono;
  return duration<double>(steady_clock::now().time_since_epoch()).count();
}

double to_double(const std::string &s, const char *what) {
  char *end = nullptr;
  double v = std::strtod(s.c_str(), &end);
  if (end == s.c_str()) throw Error(std::string("Illegal ") + what + " command");
  return v;
}

int to_int(const std::string &s, const char *what) {
  char *end = nullptr;
  long v = std::strtol(s.c_str(), &end, 10);
  if (end == s.c_str()) throw Error(std::string("Illegal ") + what + " command");
  return (int)v;
}

int lattice_id(const std::string &s) {
  if (s == "sq/4n") return SPK_HIP_SQ_4N;
  if (s == "sq/8n") return SPK_HIP_SQ_8N;
  if (s == "sc/6n") return SPK_HIP_SC_6N;
  if (s == "sc/26n") return SPK_HIP_SC_26N;
  return -1;
}

int maxneigh(const std::string &s) {
  return s == "sq/4n" ? 4 : s == "sq/8n" ? 8 : s == "sc/6n" ? 6 : 26;
}

// splits a command line into words; text in double quotes is one word
// (Input::parse, src/input.cpp:289-355)
std::vector<std::string> tokenize(const std::string &line) {
  std::vector<std::string> out;
  size_t i = 0, n = line.size();
  while (i < n) {
    while (i < n && std::strchr(" \t\n\r\f", line[i])) i++;
    if (i >= n) break;
    if (line[i] == '"') {
      size_t j = line.find('"', i + 1);
      if (j == std::string::npos) throw Error("Unbalanced quotes in input line");
      out.push_back(line.substr(i + 1, j - i - 1));
      i = j + 1;
    } else {
      size_t j = i;
      while (j < n && !std::strchr(" \t\n\r\f", line[j])) j++;
      out.push_back(line.substr(i, j - i));
      i = j;
    }
  }
  return out;
}

}  // namespace

Spparks::Spparks(FILE *screen, FILE *logfile, bool banner) : screen_(screen), logfile_(logfile) {
  if (banner) message("SPPARKS (spk_hip: strict colour sweep on MI355X)\n");
}

// Input::clear (src/input.cpp:524-529): everything is destroyed and created
// afresh (app, lattice, dumps, stats settings, master generator) except
// what the reference's Input and universe keep: variables, screen and log.
// The object is rebuilt in place; no caller holds a member across a command.
void Spparks::cmd_clear(const std::vector<std::string> &a) {
  if (!a.empty()) throw Error("Illegal clear command");
  FILE *screen = screen_, *logfile = logfile_;
  std::map<std::string, Var> vars = std::move(vars_);
  const bool quiet = quiet_, label_active = label_active_, jump_skip = jump_skip_;
  const std::string label = label_, jump_to = jump_to_;
  const int depth = file_depth_;
  this->~Spparks();
  new (this) Spparks(screen, logfile, false);
  vars_ = std::move(vars);
  quiet_ = quiet;
  label_active_ = label_active;
  jump_skip_ = jump_skip;
  label_ = label;
  jump_to_ = jump_to;
  file_depth_ = depth;
}

Spparks::~Spparks() {
  (void)writer_.drain();
  for (auto &d : dumps_)
    if (d.fp) std::fclose(d.fp);
  if (h_) spk_hip_destroy(h_);
}

void Spparks::message(const char *fmt, ...) {
  if (quiet_) return;
  va_list ap;
  if (screen_) {
    va_start(ap, fmt);
    std::vfprintf(screen_, fmt, ap);
    va_end(ap);
    std::fflush(screen_);
  }
  if (logfile_) {
    va_start(ap, fmt);
    std::vfprintf(logfile_, fmt, ap);
    va_end(ap);
    std::fflush(logfile_);
  }
}

void Spparks::hip(int rc) {
  if (rc != SPK_HIP_OK) throw Error(spk_hip_last_error());
}

// ------------------------------------------------------------ input

void Spparks::file(const char *path) {
  FILE *fp = std::fopen(path, "r");
  if (!fp) throw Error(std::string("Cannot open input script ") + path);
  file(fp, true);
}

void Spparks::file(FILE *fp) { file(fp, false); }

// Input::file (src/input.cpp:121-200): lines ending in '&' continue.  A
// jump replaces the file being read (closing it when it was opened here).
void Spparks::file(FILE *fp, bool owned) {
  struct Closer {
    FILE *&f;
    bool &own;
    int &depth;
    ~Closer() {
      if (own && f) std::fclose(f);
      depth--;
    }
  } closer{fp, owned, file_depth_};
  file_depth_++;
  std::string line;
  char buf[4096];
  for (;;) {
    while (std::fgets(buf, sizeof(buf), fp)) {
      line += buf;
      if (!line.empty() && line.back() != '\n' && !std::feof(fp)) continue;   // long line
      size_t e = line.find_last_not_of(" \t\r\n");
      if (e != std::string::npos && line[e] == '&') {
        line.erase(e);
        line += ' ';
        continue;
      }
      one(line);
      line.clear();
      if (!jump_to_.empty()) break;
    }
    if (jump_to_.empty() && !line.empty()) one(line);
    if (jump_to_.empty()) {
      // end of file while a jump still looks for its label (src/input.cpp:155-157)
      if (label_active_) {
        label_active_ = false;
        throw Error("Label wasn't found in input script");
      }
      return;
    }
    // Input::jump (src/input.cpp:621-651): read on from the named file
    const std::string path = jump_to_;
    jump_to_.clear();
    line.clear();
    if (owned) std::fclose(fp);
    fp = std::fopen(path.c_str(), "r");
    owned = fp != nullptr;
    if (!fp) throw Error("Cannot open input script " + path);
  }
}

std::string Spparks::substitute(const std::string &in) const {
  // $x and ${name} replaced from the variable table (Input::substitute)
  std::string out;
  for (size_t i = 0; i < in.size(); i++) {
    if (in[i] != '$') {
      out += in[i];
      continue;
    }
    std::string name;
    if (i + 1 < in.size() && in[i + 1] == '{') {
      size_t j = in.find('}', i + 2);
      if (j == std::string::npos) throw Error("Invalid variable name");
      name = in.substr(i + 2, j - i - 2);
      i = j;
    } else if (i + 1 < in.size()) {
      name = in.substr(i + 1, 1);
      i += 1;
    } else {
      throw Error("Invalid variable name");
    }
    auto it = vars_.find(name);
    if (it == vars_.end()) throw Error("Substitution for illegal variable " + name);
    // Variable::retrieve (src/variable.cpp:330-360): a loop variable is its
    // 1-based counter, the others their current value
    const Var &v = it->second;
    if (v.index >= v.num) throw Error("Substitution for illegal variable " + name);
    out += v.loop ? std::to_string(v.index + 1) : v.data[v.index];
  }
  return out;
}

std::string Spparks::one(const std::string &raw) {
  // strip # comment outside quotes
  std::string line;
  bool inq = false;
  for (char ch : raw) {
    if (ch == '"') inq = !inq;
    if (ch == '#' && !inq) break;
    line += ch;
  }
  // no substitution while a jump looks for its label: variables used before
  // the label may not be defined yet (src/input.cpp:314-317)
  if (!label_active_) line = substitute(line);
  std::vector<std::string> w = tokenize(line);
  if (w.empty()) return "";
  const std::string cmd = w[0];
  std::vector<std::string> a(w.begin() + 1, w.end());
  // scanning for a jump's label: every other command is skipped
  if (label_active_ && cmd != "label") return cmd;
  if (cmd == "seed") cmd_seed(a);
  else if (cmd == "app_style") cmd_app_style(a);
  else if (cmd == "dimension") cmd_dimension(a);
  else if (cmd == "lattice") cmd_lattice(a);
  else if (cmd == "region") cmd_region(a);
  else if (cmd == "boundary") cmd_boundary(a);
  else if (cmd == "create_box") cmd_create_box(a);
  else if (cmd == "create_sites") cmd_create_sites(a);
  else if (cmd == "set") cmd_set(a);
  else if (cmd == "diag_style") cmd_diag_style(a);
  else if (cmd == "stats") cmd_stats(a);
  else if (cmd == "dump") cmd_dump(a);
  else if (cmd == "read_sites") cmd_read_sites(a);
  else if (cmd == "undump") cmd_undump(a);
  else if (cmd == "dump_modify") cmd_dump_modify(a);
  else if (cmd == "clear") cmd_clear(a);
  else if (cmd == "next") cmd_next(a);
  else if (cmd == "jump") cmd_jump(a);
  else if (cmd == "label") {
    // Input::label (src/input.cpp:655-660)
    if (a.size() != 1) throw Error("Illegal label command");
    if (label_active_ && label_ == a[0]) label_active_ = false;
  }
  else if (cmd == "run") cmd_run(a);
  else if (cmd == "variable") cmd_variable(a);
  else if (cmd == "print") cmd_print(a);
  else if (cmd == "log") cmd_log(a);
  else if (cmd == "include") {
    // Input::include (src/input.cpp:599-617): the file's commands, then back
    if (a.size() != 1) throw Error("Illegal include command");
    file(a[0].c_str());
  } else if (cmd == "reset_time") {
    // Input::reset_time (src/input.cpp:878-929), App::reset_time (src/app.cpp:152-155)
    if (app_.empty()) throw Error("Reset_time command before app_style set");
    if (a.size() < 1) throw Error("Illegal reset_time command");
    if (a[0] == "stitch") throw Error("reset_time stitch is outside this build");
    if (a.size() != 1) throw Error("Illegal reset_time command");
    const double t = std::atof(a[0].c_str());
    if (t < 0.0) throw Error("Illegal reset_time command");
    time_ = t;
  } else if (cmd == "dump_one") {
    // Input::dump_one, Output::dump_one (src/output.cpp:286-299): one
    // snapshot of a dump now, at the current time
    if (app_.empty()) throw Error("Dump_one command before app_style set");
    if (a.size() != 1) throw Error("Illegal dump_one command");
    Dump *d = nullptr;
    for (auto &x : dumps_)
      if (x.id == a[0]) d = &x;
    if (!d) throw Error("Could not find dump ID in dump_one command");
    if (!h_) throw Error("Dump_one command before sites exist");
    write_dump(*d);
    drain_dumps();
  } else if (cmd == "echo" || cmd == "processors") {
    // accepted: echo is cosmetic; one GPU has no processor grid
  } else if (cmd == "sweep" || cmd == "temperature") {
    // app-specific commands fall through to the app (src/input.cpp:512-514)
    need_app(cmd.c_str());
    if (cmd == "sweep") cmd_sweep(a);
    else cmd_temperature(a);
  } else if (cmd == "sector" || cmd == "solve_style") {
    throw Error("Command " + cmd + " is outside this build (strict colour sweep only)");
  } else {
    throw Error("Unknown command: " + line);
  }
  return cmd;
}

void Spparks::set_variable(const std::string &name, const std::string &value) {
  Var v;
  v.data = {value};
  v.num = 1;
  vars_[name] = v;
}

void Spparks::need_app(const char *cmd) const {
  if (app_.empty()) throw Error(std::string("Command ") + cmd + " before app_style set");
}

// ------------------------------------------------------------ commands

// Input::seed (src/input.cpp:942-951)
void Spparks::cmd_seed(const std::vector<std::string> &a) {
  if (a.size() != 1) throw Error("Illegal seed command");
  int s = to_int(a[0], "seed");
  if (s <= 0) throw Error("Illegal seed command");
  ranmaster_ = std::make_unique<spk::MasterRng>(s);
}

// Input::app_style (src/input.cpp:727-747), the app constructors
// (src/app_potts.cpp:30-54, src/app_ising.cpp:30-47,
// src/app_potts_neigh.cpp:29-40, src/app_potts_neighonly.cpp:29-40,
// src/app_ising_single.cpp:26-30)
void Spparks::cmd_app_style(const std::vector<std::string> &a) {
  if (a.empty()) throw Error("Illegal app_style command");
  if (a[0] == "potts" || a[0] == "potts/neigh" || a[0] == "potts/neighonly") {
    if (a.size() != 2) throw Error("Illegal app_style command");
    nspins_ = to_int(a[1], "app_style");
    if (nspins_ <= 0) throw Error("Illegal app_style command");
    // potts/neigh[only] set dt_sweep = 1/maxneigh in init_app, at each run
    dt_sweep_ = a[0] == "potts" ? 1.0 / nspins_ : 0.0;
  } else if (a[0] == "ising" || a[0] == "ising/single") {
    if (a.size() != 1) throw Error("Illegal app_style command");
    nspins_ = 2;
    dt_sweep_ = a[0] == "ising" ? 1.0 / 2.0 : 1.0;
  } else {
    throw Error("Unrecognized app style " + a[0] +
                " (this build runs potts, potts/neigh, potts/neighonly, ising and ising/single)");
  }
  app_ = a[0];
  if (h_) {
    spk_hip_destroy(h_);
    h_ = nullptr;
  }
  nsites_ = 0;
  values_set_ = false;
  seeds_loaded_ = false;
  sweepflag_ = 0;
  mask_ = false;
  temperature_ = 0.0;
  ranapp_made_ = ranstrict_made_ = false;
  time_ = 0.0;
  naccept_ = nattempt_ = 0;
  nsweeps_ = 0;
  first_run_ = true;
}

void Spparks::cmd_dimension(const std::vector<std::string> &a) {
  if (a.size() != 1) throw Error("Illegal dimension command");
  if (box_exist_) throw Error("Dimension command after simulation box is defined");
  int d = to_int(a[0], "dimension");
  if (d != 2 && d != 3) throw Error("Illegal simulation dimension (this build runs 2d and 3d lattices)");
  dimension_ = d;
}

// Lattice constructor (src/lattice.cpp:37-160)
void Spparks::cmd_lattice(const std::vector<std::string> &a) {
  if (a.empty()) throw Error("Illegal lattice command");
  if (lattice_id(a[0]) < 0)
    throw Error("Lattice style " + a[0] + " is outside this build (sq/4n, sq/8n, sc/6n, sc/26n)");
  if (a.size() != 2) throw Error("Illegal lattice command");
  const bool is3 = a[0] == "sc/6n" || a[0] == "sc/26n";
  if ((is3 && dimension_ != 3) || (!is3 && dimension_ != 2))
    throw Error("Lattice style does not match dimension");
  lattice_ = a[0];
  latconst_ = to_double(a[1], "lattice");
  if (latconst_ <= 0.0) throw Error("Illegal lattice command");
}

// RegBlock, scaled by the lattice spacing (src/region.cpp:71-78)
void Spparks::cmd_region(const std::vector<std::string> &a) {
  if (a.size() < 2) throw Error("Illegal region command");
  if (a[1] != "block") throw Error("Region style " + a[1] + " is outside this build (block)");
  if (a.size() < 8) throw Error("Illegal region command");
  if (lattice_.empty()) throw Error("Use of region with undefined lattice");
  Region r;
  for (int d = 0; d < 3; d++) {
    r.lo[d] = to_double(a[2 + 2 * d], "region") * latconst_;
    r.hi[d] = to_double(a[3 + 2 * d], "region") * latconst_;
  }
  for (size_t i = 8; i < a.size(); i += 2)
    if (a[i] != "side" || i + 1 >= a.size() || a[i + 1] != "in") throw Error("Illegal region command");
  regions_[a[0]] = r;
}

void Spparks::cmd_boundary(const std::vector<std::string> &a) {
  if (a.size() != 3) throw Error("Illegal boundary command");
  for (int d = 0; d < 3; d++) {
    if (a[d] == "p") periodic_[d] = true;
    else if (a[d] == "n") periodic_[d] = false;
    else throw Error("Illegal boundary command");
  }
}

// CreateBox (src/create_box.cpp)
void Spparks::cmd_create_box(const std::vector<std::string> &a) {
  if (a.size() != 1) throw Error("Illegal create_box command");
  if (box_exist_) throw Error("Cannot create_box after simulation box is defined");
  need_app("create_box");
  auto it = regions_.find(a[0]);
  if (it == regions_.end()) throw Error("Create_box region ID does not exist");
  for (int d = 0; d < 3; d++) {
    boxlo_[d] = it->second.lo[d];
    boxhi_[d] = it->second.hi[d];
  }
  box_exist_ = true;
  message("Created box = (%g %g %g) to (%g %g %g)\n", boxlo_[0], boxlo_[1], boxlo_[2], boxhi_[0], boxhi_[1],
          boxhi_[2]);
  message("  1 by 1 by 1 processor grid\n");
}

// CreateSites::structured_lattice for a fully periodic box
// (src/create_sites.cpp:198-409)
void Spparks::cmd_create_sites(const std::vector<std::string> &a) {
  if (a.size() != 1 || a[0] != "box")
    throw Error("create_sites style is outside this build (create_sites box)");
  if (!box_exist_) throw Error("Cannot create sites before simulation box is defined");
  if (lattice_.empty()) throw Error("Cannot create sites with undefined lattice");
  if (h_) throw Error("Cannot create sites after sites already exist");
  for (int d = 0; d < dimension_; d++)
    if (!periodic_[d]) throw Error("Non-periodic boundaries are outside this build");
  message("Creating sites ...\n");
  const double xprd = boxhi_[0] - boxlo_[0], yprd = boxhi_[1] - boxlo_[1], zprd = boxhi_[2] - boxlo_[2];
  nx_ = (int)(xprd / latconst_);
  ny_ = (int)(yprd / latconst_);
  nz_ = dimension_ == 3 ? (int)(zprd / latconst_) : 1;
  if (std::fabs(nx_ * latconst_ - xprd) > kEpsilon * latconst_ ||
      std::fabs(ny_ * latconst_ - yprd) > kEpsilon * latconst_ ||
      (dimension_ == 3 && std::fabs(nz_ * latconst_ - zprd) > kEpsilon * latconst_))
    throw Error("Periodic box is not a multiple of lattice spacing");
  if (nx_ < 1 || ny_ < 1 || nz_ < 1) throw Error("Simulation box holds no lattice sites");
  nsites_ = (int64_t)nx_ * ny_ * nz_;
  if (nsites_ > 2147483647LL) throw Error("Site IDs may exceed max ID value");
  // the device lattice needs even edges; odd ones fail at the first run,
  // where the reference checks colourability (src/lattice.cpp:199-220)
  if (nx_ % 2 || ny_ % 2 || (dimension_ == 3 && nz_ % 2)) {
    message("  %lld sites\n", (long long)nsites_);
    h_ = nullptr;
    return;
  }
  spk_hip_desc d;
  d.app = app_ == "potts"             ? SPK_HIP_POTTS
          : app_ == "potts/neigh"     ? SPK_HIP_POTTS_NEIGH
          : app_ == "potts/neighonly" ? SPK_HIP_POTTS_NEIGHONLY
          : app_ == "ising/single"    ? SPK_HIP_ISING_SINGLE
                                      : SPK_HIP_ISING;
  d.lattice = lattice_id(lattice_);
  d.nx = nx_;
  d.ny = ny_;
  d.nz = nz_;
  d.nspins = nspins_;
  d.temperature = 0.0;
  hip(spk_hip_create(&d, &h_));
  message("  %lld sites\n", (long long)nsites_);
  message("  %lld sites have %d neighbors\n", (long long)nsites_, maxneigh(lattice_));
}

// Set::command with site range / value, loop all (src/set.cpp:100-300,
// :317-345, :582-627): one master draw per set command.
void Spparks::cmd_set(const std::vector<std::string> &a) {
  if (a.size() < 2) throw Error("Illegal set command");
  if (nsites_ == 0) throw Error("Set command before sites exist");
  if (a[0] != "site") throw Error("set " + a[0] + " is outside this build (set site ...)");
  if (!ranmaster_) throw Error("Seed command has not been used");
  message("Setting site values ...\n");
  size_t iarg;
  int lo = 0, hi = 0;
  bool range;
  if (a[1] == "range") {
    if (a.size() < 4) throw Error("Illegal set command");
    lo = to_int(a[2], "set");
    hi = to_int(a[3], "set");
    range = true;
    iarg = 4;
  } else if (a[1] == "value") {
    if (a.size() < 3) throw Error("Illegal set command");
    lo = hi = to_int(a[2], "set");
    range = false;
    iarg = 3;
  } else {
    throw Error("set site " + a[1] + " is outside this build (range, value)");
  }
  for (; iarg < a.size(); iarg += 2) {
    if (a[iarg] == "loop" && iarg + 1 < a.size() && a[iarg + 1] == "all") continue;
    throw Error("set option " + a[iarg] + " is outside this build");
  }
  const int seed0 = spk::park_seed_from_draw(ranmaster_->uniform());
  if (!h_) throw Error("Color stencil is incommensurate with lattice size");
  // the reference stores any int; a value outside 1..Q fails at the run's
  // init_app check, which the device enforces here
  if (range) hip(spk_hip_set_range(h_, seed0, lo, hi));
  else hip(spk_hip_set_value(h_, lo));
  values_set_ = true;
  message("  %lld settings made for site\n", (long long)nsites_);
}

// AppLattice::set_sweep (src/app_lattice.cpp:867-889)
void Spparks::cmd_sweep(const std::vector<std::string> &a) {
  if (a.empty()) throw Error("Illegal sweep command");
  if (a[0] == "color/strict" || a[0] == "color/strict/hip") sweepflag_ = 1;
  else if (a[0] == "none") sweepflag_ = 0;
  else if (a[0] == "random" || a[0] == "raster" || a[0] == "color")
    throw Error("sweep " + a[0] + " is outside this build: it runs sweep color/strict on the GPU");
  else throw Error("Illegal sweep command");
  mask_ = false;
  for (size_t i = 1; i < a.size(); i += 2) {
    if (a[i] != "mask" || i + 1 >= a.size()) throw Error("Illegal sweep command");
    if (a[i + 1] == "yes") mask_ = true;
    else if (a[i + 1] == "no") mask_ = false;
    else throw Error("Illegal sweep command");
  }
}

// AppLattice::set_temperature (src/app_lattice.cpp:893-898)
void Spparks::cmd_temperature(const std::vector<std::string> &a) {
  if (a.size() != 1) throw Error("Illegal temperature command");
  temperature_ = to_double(a[0], "temperature");
}

// diag_style energy [stats yes|no] (src/diag_energy.cpp, src/diag.cpp)
void Spparks::cmd_diag_style(const std::vector<std::string> &a) {
  if (a.empty()) throw Error("Illegal diag_style command");
  if (a[0] != "energy") throw Error("diag_style " + a[0] + " is outside this build (energy)");
  need_app("diag_style");
  diag_energy_ = true;
  diag_stats_ = true;
  for (size_t i = 1; i < a.size(); i += 2) {
    if (i + 1 >= a.size()) throw Error("Illegal diag_style command");
    if (a[i] == "stats") diag_stats_ = a[i + 1] == "yes";
    else throw Error("diag_style option " + a[i] + " is outside this build");
  }
}

// Output::set_stats (src/output.cpp:206-246)
void Spparks::cmd_stats(const std::vector<std::string> &a) {
  if (a.empty()) throw Error("Illegal stats command");
  stats_delta_ = to_double(a[0], "stats");
  if (stats_delta_ < 0.0) throw Error("Illegal stats command");
  if (stats_delta_ == 0.0 && a.size() > 1) throw Error("Illegal stats command");
  stats_delay_ = 0.0;
  stats_logfreq_ = 0;
  for (size_t i = 1; i < a.size();) {
    if (a[i] == "delay" && i + 1 < a.size()) {
      stats_delay_ = to_double(a[i + 1], "stats");
      i += 2;
    } else if ((a[i] == "logfreq" || a[i] == "loglinfreq") && i + 2 < a.size()) {
      stats_nrepeat_ = to_int(a[i + 1], "stats");
      stats_scale_ = to_double(a[i + 2], "stats");
      if (stats_nrepeat_ < 0 || (a[i] == "logfreq" && stats_scale_ <= 0)) throw Error("Illegal stats command");
      stats_logfreq_ = stats_nrepeat_ == 0 ? 0 : (a[i] == "logfreq" ? 1 : 2);
      i += 3;
    } else if (a[i] == "tol" && i + 1 < a.size()) {
      stats_tol_ = to_double(a[i + 1], "stats");
      if (stats_tol_ < 0.0) throw Error("Illegal stats command");
      i += 2;
    } else {
      throw Error("Illegal stats command");
    }
  }
}

// dump ID text|sites delta file [fields] (src/dump.cpp:41-130,
// src/dump_text.cpp:40-130, src/dump_sites.cpp:25-48).  A '*' in the file
// name writes one file per snapshot, '*' replaced by the snapshot index.
// Field `seed` (the site's strict RNG state) is this build's addition: with
// it a `dump sites` file is a checkpoint that read_sites restores exactly.
void Spparks::cmd_dump(const std::vector<std::string> &a) {
  if (a.size() < 4) throw Error("Illegal dump command");
  if (a[1] != "text" && a[1] != "sites")
    throw Error("dump style " + a[1] + " is outside this build (text, sites)");
  if (!box_exist_) throw Error("Dump command before simulation box is defined");
  for (auto &d : dumps_)
    if (d.id == a[0]) throw Error("Reuse of dump ID");
  Dump d;
  d.id = a[0];
  d.style = a[1];
  d.delta = to_double(a[2], "dump");
  if (d.delta <= 0.0) throw Error("Illegal dump command");
  d.file = a[3];
  d.multifile = d.file.find('*') != std::string::npos;
  if (a.size() == 4) d.fields = {"id", "site", "x", "y", "z"};
  for (size_t i = 4; i < a.size(); i++) {
    const std::string &f = a[i];
    if (f != "id" && f != "site" && f != "i1" && f != "x" && f != "y" && f != "z" && f != "energy" && f != "seed")
      throw Error("Invalid attribute " + f + " in dump " + d.style + " command");
    d.fields.push_back(f);
  }
  if (d.style == "sites") {
    if (!d.multifile) throw Error("Dump sites must write one file per snapshot");
    if (d.fields[0] != "id") throw Error("Dump sites requires ID as first field");
    if (d.fields.size() < 2) throw Error("Dump sites requires two or more fields");
  }
  if (!d.multifile) {
    drain_dumps();   // the file may be one a queued snapshot still writes
    d.fp = std::fopen(d.file.c_str(), "w");
    if (!d.fp) throw Error("Cannot open dump file");
  }
  dumps_.push_back(std::move(d));
}

void Spparks::cmd_undump(const std::vector<std::string> &a) {
  if (a.size() != 1) throw Error("Illegal undump command");
  drain_dumps();
  for (size_t i = 0; i < dumps_.size(); i++)
    if (dumps_[i].id == a[0]) {
      if (dumps_[i].fp) std::fclose(dumps_[i].fp);
      dumps_.erase(dumps_.begin() + i);
      return;
    }
  throw Error("Could not find undump ID");
}

// Output::dump_modify (src/output.cpp:303-316) and Dump::modify_params
// (src/dump.cpp:644-790).  Snapshots are always in id order, so sort off and
// sort id are accepted; per-processor files and the text style's region and
// threshold filters are outside this build.
void Spparks::cmd_dump_modify(const std::vector<std::string> &a) {
  if (a.empty()) throw Error("Illegal dump_modify command");
  Dump *d = nullptr;
  for (auto &x : dumps_)
    if (x.id == a[0]) d = &x;
  if (!d) throw Error("Could not find dump ID in dump_modify command");
  if (a.size() < 2) throw Error("Illegal dump_modify command");
  auto need = [&](size_t i, size_t n) {
    if (i + n > a.size()) throw Error("Illegal dump_modify command");
  };
  for (size_t i = 1; i < a.size();) {
    const std::string &k = a[i];
    if (k == "delay") {
      need(i, 2);
      d->delay = std::atof(a[i + 1].c_str());
      i += 2;
    } else if (k == "delta") {
      need(i, 2);
      d->delta = std::atof(a[i + 1].c_str());
      if (d->delta <= 0.0) throw Error("Illegal dump_modify command");
      i += 2;
    } else if (k == "first") {
      need(i, 2);
      d->idump = std::atoi(a[i + 1].c_str());
      if (d->idump < 0) throw Error("Illegal dump_modify command");
      i += 2;
    } else if (k == "flush") {
      need(i, 2);
      if (a[i + 1] != "yes" && a[i + 1] != "no") throw Error("Illegal dump_modify command");
      i += 2;   // every snapshot is flushed when written
    } else if (k == "logfreq" || k == "loglinfreq") {
      need(i, 3);
      d->nrepeat = std::atoi(a[i + 1].c_str());
      d->scale = std::atof(a[i + 2].c_str());
      if (d->nrepeat < 0) throw Error("Illegal dump_modify command");
      d->logfreq = d->nrepeat == 0 ? 0 : k == "logfreq" ? 1 : 2;
      i += 3;
    } else if (k == "pad") {
      need(i, 2);
      d->pad = std::atoi(a[i + 1].c_str());
      if (d->pad < 0) throw Error("Illegal dump_modify command");
      i += 2;
    } else if (k == "sort") {
      need(i, 2);
      if (a[i + 1] != "off" && a[i + 1] != "id")
        throw Error("dump_modify sort by a column is outside this build (snapshots are in id order)");
      i += 2;
    } else if (k == "tol") {
      need(i, 2);
      d->tol = std::atof(a[i + 1].c_str());
      if (d->tol < 0.0) throw Error("Illegal dump_modify command");
      i += 2;
    } else if ((k == "region" || k == "thresh") && i + 1 < a.size() && a[i + 1] == "none") {
      i += 2;   // no filter: the default
    } else if (k == "fileper" || k == "nfile" || k == "region" || k == "thresh") {
      throw Error("dump_modify " + k + " is outside this build");
    } else {
      throw Error("Illegal dump_modify command");
    }
  }
}

// Variable::set (src/variable.cpp:80-160): an existing variable is kept
// (so -var on the command line wins), except that equal re-defines
void Spparks::cmd_variable(const std::vector<std::string> &a) {
  if (a.size() < 3) throw Error("Illegal variable command");
  const std::string &style = a[1];
  if (style != "index" && style != "loop" && style != "string" && style != "equal")
    throw Error("variable style " + style + " is outside this build (index, loop, equal numbers)");
  auto it = vars_.find(a[0]);
  if (it != vars_.end() && style != "equal") return;
  Var v;
  if (style == "index" || style == "string") {
    if (style == "string" && a.size() != 3) throw Error("Illegal variable command");
    v.data.assign(a.begin() + 2, a.end());
    v.num = (int)v.data.size();
  } else if (style == "loop") {
    if (a.size() != 3) throw Error("Illegal variable command");
    v.loop = true;
    v.num = std::atoi(a[2].c_str());
  } else {
    if (a.size() != 3) throw Error("Illegal variable command");
    if (it != vars_.end() && !it->second.equal) throw Error("Cannot redefine variable as a different style");
    char *end = nullptr;
    const double x = std::strtod(a[2].c_str(), &end);
    if (*end) throw Error("equal-style formulas are outside this build; give a number");
    // an equal-style variable substitutes its value as "%.10g"
    // (Variable::retrieve, src/variable.cpp:349-352)
    char buf[64];
    std::snprintf(buf, sizeof buf, "%.10g", x);
    v.data = {buf};
    v.num = 1;
    v.equal = true;
  }
  vars_[a[0]] = v;
}

// Input::next_command, Variable::next (src/variable.cpp:242-280): every
// named index or loop variable advances; one that runs out is removed and
// the next jump is skipped, which ends a loop
void Spparks::cmd_next(const std::vector<std::string> &a) {
  if (a.empty()) throw Error("Illegal next command");
  for (const auto &n : a)
    if (!vars_.count(n)) throw Error("Invalid variable in next command");
  const Var &first = vars_.at(a[0]);
  for (const auto &n : a)
    if (vars_.at(n).loop != first.loop || vars_.at(n).equal != first.equal)
      throw Error("All variables in next command must be same style");
  if (first.equal) throw Error("Invalid variable style with next command");
  bool done = false;
  for (const auto &n : a) {
    Var &v = vars_.at(n);
    if (++v.index >= v.num) {
      done = true;
      vars_.erase(n);
    }
  }
  if (done) jump_skip_ = true;
}

// Input::jump (src/input.cpp:621-651): read on from the start of the named
// file, skipping to the label when one is given
void Spparks::cmd_jump(const std::vector<std::string> &a) {
  if (a.empty() || a.size() > 2) throw Error("Illegal jump command");
  if (file_depth_ == 0) throw Error("Jump command outside an input script");
  if (jump_skip_) {
    jump_skip_ = false;
    return;
  }
  jump_to_ = a[0];
  if (a.size() == 2) {
    label_active_ = true;
    label_ = a[1];
  }
}

void Spparks::cmd_print(const std::vector<std::string> &a) {
  std::string s;
  for (size_t i = 0; i < a.size(); i++) s += (i ? " " : "") + a[i];
  message("%s\n", s.c_str());
}

// Input::log (src/input.cpp:664-686)
void Spparks::cmd_log(const std::vector<std::string> &a) {
  if (a.size() != 1) throw Error("Illegal log command");
  if (logfile_ && logfile_ != stdout && logfile_ != stderr) std::fclose(logfile_);
  logfile_ = a[0] == "none" ? nullptr : std::fopen(a[0].c_str(), "w");
  if (a[0] != "none" && !logfile_) throw Error("Cannot open logfile " + a[0]);
  if (logfile_) std::fprintf(logfile_, "SPPARKS (spk_hip) -- continued logfile from same input script\n");
}

// ------------------------------------------------------------ run

// AppLattice::init (src/app_lattice.cpp:133-328) for sweep color/strict
void Spparks::init_run() {
  if (sweepflag_ == 0) throw Error("App needs a KMC or rejection KMC solver");
  if (lattice_.empty()) throw Error("Cannot color without a lattice definition of sites");
  if (!h_) throw Error("Color stencil is incommensurate with lattice size");
  if (!ranmaster_) throw Error("Seed command has not been used");
  if (!ranapp_made_) {          // ranapp: RandomPark(uniform()) then reset(uniform(),me,100)
    ranmaster_->uniform();
    ranmaster_->uniform();
    ranapp_made_ = true;
  }
  if (!ranstrict_made_) {       // ranstrict: RandomPark(uniform()); seed = uniform()
    ranmaster_->uniform();
    const double rseed = ranmaster_->uniform();
    if (!seeds_loaded_) hip(spk_hip_init_seeds(h_, rseed));   // a checkpoint's seeds are kept
    ranstrict_made_ = true;
  }
  // init_app: site-value check; potts/neigh[only] take dt_sweep = 1/maxneigh
  // (src/app_potts_neigh.cpp:54, src/app_potts_neighonly.cpp:54)
  if (!values_set_) throw Error("One or more sites have invalid values");
  if (app_ == "potts/neigh" || app_ == "potts/neighonly") dt_sweep_ = 1.0 / maxneigh(lattice_);
  hip(spk_hip_set_temperature(h_, temperature_));
  if (dt_sweep_ == 0.0) throw Error("App did not set dt_sweep");
}

double Spparks::next_time(double tcurrent, int logfreq, double delta, int nrepeat, double scale,
                          double delay) const {
  // Output::next_time (src/output.cpp:414-437)
  double tnew = 0.0;
  if (logfreq == 0) {
    tnew = std::ceil(tcurrent / delta) * delta;
    if (tnew == tcurrent) tnew = tcurrent + delta;
  } else if (logfreq == 1) {
    while (tcurrent >= delta * scale) delta *= scale;
    double ktmp = std::pow(scale, 1.0 / nrepeat);
    tnew = delta;
    while (tcurrent >= tnew) tnew *= ktmp;
  } else {
    double start = delta;
    while (tcurrent >= start * scale) start *= scale;
    tnew = std::ceil(tcurrent / start) * start;
    if (tnew == tcurrent) tnew = tcurrent + start;
    if (static_cast<int>(tnew / start) > nrepeat) tnew = start * scale;
  }
  return tnew > delay ? tnew : delay;
}

void Spparks::stats_header() {
  char line[256];
  int n = std::snprintf(line, sizeof line, "%10s %10s %10s %10s %10s", "Time", "Naccept", "Nreject", "Nsweeps",
                        "CPU");
  if (diag_energy_ && diag_stats_) std::snprintf(line + n, sizeof line - n, " %10s", "Energy");
  message("%s\n", line);
}

// Output::stats (src/output.cpp:346-379) with AppLattice::stats
// (src/app_lattice.cpp:917-934) and DiagEnergy::stats
void Spparks::stats_line(bool timeflag) {
  char line[256];
  int n = std::snprintf(line, sizeof line, "%10g %10lld %10lld %10d %10.3g", time_, (long long)naccept_,
                        (long long)(nattempt_ - naccept_), nsweeps_, timeflag ? wall() - loop_t0_ : 0.0);
  if (diag_energy_ && diag_stats_) std::snprintf(line + n, sizeof line - n, " %10g", energy_);
  message("%s\n", line);
}

void Spparks::download() {
  host_spin_.resize(nsites_);
  host_seed_.resize(nsites_);
  hip(spk_hip_download(h_, host_spin_.data(), host_seed_.data()));
}

// Dump::write (src/dump.cpp:250-356): open the snapshot's file when there
// is one per snapshot, the style's header (DumpText::write_header,
// src/dump_text.cpp:380-392; DumpSites::write_header,
// src/dump_sites.cpp:52-66), then one line per site in id order, each value
// followed by a space (DumpText::write_text, src/dump_text.cpp:405-421).
// The state is copied from the device now; formatting and writing happen on
// the dump writer's thread while the run goes on.
void Spparks::write_dump(Dump &d) {
  DumpJob job;
  bool seeds = false;
  for (const auto &f : d.fields) {
    DumpField c = DumpField::Site;
    if (f == "id") c = DumpField::Id;
    else if (f == "seed") c = DumpField::Seed;
    else if (f == "x") c = DumpField::X;
    else if (f == "y") c = DumpField::Y;
    else if (f == "z") c = DumpField::Z;
    else if (f == "energy") c = DumpField::Energy;
    seeds |= c == DumpField::Seed;
    job.fields.push_back(c);
  }
  job.spin.resize(nsites_);
  if (seeds) job.seed.resize(nsites_);
  hip(spk_hip_download(h_, job.spin.data(), seeds ? job.seed.data() : nullptr));
  if (d.multifile) {
    const size_t star = d.file.find('*');
    char idx[32];
    std::snprintf(idx, sizeof idx, "%0*d", d.pad, d.idump);   // dump_modify pad (src/dump.cpp:377-383)
    const std::string name = d.file.substr(0, star) + idx + d.file.substr(star + 1);
    job.fp = std::fopen(name.c_str(), "w");
    if (!job.fp) throw Error("Cannot open dump file");
    job.close = true;
  } else {
    job.fp = d.fp;
  }
  std::string cols;   // "site" is written as "type" in text dumps, as is in site files
  for (size_t i = 0; i < d.fields.size(); i++)
    cols += (i ? " " : "") + ((d.fields[i] == "site" && d.style == "text") ? std::string("type") : d.fields[i]);
  auto add = [&](const char *fmt, auto... v) {
    char line[512];
    std::snprintf(line, sizeof line, fmt, v...);
    job.header += line;
  };
  if (d.style == "text") {
    add("ITEM: TIMESTEP\n%d %10g\n", d.idump, time_);
    add("ITEM: NUMBER OF ATOMS\n%lld\n", (long long)nsites_);
    add("ITEM: BOX BOUNDS\n%g %g\n%g %g\n%g %g\n", boxlo_[0], boxhi_[0], boxlo_[1], boxhi_[1], boxlo_[2],
        boxhi_[2]);
    job.header += "ITEM: ATOMS " + cols + "\n";
  } else {
    job.header += "Site file written by dump sites " + d.id + " command at time: ";
    add("%d %g\n\n", d.idump, time_);
    add("%d dimension\n", dimension_);
    add("%lld sites\n", (long long)nsites_);
    job.header += cols + " values\n";
    add("%g %g xlo xhi\n", boxlo_[0], boxhi_[0]);
    add("%g %g ylo yhi\n", boxlo_[1], boxhi_[1]);
    add("%g %g zlo zhi\n", boxlo_[2], boxhi_[2]);
    job.header += "\nValues\n\n";
  }
  job.nx = nx_;
  job.ny = ny_;
  job.nz = nz_;
  job.dimension = dimension_;
  job.nearest_only = lattice_ == "sq/4n" || lattice_ == "sc/6n";
  job.latconst = latconst_;
  for (int m = 0; m < 3; m++) {
    job.boxlo[m] = boxlo_[m];
    job.boxhi[m] = boxhi_[m];
  }
  writer_.submit(std::move(job));
  d.idump++;
  d.written = true;
}

// Waits for every snapshot handed to the dump writer.
void Spparks::drain_dumps() {
  if (!writer_.drain()) throw Error("Error writing dump file");
}

// read_sites file (src/read_sites.cpp:69-186): the free-format header
// (:199-314) checked against the existing box and sites, then the Values
// section (:507-640), one line per site: id and its values, in the columns
// of the header's `values` line (site or i1; default: the one integer per
// site).  Sites come from create_sites here, so Sites and Neighbors sections
// are outside this build.  Column `seed` is this build's addition (see
// cmd_dump): it sets the strict seeds, and the first colour/strict run then
// keeps them instead of seeding by tagreset.
void Spparks::cmd_read_sites(const std::vector<std::string> &a) {
  if (app_.empty()) throw Error("Read_sites command before app_style set");
  if (a.size() != 1) throw Error("Illegal read_sites command");
  message("Reading site file ...\n");
  drain_dumps();   // the file may be a snapshot still being written
  FILE *fp = std::fopen(a[0].c_str(), "r");
  if (!fp) throw Error("Cannot open file " + a[0]);
  struct Closer {
    FILE *f;
    ~Closer() { std::fclose(f); }
  } closer{fp};
  char line[4096];
  auto blank = [](const char *l) { return std::strspn(l, " \t\n\r") == std::strlen(l); };
  if (!std::fgets(line, sizeof line, fp)) throw Error("Unexpected end of data file");   // title line
  double blo[3] = {-0.5, -0.5, -0.5}, bhi[3] = {0.5, 0.5, 0.5};
  std::string columns;
  bool have_columns = false;
  std::string keyword;
  while (true) {
    if (!std::fgets(line, sizeof line, fp)) break;
    if (char *h = std::strchr(line, '#')) *h = '\0';
    if (blank(line)) continue;
    if (std::strstr(line, "dimension")) {
      int dim = 0;
      std::sscanf(line, "%d", &dim);
      if (box_exist_ && dim != dimension_) throw Error("Data file dimension does not match existing box");
    } else if (std::strstr(line, "sites")) {
      long long n = 0;
      std::sscanf(line, "%lld", &n);
      if (nsites_ && n != nsites_) throw Error("Data file number of sites does not match existing sites");
    } else if (std::strstr(line, "max neighbors")) {
      throw Error("read_sites Neighbors data is outside this build (sites come from create_sites)");
    } else if (char *v = std::strstr(line, "values")) {
      *v = '\0';
      columns = line;
      have_columns = true;
    } else if (std::strstr(line, "xlo xhi") || std::strstr(line, "ylo yhi") || std::strstr(line, "zlo zhi")) {
      const int dim = std::strstr(line, "xlo") ? 0 : std::strstr(line, "ylo") ? 1 : 2;
      std::sscanf(line, "%lg %lg", &blo[dim], &bhi[dim]);
      if (box_exist_ && (std::fabs(boxlo_[dim] - blo[dim]) > 1.0e-6 || std::fabs(boxhi_[dim] - bhi[dim]) > 1.0e-6))
        throw Error("Data file simulation box different that current box");
    } else {
      // ReadSites::parse_keyword (:680-720): the whole trimmed line is the
      // keyword, and the one line after it is skipped
      keyword = line;
      keyword.erase(0, keyword.find_first_not_of(" \t\n\r"));
      keyword.erase(keyword.find_last_not_of(" \t\n\r") + 1);
      if (!std::fgets(line, sizeof line, fp)) keyword.clear();
      break;
    }
  }
  if (keyword.empty()) throw Error("Site file has no Sites, Neighbors, or Values");
  if (keyword == "Sites" || keyword == "Neighbors")
    throw Error("read_sites " + keyword + " section is outside this build (sites come from create_sites)");
  if (keyword != "Values") throw Error("Unknown identifier in data file: " + keyword);
  if (nsites_ == 0) throw Error("Cannot read Values before sites exist or are read");
  if (!h_) throw Error("Color stencil is incommensurate with lattice size");

  // column kinds: 0 spin (site / i1), 1 seed
  std::vector<int> kind;
  if (!have_columns) {
    kind = {0};   // ninteger + ndouble = 1 for the potts and ising styles
  } else {
    std::istringstream ss(columns);
    std::string w;
    std::vector<std::string> words;
    while (ss >> w) words.push_back(w);
    if (words.size() < 2) throw Error("Read sites Values section has no values per site");
    if (words[0] != "id") throw Error("Read sites values header keyword must start with id");
    for (size_t i = 1; i < words.size(); i++) {
      if (words[i] == "site" || words[i] == "i1") kind.push_back(0);
      else if (words[i] == "seed") kind.push_back(1);
      else throw Error("Invalid keyword in read sites values header");
    }
  }
  download();
  bool seeds = false;
  for (int k : kind) seeds |= k == 1;
  // nsites lines of id + one word per column
  int64_t nread = 0;
  while (nread < nsites_) {
    if (!std::fgets(line, sizeof line, fp)) throw Error("Unexpected end of data file");
    std::istringstream ls(line);
    std::vector<std::string> w;
    std::string t;
    while (ls >> t) w.push_back(t);
    if (w.size() != kind.size() + 1) throw Error("Incorrect value format in data file");
    const long long id = std::atoll(w[0].c_str());
    if (id >= 1 && id <= nsites_) {
      for (size_t m = 0; m < kind.size(); m++) {
        const int v = std::atoi(w[m + 1].c_str());
        if (kind[m] == 0) host_spin_[id - 1] = v;
        else host_seed_[id - 1] = v;
      }
    }
    nread++;
  }
  int rc = spk_hip_upload(h_, host_spin_.data(), seeds ? host_seed_.data() : nullptr);
  if (rc == SPK_HIP_EVALUE && !seeds) throw Error("One or more sites have invalid values");
  hip(rc);
  values_set_ = true;
  if (seeds) seeds_loaded_ = true;
}

// Output::setup (src/output.cpp:82-137)
double Spparks::output_setup() {
  double dump_time = stoptime_;
  for (auto &d : dumps_) {
    // Output::setup (src/output.cpp:82-100): the first snapshot once the
    // delay has passed, then the schedule shifted by the dump's tolerance
    if (!d.written && time_ >= d.delay) write_dump(d);
    d.next_time = next_time(time_ + d.tol, d.logfreq, d.delta, d.nrepeat, d.scale, d.delay) - d.tol;
    dump_time = std::min(dump_time, d.next_time);
  }
  if (diag_energy_ && diag_stats_) hip(spk_hip_energy(h_, &energy_));
  if (first_run_) {
    message("Running with 32-bit site IDs\n");
    message("Memory usage per processor = %g Mbytes\n", nsites_ * 8.0 / 1048576.0);
  }
  stats_header();
  stats_line(false);
  stats_time_ = stoptime_;
  if (stats_delta_ > 0.0) stats_time_ = stats_next(time_);
  return std::min(std::min(stoptime_, dump_time), stats_time_);
}

// The stats time after a stats line at time t (src/output.cpp:172-177).
double Spparks::stats_next(double t) const {
  if (stats_delta_ == 0.0) return stoptime_;
  return next_time(t + stats_tol_, stats_logfreq_, stats_delta_, stats_nrepeat_, stats_scale_, stats_delay_) -
         stats_tol_;
}

// Output::compute's return value: the next time anything is due.  A dump's
// next_time already has its tolerance taken off once (src/output.cpp:189,
// :194-198).
double Spparks::output_next(double stats_time) const {
  double t = std::min(stoptime_, stats_time);
  for (const auto &d : dumps_) t = std::min(t, d.next_time);
  return t;
}

// Output::compute (src/output.cpp:145-202), in two halves.
// The stats half: a line when it is due.  energy_now is
// the energy after the current sweep when the caller has it from the
// per-sweep counts of a sweeps call (null: ask the device).
void Spparks::stats_compute(bool done, const double *energy_now) {
  const bool sflag = time_ >= stats_time_ || done;
  if (!sflag) return;
  if (diag_energy_ && diag_stats_) {
    if (energy_now) energy_ = *energy_now;
    else hip(spk_hip_energy(h_, &energy_));
  }
  stats_line(true);
  stats_time_ = stats_next(time_);
}

// The dump half: every dump due at the current time.
void Spparks::dumps_compute() {
  for (auto &d : dumps_) {
    if (time_ >= d.next_time - d.tol) {   // Output::compute (src/output.cpp:180-192)
      write_dump(d);
      d.next_time = next_time(time_ + d.tol, d.logfreq, d.delta, d.nrepeat, d.scale, d.delay) - d.tol;
    }
  }
}

// App::run (src/app.cpp:86-148) -> AppLattice::setup / iterate_rejection
// (src/app_lattice.cpp:332-440, :632-711) -> Finish
void Spparks::cmd_run(const std::vector<std::string> &a) {
  if (app_.empty()) throw Error("Run command before app_style set");
  if (!box_exist_) throw Error("Cannot run application until simulation box is defined");
  if (a.empty()) throw Error("Illegal run command");
  const double delta = to_double(a[0], "run");
  stoptime_ = time_ + delta;
  if (stoptime_ < time_) throw Error("Illegal run command");
  bool upto = false, pre = true, post = true;
  for (size_t i = 1; i < a.size();) {
    if (a[i] == "upto") {
      upto = true;
      i += 1;
    } else if ((a[i] == "pre" || a[i] == "post") && i + 1 < a.size()) {
      bool v;
      if (a[i + 1] == "yes") v = true;
      else if (a[i + 1] == "no") v = false;
      else throw Error("Illegal run command");
      (a[i] == "pre" ? pre : post) = v;
      i += 2;
    } else {
      throw Error("Illegal run command");
    }
  }
  // a stop time before the current time runs no sweep (src/app.cpp:126-129)
  if (upto) stoptime_ -= time_;
  if (upto && stoptime_ < 0.0) throw Error("Run upto value is before current time");

  if (pre || first_run_) init_run();
  message("Setting up run ...\n");
  t_solve_ = t_output_ = t_loop_ = 0.0;

  // setup: one colour sweep per dt_rkmc = dt_sweep, cut at the stop time
  double dt_rkmc = dt_sweep_;
  if (stoptime_ - time_ < dt_rkmc) dt_rkmc = stoptime_ - time_;
  double nextoutput = output_setup();
  first_run_ = false;

  loop_t0_ = wall();
  if (stoptime_ > time_) {
    bool done = false;
    const bool want_e = diag_energy_ && diag_stats_;
    std::vector<int64_t> acc_sw;
    std::vector<double> eng_sw;
    while (!done) {
      // The sweeps up to the next dump snapshot or the stop time (at most
      // kCallSweeps) run as one device call.  After each sweep
      // iterate_rejection calls Output::compute when the time has reached
      // nextoutput (src/app_lattice.cpp:705-708); a dump is written there
      // when time >= next_time - tol (src/output.cpp:185).  This first pass
      // replays that schedule to find the first sweep that writes a dump;
      // time advances by the same additions as one sweep at a time.
      int n = 0;
      double t = time_, st = stats_time_, no = nextoutput;
      for (;;) {
        n++;
        t += dt_rkmc;
        const bool last = t >= stoptime_;
        bool dump = false;
        if (last || t >= no) {
          if (t >= st || last) st = stats_next(t);
          for (const auto &d : dumps_) dump |= t >= d.next_time - d.tol;
          no = output_next(st);
        }
        if (last) {
          done = true;
          break;
        }
        if (dump || n == kCallSweeps) break;
      }
      acc_sw.resize(n);
      eng_sw.resize(want_e ? n : 0);
      const double t0 = wall();
      int64_t nacc = 0;
      hip(spk_hip_sweeps(h_, n, &nacc, acc_sw.data(), want_e ? eng_sw.data() : nullptr));
      t_solve_ += wall() - t0;
      const double t1 = wall();
      const int64_t acc0 = naccept_, att0 = nattempt_;
      const int sw0 = nsweeps_;
      // the same schedule with output: stats lines from the call's per-sweep
      // Naccept and energies; dumps can only be due after its last sweep
      double tk = time_;
      for (int k = 0; k < n; k++) {
        tk += dt_rkmc;
        const bool last = done && k == n - 1;
        if (!(last || tk >= nextoutput)) continue;
        time_ = tk;
        naccept_ = acc0 + acc_sw[k];
        nattempt_ = att0 + (int64_t)(k + 1) * nsites_;
        nsweeps_ = sw0 + k + 1;
        stats_compute(last, want_e ? &eng_sw[k] : nullptr);
        if (k == n - 1) dumps_compute();
        nextoutput = output_next(stats_time_);
      }
      time_ = t;
      naccept_ = acc0 + nacc;
      nattempt_ = att0 + (int64_t)n * nsites_;
      nsweeps_ = sw0 + n;
      t_output_ += wall() - t1;
    }
  }
  {
    const double t1 = wall();
    drain_dumps();
    t_output_ += wall() - t1;
  }
  t_loop_ = wall() - loop_t0_;
  // Finish (src/finish.cpp:30-120)
  message("Loop time of %g on %d procs\n", t_loop_, 1);
  if (!post) return;
  const double tl = t_loop_ == 0.0 ? 1.0 : t_loop_;
  message("\n");
  message("Solve time (%%) = %g (%g)\n", t_solve_, t_solve_ / tl * 100.0);
  message("Update time (%%) = %g (%g)\n", 0.0, 0.0);
  message("Comm  time (%%) = %g (%g)\n", 0.0, 0.0);
  message("Outpt time (%%) = %g (%g)\n", t_output_, t_output_ / tl * 100.0);
  message("App   time (%%) = %g (%g)\n", 0.0, 0.0);
  const double other = t_loop_ - t_solve_ - t_output_;
  message("Other time (%%) = %g (%g)\n", other, other / tl * 100.0);
}

// ------------------------------------------------------------ library

double Spparks::energy() {
  if (!h_) throw Error("No sites on the device");
  double e = 0.0;
  hip(spk_hip_energy(h_, &e));
  return e;
}

int32_t *Spparks::site_values() {
  if (!h_) return nullptr;
  download();
  spin_out_ = true;
  spin_ref_ = host_spin_;
  if (seed_out_) seed_ref_ = host_seed_;
  return host_spin_.data();
}

int32_t *Spparks::site_seeds() {
  if (!h_) return nullptr;
  download();
  seed_out_ = true;
  seed_ref_ = host_seed_;
  if (spin_out_) spin_ref_ = host_spin_;
  return host_seed_.data();
}

// Site coordinates as CreateSites::structured_lattice gives them
// (src/create_sites.cpp:344-351): x = i*latconst + boxxlo in a periodic
// dimension; z of a 2d lattice sits at the middle of the box (:213).
double **Spparks::site_xyz() {
  if (nsites_ == 0) return nullptr;
  xyz_data_.resize(3 * (size_t)nsites_);
  xyz_rows_.resize(nsites_);
  const double zmid = 0.5 * (boxlo_[2] + boxhi_[2]);
  for (int64_t g = 0; g < nsites_; g++) {
    const int64_t i = g % nx_, j = (g / nx_) % ny_, k = g / ((int64_t)nx_ * ny_);
    double *r = &xyz_data_[3 * (size_t)g];
    r[0] = (double)i * latconst_ + boxlo_[0];
    r[1] = (double)j * latconst_ + boxlo_[1];
    r[2] = dimension_ == 3 ? (double)k * latconst_ + boxlo_[2] : zmid;
    xyz_rows_[g] = r;
  }
  return xyz_rows_.data();
}

void Spparks::library_enter() {
  if (!spin_out_ && !seed_out_) return;
  if (!h_ || (int64_t)host_spin_.size() != nsites_) {   // the lattice was recreated
    spin_out_ = seed_out_ = false;
    return;
  }
  const bool sp = spin_out_ && host_spin_ != spin_ref_;
  const bool sd = seed_out_ && host_seed_ != seed_ref_;
  if (!sp && !sd) return;
  const int rc = spk_hip_upload(h_, sp ? host_spin_.data() : nullptr, sd ? host_seed_.data() : nullptr);
  if (rc == SPK_HIP_EVALUE && sp) throw Error("One or more sites have invalid values");
  hip(rc);
  if (sp) {
    spin_ref_ = host_spin_;
    values_set_ = true;
  }
  if (sd) {
    seed_ref_ = host_seed_;
    seeds_loaded_ = true;
  }
}

void Spparks::library_leave() {
  if (!spin_out_ && !seed_out_) return;
  if (!h_ || (int64_t)host_spin_.size() != nsites_) {
    spin_out_ = seed_out_ = false;
    return;
  }
  download();
  if (spin_out_) spin_ref_ = host_spin_;
  if (seed_out_) seed_ref_ = host_seed_;
}

}  // namespace spkfront
