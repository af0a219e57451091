// spk_library.cpp -- libspparks_hip.so: the reference's library interface
// (src/library.h:31-38) over the spk_hip front end; see spparks_hip.h.
#include <cstdio>
#include <cstring>
#include <new>
#include <string>
#include <vector>

#include "spk_front.h"
#include "spparks_hip.h"

namespace {

struct Instance {
  spkfront::Spparks spk;
  std::string error, last_cmd;
  int32_t ivalue = 0;
  double dvalue = 0.0;
  std::vector<int32_t> ids;
  Instance(FILE *screen, FILE *log) : spk(screen, log) {}
};

void open_instance(int argc, char **argv, void **ptr) {
  FILE *screen = stdout;
  FILE *log = nullptr;
  const char *logname = "log.spparks";
  for (int i = 1; argv && i < argc; i++) {
    if ((!std::strcmp(argv[i], "-screen") || !std::strcmp(argv[i], "-s")) && i + 1 < argc) {
      screen = std::strcmp(argv[i + 1], "none") ? std::fopen(argv[i + 1], "w") : nullptr;
      i++;
    } else if ((!std::strcmp(argv[i], "-log") || !std::strcmp(argv[i], "-l")) && i + 1 < argc) {
      logname = argv[++i];
    }
  }
  if (std::strcmp(logname, "none")) log = std::fopen(logname, "w");
  *ptr = new (std::nothrow) Instance(screen, log);
}

// After a command or file, also one that failed part way: arrays handed
// out through spparks_extract take the device's state.
void leave(Instance *in) {
  try {
    in->spk.library_leave();
  } catch (const std::exception &e) {
    if (in->error.empty()) in->error = e.what();
  }
}

}  // namespace

extern "C" {

// the communicator slot is never read: one GPU per instance, no MPI
void spparks_open(int argc, char **argv, void *comm, void **ptr) {
  (void)comm;
  open_instance(argc, argv, ptr);
}
void spparks_open_no_mpi(int argc, char **argv, void **ptr) { open_instance(argc, argv, ptr); }

void spparks_close(void *ptr) { delete static_cast<Instance *>(ptr); }

void spparks_file(void *ptr, char *file) {
  auto *in = static_cast<Instance *>(ptr);
  try {
    in->error.clear();
    in->spk.library_enter();
    in->spk.file(file);
  } catch (const std::exception &e) {
    in->error = e.what();
  }
  leave(in);
}

char *spparks_command(void *ptr, char *str) {
  auto *in = static_cast<Instance *>(ptr);
  char *out = nullptr;
  try {
    in->error.clear();
    in->spk.library_enter();
    in->last_cmd = in->spk.one(str);
    out = const_cast<char *>(in->last_cmd.c_str());
  } catch (const std::exception &e) {
    in->error = e.what();
  }
  leave(in);
  return out;
}

void *spparks_extract(void *ptr, char *name) {
  auto *in = static_cast<Instance *>(ptr);
  try {
    in->error.clear();
    in->spk.library_enter();
    // App::extract names (src/app.cpp:163-196) plus "seeds" and "time"
    if (!std::strcmp(name, "site") || !std::strcmp(name, "iarray1")) return in->spk.site_values();
    if (!std::strcmp(name, "seeds")) return in->spk.site_seeds();
    if (!std::strcmp(name, "xyz")) return in->spk.site_xyz();
    // one integer and no double per site (ninteger = 1, ndouble = 0 in
    // src/app_potts.cpp:33-34, src/app_ising.cpp:33-34)
    if (!std::strncmp(name, "iarray", 6) || !std::strncmp(name, "darray", 6)) return nullptr;
    if (!std::strcmp(name, "nglobal") || !std::strcmp(name, "nlocal")) {
      in->ivalue = (int32_t)in->spk.nglobal();   // tagint / int in the reference
      return &in->ivalue;
    }
    if (!std::strcmp(name, "dimension")) {
      in->ivalue = in->spk.dimension();
      return &in->ivalue;
    }
    if (!std::strcmp(name, "id")) {
      in->ids.resize(in->spk.nglobal());
      for (size_t g = 0; g < in->ids.size(); g++) in->ids[g] = (int32_t)(g + 1);
      return in->ids.data();
    }
    const char *bounds[6] = {"boxxlo", "boxxhi", "boxylo", "boxyhi", "boxzlo", "boxzhi"};
    for (int b = 0; b < 6; b++)
      if (!std::strcmp(name, bounds[b])) {
        in->dvalue = in->spk.box_bound(b);
        return &in->dvalue;
      }
    if (!std::strcmp(name, "time")) {
      in->dvalue = in->spk.time();
      return &in->dvalue;
    }
  } catch (const std::exception &e) {
    in->error = e.what();
  }
  return nullptr;
}

double spparks_energy(void *ptr) {
  auto *in = static_cast<Instance *>(ptr);
  try {
    in->error.clear();
    in->spk.library_enter();
    return in->spk.energy();
  } catch (const std::exception &e) {
    in->error = e.what();
    return 0.0;
  }
}

const char *spparks_error_message(void *ptr) { return static_cast<Instance *>(ptr)->error.c_str(); }

}  // extern "C"
