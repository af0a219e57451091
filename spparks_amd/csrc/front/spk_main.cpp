// spk_main.cpp -- the spk_hip program: `spk_hip -in in.script`.
// Command-line switches follow SPPARKS (src/spparks.cpp:56-96): -in, -log,
// -screen, -var, -echo; plus -sf hip, accepted for scripts written against
// a reference build with the GPU style (this program only runs the GPU
// style, so the suffix is implied).
#include <cstdio>
#include <cstring>
#include <string>
#include <vector>

#include "spk_front.h"

int main(int argc, char **argv) {
  const char *in = nullptr, *log = "log.spparks", *screen = nullptr;
  std::vector<std::pair<std::string, std::string>> vars;
  for (int i = 1; i < argc; i++) {
    const std::string a = argv[i];
    if ((a == "-in" || a == "-i") && i + 1 < argc) in = argv[++i];
    else if ((a == "-log" || a == "-l") && i + 1 < argc) log = argv[++i];
    else if ((a == "-screen" || a == "-s") && i + 1 < argc) screen = argv[++i];
    else if ((a == "-echo" || a == "-e") && i + 1 < argc) ++i;
    else if ((a == "-sf" || a == "-suffix") && i + 1 < argc) {
      if (std::strcmp(argv[++i], "hip") != 0) {
        std::fprintf(stderr, "ERROR: only the hip suffix exists in this build\n");
        return 1;
      }
    } else if ((a == "-var" || a == "-v") && i + 2 < argc) {
      vars.emplace_back(argv[i + 1], argv[i + 2]);
      i += 2;
    } else {
      std::fprintf(stderr, "ERROR: Invalid command-line argument %s\n", a.c_str());
      return 1;
    }
  }
  FILE *scr = stdout;
  if (screen) scr = std::strcmp(screen, "none") == 0 ? nullptr : std::fopen(screen, "w");
  FILE *lg = std::strcmp(log, "none") == 0 ? nullptr : std::fopen(log, "w");
  int rc = 0;
  {
    spkfront::Spparks spk(scr, lg);
    try {
      for (auto &v : vars) spk.set_variable(v.first, v.second);
      if (in) spk.file(in);
      else spk.file(stdin);
    } catch (const spkfront::Error &e) {
      if (scr) std::fprintf(scr, "ERROR: %s\n", e.what());
      if (lg) std::fprintf(lg, "ERROR: %s\n", e.what());
      std::fprintf(stderr, "ERROR: %s\n", e.what());
      rc = 1;
    }
  }
  if (lg) std::fclose(lg);
  if (scr && scr != stdout) std::fclose(scr);
  return rc;
}
