// ORACLE -- test infrastructure.  Command-line driver of the CPU oracle:
//   refcpu_run <donnees.txt> <output_dir|none> [max_steps]
// mirrors the reference executable (main.cc:13-76) for fixture comparison.
#include <cstdio>
#include <cstdlib>

#include "refcpu.h"

int main(int argc, char** argv) {
  if (argc < 3) {
    std::fprintf(stderr, "usage: %s donnees.txt output_dir|none [max_steps]\n", argv[0]);
    return 2;
  }
  mhy_config cfg;
  int st = ref_read_donnees(argv[1], &cfg);
  if (st != MHY_OK) {
    std::fprintf(stderr, "cannot read %s (%d)\n", argv[1], st);
    return 1;
  }
  ref_ctx* r = ref_create(&cfg, &st);
  if (!r) {
    std::fprintf(stderr, "init failed (%d)\n", st);
    return 1;
  }
  st = ref_simulate(r, argv[2], argc > 3 ? std::atoi(argv[3]) : 0);
  mhy_time_state ts;
  ref_get_time_state(r, &ts);
  std::printf("status %d iterations %d t_n %.17g\n", st, ts.iteration, ts.t_n);
  ref_destroy(r);
  return st == MHY_OK ? 0 : 1;
}
