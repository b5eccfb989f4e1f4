// The mahyco executable: same command line and output as the reference's
// main (main.cc:13-80), `mahyco donnees.txt` writing ./output/VnrRemap.*.vtp
// and ./output/VnrRemap.pvd for the VNR scheme, ./output/EucclhydRemap.* and
// the empty ./output/Particules.* files for the Eucclhyd scheme, on one
// MI355X.
// An optional second argument names the output base directory ("none"
// disables output).
//
// A JSON input (mhy_read_json) also reaches the keys the positional format
// has no place for.  One problem runs as above.  An ensemble ("problems")
// runs as one batch for max_time_iterations steps (every problem must give
// the same count) and prints each problem's time and conservation sums.
#include <cstdio>
#include <iostream>
#include <string>
#include <vector>

#include "mahyco.hpp"

namespace {

bool endsWith(const std::string& s, const std::string& suffix) {
  return s.size() >= suffix.size() && s.compare(s.size() - suffix.size(), suffix.size(), suffix) == 0;
}

int runBatch(const std::vector<mhy_config>& cfgs) {
  const int steps = cfgs[0].max_time_iterations;
  for (const auto& c : cfgs)
    if (c.max_time_iterations != steps || steps < 1 || steps > 10000000) {
      std::cout << "ERREUR: un ensemble demande le meme max_time_iterations (1 .. 1e7) pour chaque probleme"
                << std::endl;
      return 1;
    }
  mhy_batch* b = nullptr;
  int st = mhy_batch_create(cfgs.data(), (int)cfgs.size(), &b);
  if (st != MHY_OK) {
    std::cout << "ERREUR: creation de l'ensemble (status " << st << "): " << mhy_batch_last_error(b) << std::endl;
    mhy_batch_destroy(b);
    return 1;
  }
  st = mhy_batch_run(b, steps);
  if (st != MHY_OK) {
    std::cout << "ERREUR: calcul de l'ensemble (status " << st << "): " << mhy_batch_last_error(b) << std::endl;
    mhy_batch_destroy(b);
    return 1;
  }
  for (int i = 0; i < (int)cfgs.size(); ++i) {
    mhy_ctx* x = nullptr;
    mhy_time_state ts;
    double sums[2 + MHY_NBMATMAX] = {0};
    if (mhy_batch_context(b, i, &x) != MHY_OK || mhy_get_time_state(x, &ts) != MHY_OK ||
        mhy_conservation(x, sums) != MHY_OK) {
      std::cout << "ERREUR: lecture du probleme " << i << std::endl;
      mhy_batch_destroy(b);
      return 1;
    }
    std::printf("probleme %d %dx%d: t = %.8e masse = %.17g energie interne = %.17g\n", i, cfgs[i].nx, cfgs[i].ny,
                ts.t_n, sums[0], sums[1]);
  }
  mhy_batch_destroy(b);
  return 0;
}

}  // namespace

int main(int argc, char* argv[]) {
  if (argc < 2) {
    std::cerr << "[ERREUR] Fichier de donnees non passé en argument " << std::endl;
    return 1;
  }
  const std::string input = argv[1];
  const std::string output = (argc >= 3) ? argv[2] : "";
  mhy_config cfg;
  mhy_config_default(&cfg);
  if (endsWith(input, ".json")) {
    int n = 0;
    if (mhy_read_json(input.c_str(), nullptr, 0, &n) != MHY_OK || n < 1) {
      std::cout << "ERREUR: lecture du fichier JSON " << input << " impossible" << std::endl;
      return 1;
    }
    std::vector<mhy_config> cfgs(n);
    if (mhy_read_json(input.c_str(), cfgs.data(), n, &n) != MHY_OK) {
      std::cout << "ERREUR: lecture du fichier JSON " << input << " impossible" << std::endl;
      return 1;
    }
    if (n > 1) return runBatch(cfgs);
    cfg = cfgs[0];
  } else {
    mahyco::LectureDonneesClass lecture;
    lecture.LectureDonnees(input, &cfg);
  }
  if (cfg.schema == MHY_SCHEMA_EUCCLHYD) {
    mahyco::VariablesLagRemap varlp(cfg);
    mahyco::Eucclhyd c(&varlp, output);
    c.simulate();
    return 0;
  }
  if (cfg.schema != MHY_SCHEMA_VNR) {
    std::cerr << "[ERREUR] schema Lagrange inconnu" << std::endl;
    return 1;
  }
  mahyco::VariablesLagRemap varlp(cfg);
  mahyco::Remap proj(&varlp);
  mahyco::Vnr c(&varlp, &proj, output);
  c.simulate();
  return 0;
}
