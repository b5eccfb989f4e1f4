// The mahyco executable: same command line and output as the reference's
// main (main.cc:13-80) for the VNR scheme, `mahyco donnees.txt` writing
// ./output/VnrRemap.*.vtp and ./output/VnrRemap.pvd, on one MI355X.
// An optional second argument names the output base directory ("none"
// disables output).
#include <iostream>
#include <string>

#include "mahyco.hpp"

int main(int argc, char* argv[]) {
  mhy_config cfg;
  mhy_config_default(&cfg);
  if (argc >= 2) {
    mahyco::LectureDonneesClass lecture;
    lecture.LectureDonnees(argv[1], &cfg);
  } else {
    std::cerr << "[ERREUR] Fichier de donnees non passé en argument " << std::endl;
    return 1;
  }
  if (cfg.schema != MHY_SCHEMA_VNR) {
    std::cerr << "[ERREUR] seul le schema VNR est disponible" << std::endl;
    return 1;
  }
  const std::string output = (argc >= 3) ? argv[2] : "";
  mahyco::VariablesLagRemap varlp(cfg);
  mahyco::Remap proj(&varlp);
  mahyco::Vnr c(&varlp, &proj, output);
  c.simulate();
  return 0;
}
