// Positional donnees.txt reader and option defaults: the drop-in for
// LectureDonneesClass::LectureDonnees (lecture_donnees/LectureDonnees.cc:12-226)
// and the default member values of includes/ (Options.h, Eos.h, GestionTemps.h,
// Limiteurs.h, CstMesh.h).
//
// Format: every value sits on the line after a comment line.  A value is
// read with operator>>, one character is skipped, the next line is the
// comment of the following value; option blocks exist only when earlier
// values ask for them.  Unknown keywords map to 0, as unordered_map::
// operator[] does in the reference.
#include <cstring>
#include <fstream>
#include <string>
#include <unordered_map>

#include "../../../include/mahyco.h"

namespace {

int lookup(const std::unordered_map<std::string, int>& table, const std::string& key) {
  auto it = table.find(key);
  return it == table.end() ? 0 : it->second;
}

const std::unordered_map<std::string, int>& caseTable() {
  static const std::unordered_map<std::string, int> t = {
      {"UnitTestCase", MHY_CASE_UNIT_TEST},       {"SedovTestCase", MHY_CASE_SEDOV},
      {"TriplePoint", MHY_CASE_TRIPLE_POINT},     {"SodCaseX", MHY_CASE_SOD_X},
      {"SodCaseY", MHY_CASE_SOD_Y},               {"NohTestCase", MHY_CASE_NOH},
      {"AdvectionX", MHY_CASE_ADVECTION_X},       {"AdvectionY", MHY_CASE_ADVECTION_Y},
      {"AdvectionVitX", MHY_CASE_ADVECTION_VIT_X}, {"AdvectionVitY", MHY_CASE_ADVECTION_VIT_Y},
      {"BiUnitTestCase", MHY_CASE_BI_UNIT_TEST},  {"BiSedovTestCase", MHY_CASE_BI_SEDOV},
      {"BiTriplePoint", MHY_CASE_BI_TRIPLE_POINT}, {"BiShockBubble", MHY_CASE_BI_SHOCK_BUBBLE},
      {"BiSodCaseX", MHY_CASE_BI_SOD_X},          {"BiSodCaseY", MHY_CASE_BI_SOD_Y},
      {"BiNohTestCase", MHY_CASE_BI_NOH},         {"BiAdvectionX", MHY_CASE_BI_ADVECTION_X},
      {"BiAdvectionY", MHY_CASE_BI_ADVECTION_Y},  {"BiAdvectionVitX", MHY_CASE_BI_ADVECTION_VIT_X},
      {"BiAdvectionVitY", MHY_CASE_BI_ADVECTION_VIT_Y}};
  return t;
}

// Reads "comment line, value" and skips the end of the value line.
template <class T>
void field(std::ifstream& in, T& v) {
  std::string comment;
  std::getline(in, comment);
  in >> v;
  in.ignore();
}

}  // namespace

extern "C" void mhy_config_default(mhy_config* c) {
  std::memset(c, 0, sizeof(*c));
  c->test_case = -1;
  c->nx = 30, c->ny = 30;
  c->dx = 1.2 / 30, c->dy = 1.2 / 30;
  c->final_time = 1.0, c->output_time = 1.0;
  c->deltat_init = 0.0, c->deltat_min = 1.0E-10;
  c->max_time_iterations = 500000000;
  c->nbmat = -1;
  for (int k = 0; k < MHY_NBMATMAX; ++k) c->eos[k] = MHY_EOS_PERFECT_GAS, c->gammap[k] = 1.4;
  c->gamma = 1.4;
  c->schema = -1;
  c->space_order = 2;
  c->avec_projection = 1;
  c->threshold = 1.0E-16;
  c->projection_order = 2;
  c->limiter = -1, c->limiter_pure = -1;
  c->pseudo_centree = 1;
  c->abort_on_energy = 1;
}

extern "C" int mhy_read_donnees(const char* path, mhy_config* c) {
  static const std::unordered_map<std::string, int> schemas = {
      {"Eucclhyd", MHY_SCHEMA_EUCCLHYD}, {"VNR", MHY_SCHEMA_VNR}, {"CSTS", 2002}, {"MYR", 2003}};
  static const std::unordered_map<std::string, int> limiters = {
      {"minmod", MHY_LIM_MINMOD},           {"superBee", MHY_LIM_SUPERBEE},
      {"vanLeer", MHY_LIM_VANLEER},         {"minmodG", MHY_LIM_MINMODG},
      {"superBeeG", MHY_LIM_SUPERBEEG},     {"vanLeerG", MHY_LIM_VANLEERG},
      {"arithmeticG", MHY_LIM_ARITHMETICG}, {"ultrabeeG", MHY_LIM_ULTRABEEG}};
  static const std::unordered_map<std::string, int> yesno = {{"non", 0}, {"oui", 1}};
  static const std::unordered_map<std::string, int> dual = {{"A1", 0}, {"A2", 1}, {"PB", 2}};
  static const std::unordered_map<std::string, int> eos = {
      {"Void", MHY_EOS_VOID},           {"PerfectGas", MHY_EOS_PERFECT_GAS}, {"StiffenedGas", MHY_EOS_STIFFENED_GAS},
      {"Murnhagan", MHY_EOS_MURNHAGAN}, {"SolidLinear", MHY_EOS_SOLID_LINEAR}};
  if (!c) return MHY_ERR_ARG;
  mhy_config_default(c);
  std::ifstream in(path ? path : "");
  if (!in) return MHY_ERR_IO;
  std::string word;
  field(in, word);
  c->test_case = lookup(caseTable(), word);
  // number of materials from the case id (LectureDonnees.cc:34-38)
  c->nbmat = (c->test_case < MHY_CASE_BI_UNIT_TEST) ? 1 : 2;
  if (c->test_case == MHY_CASE_BI_TRIPLE_POINT) c->nbmat = 3;
  switch (c->test_case) {
    case MHY_CASE_ADVECTION_X: case MHY_CASE_ADVECTION_Y: case MHY_CASE_BI_ADVECTION_X:
    case MHY_CASE_BI_ADVECTION_Y: case MHY_CASE_ADVECTION_VIT_X: case MHY_CASE_ADVECTION_VIT_Y:
    case MHY_CASE_BI_ADVECTION_VIT_X: case MHY_CASE_BI_ADVECTION_VIT_Y:
      c->sans_lagrange = 1;
      break;
    default:
      break;
  }
  field(in, c->nx);
  field(in, c->ny);
  field(in, c->dx);
  field(in, c->dy);
  field(in, c->final_time);
  field(in, c->output_time);
  field(in, c->deltat_init);
  field(in, c->deltat_min);
  for (int k = 0; k < c->nbmat; ++k) {
    field(in, word);
    c->eos[k] = lookup(eos, word);
  }
  field(in, word);
  c->schema = lookup(schemas, word);
  field(in, c->space_order);
  field(in, word);
  c->avec_projection = lookup(yesno, word);
  field(in, word);
  c->projection_conservative = lookup(yesno, word);
  field(in, c->threshold);
  field(in, c->projection_order);
  {
    std::string comment;
    std::getline(in, comment);
    in >> word;
    c->limiter = lookup(limiters, word);
    if (c->limiter == -1 || c->limiter == 0) return MHY_ERR_IO;  // "Limiteur non prévu"
    in.ignore();
  }
  field(in, word);
  c->pp = lookup(yesno, word);
  if (c->pp == 1) {
    field(in, word);
    c->pp_mixte = lookup(yesno, word);
    {
      std::string comment;
      std::getline(in, comment);
      in >> word;
      c->limiter_pure = lookup(limiters, word);
      if (c->limiter == -1 && c->pp_mixte == 1) return MHY_ERR_IO;
      in.ignore();
    }
    field(in, word);
    c->pp_complet = lookup(yesno, word);
  }
  if (c->schema == MHY_SCHEMA_VNR) {
    field(in, word);
    c->pseudo_centree = lookup(yesno, word);
    if (c->avec_projection == 1) {
      field(in, word);
      c->methode_flux_masse = lookup(dual, word);
    }
  }
  return MHY_OK;
}
