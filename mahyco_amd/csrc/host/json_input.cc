// JSON front-end of the input: the keys of the positional donnees.txt
// (lecture_donnees/LectureDonnees.cc:12-226) plus the keys the reference has
// no input for (per-material gamma, the initialisers of the extension test
// cases and their parameters, seed, precision form, |e| > 5 abort), and
// ensembles of problems.
//
//   { "donnees": "SOD/donnees.txt",       optional base file, relative to this one
//     "test_case": "RiemannPerturbed",    or a reference case name, or its id
//     "nx": 2048, "ny": 2048, "dx": 4.8828125e-4, "dy": 4.8828125e-4,
//     "gammap": [1.4, 1.67], "limiter": "superBee", "pp": true,
//     "precision": "mixed", "seed": 20261202, "param": [0.5, 0.01, 8],
//     "problems": [ {...}, {...} ] }    optional: one problem per entry, each
//                                        starting from the keys around it
//
// Unknown keys and malformed documents are errors (MHY_ERR_IO), so that a
// misspelt key cannot silently fall back to a default.
#include <cctype>
#include <cmath>
#include <cstdlib>
#include <cstring>
#include <fstream>
#include <map>
#include <memory>
#include <sstream>
#include <string>
#include <vector>

#include "../../../include/mahyco.h"

namespace {

struct JVal {
  enum Kind { NUL, BOOL, NUM, STR, ARR, OBJ } kind = NUL;
  bool b = false;
  double num = 0.0;
  std::string str;  // also the literal text of a number (integers beyond 2^53)
  std::vector<JVal> arr;
  std::vector<std::pair<std::string, JVal>> obj;  // document order
};

class Parser {
 public:
  explicit Parser(const std::string& t) : t_(t) {}
  bool parse(JVal& v) {
    if (!value(v)) return false;
    ws();
    return i_ == t_.size();
  }

 private:
  void ws() {
    while (i_ < t_.size() && std::isspace((unsigned char)t_[i_])) ++i_;
  }
  bool lit(const char* w) {
    const size_t n = std::strlen(w);
    if (t_.compare(i_, n, w) != 0) return false;
    i_ += n;
    return true;
  }
  bool string(std::string& s) {
    if (i_ >= t_.size() || t_[i_] != '"') return false;
    ++i_;
    s.clear();
    while (i_ < t_.size() && t_[i_] != '"') {
      char ch = t_[i_++];
      if (ch == '\\') {
        if (i_ >= t_.size()) return false;
        const char e = t_[i_++];
        switch (e) {
          case '"': case '\\': case '/': ch = e; break;
          case 'n': ch = '\n'; break;
          case 't': ch = '\t'; break;
          case 'r': ch = '\r'; break;
          case 'b': ch = '\b'; break;
          case 'f': ch = '\f'; break;
          default: return false;  // \u escapes are not needed by any key or value
        }
      }
      s.push_back(ch);
    }
    if (i_ >= t_.size()) return false;
    ++i_;
    return true;
  }
  bool value(JVal& v) {
    ws();
    if (i_ >= t_.size()) return false;
    const char ch = t_[i_];
    if (ch == '{') {
      ++i_;
      v.kind = JVal::OBJ;
      ws();
      if (i_ < t_.size() && t_[i_] == '}') return ++i_, true;
      while (true) {
        ws();
        std::string k;
        if (!string(k)) return false;
        ws();
        if (i_ >= t_.size() || t_[i_++] != ':') return false;
        JVal e;
        if (!value(e)) return false;
        v.obj.emplace_back(k, std::move(e));
        ws();
        if (i_ < t_.size() && t_[i_] == ',') { ++i_; continue; }
        if (i_ < t_.size() && t_[i_] == '}') return ++i_, true;
        return false;
      }
    }
    if (ch == '[') {
      ++i_;
      v.kind = JVal::ARR;
      ws();
      if (i_ < t_.size() && t_[i_] == ']') return ++i_, true;
      while (true) {
        JVal e;
        if (!value(e)) return false;
        v.arr.push_back(std::move(e));
        ws();
        if (i_ < t_.size() && t_[i_] == ',') { ++i_; continue; }
        if (i_ < t_.size() && t_[i_] == ']') return ++i_, true;
        return false;
      }
    }
    if (ch == '"') return v.kind = JVal::STR, string(v.str);
    if (lit("true")) return v.kind = JVal::BOOL, v.b = true, true;
    if (lit("false")) return v.kind = JVal::BOOL, v.b = false, true;
    if (lit("null")) return v.kind = JVal::NUL, true;
    const size_t s = i_;
    while (i_ < t_.size() && (std::isdigit((unsigned char)t_[i_]) || std::strchr("+-.eE", t_[i_]))) ++i_;
    if (i_ == s) return false;
    v.kind = JVal::NUM;
    v.str = t_.substr(s, i_ - s);
    char* end = nullptr;
    v.num = std::strtod(v.str.c_str(), &end);
    return end && *end == '\0';
  }
  const std::string& t_;
  size_t i_ = 0;
};

const std::map<std::string, int>& caseNames() {
  static const std::map<std::string, int> t = {
      {"UnitTestCase", MHY_CASE_UNIT_TEST},        {"SedovTestCase", MHY_CASE_SEDOV},
      {"TriplePoint", MHY_CASE_TRIPLE_POINT},      {"SodCaseX", MHY_CASE_SOD_X},
      {"SodCaseY", MHY_CASE_SOD_Y},                {"NohTestCase", MHY_CASE_NOH},
      {"AdvectionX", MHY_CASE_ADVECTION_X},        {"AdvectionY", MHY_CASE_ADVECTION_Y},
      {"AdvectionVitX", MHY_CASE_ADVECTION_VIT_X}, {"AdvectionVitY", MHY_CASE_ADVECTION_VIT_Y},
      {"BiUnitTestCase", MHY_CASE_BI_UNIT_TEST},   {"BiSedovTestCase", MHY_CASE_BI_SEDOV},
      {"BiTriplePoint", MHY_CASE_BI_TRIPLE_POINT}, {"BiShockBubble", MHY_CASE_BI_SHOCK_BUBBLE},
      {"BiSodCaseX", MHY_CASE_BI_SOD_X},           {"BiSodCaseY", MHY_CASE_BI_SOD_Y},
      {"BiNohTestCase", MHY_CASE_BI_NOH},          {"BiAdvectionX", MHY_CASE_BI_ADVECTION_X},
      {"BiAdvectionY", MHY_CASE_BI_ADVECTION_Y},   {"BiAdvectionVitX", MHY_CASE_BI_ADVECTION_VIT_X},
      {"BiAdvectionVitY", MHY_CASE_BI_ADVECTION_VIT_Y},
      {"RiemannPerturbed", MHY_CASE_RIEMANN_PERTURBED},
      {"PointExplosion", MHY_CASE_POINT_EXPLOSION},
      {"TripleLayout", MHY_CASE_TRIPLE_LAYOUT},
      {"Riemann", MHY_CASE_RIEMANN}};
  return t;
}
const std::map<std::string, int>& limiterNames() {
  static const std::map<std::string, int> t = {
      {"minmod", MHY_LIM_MINMOD},   {"superBee", MHY_LIM_SUPERBEE},   {"vanLeer", MHY_LIM_VANLEER},
      {"minmodG", MHY_LIM_MINMODG}, {"superBeeG", MHY_LIM_SUPERBEEG}, {"vanLeerG", MHY_LIM_VANLEERG},
      {"arithmeticG", MHY_LIM_ARITHMETICG}, {"ultrabeeG", MHY_LIM_ULTRABEEG}};
  return t;
}
const std::map<std::string, int>& eosNames() {
  static const std::map<std::string, int> t = {
      {"Void", MHY_EOS_VOID}, {"PerfectGas", MHY_EOS_PERFECT_GAS}, {"StiffenedGas", MHY_EOS_STIFFENED_GAS},
      {"Murnhagan", MHY_EOS_MURNHAGAN}, {"SolidLinear", MHY_EOS_SOLID_LINEAR}};
  return t;
}
const std::map<std::string, int>& schemaNames() {
  static const std::map<std::string, int> t = {{"Eucclhyd", MHY_SCHEMA_EUCCLHYD}, {"VNR", MHY_SCHEMA_VNR}};
  return t;
}
const std::map<std::string, int>& dualNames() {
  static const std::map<std::string, int> t = {{"A1", 0}, {"A2", 1}, {"PB", 2}};
  return t;
}
const std::map<std::string, int>& precisionNames() {
  static const std::map<std::string, int> t = {{"fp64", 0}, {"mixed", 1}};
  return t;
}

// Materials implied by a test case (LectureDonnees.cc:34-38 for the
// reference's cases; fixed by each extension initialiser).
int materialsOf(int tc) {
  switch (tc) {
    case MHY_CASE_RIEMANN_PERTURBED: case MHY_CASE_RIEMANN: return 2;
    case MHY_CASE_POINT_EXPLOSION: return 1;
    case MHY_CASE_TRIPLE_LAYOUT: return 3;
    case MHY_CASE_BI_TRIPLE_POINT: return 3;
    default: return tc < MHY_CASE_BI_UNIT_TEST ? 1 : 2;
  }
}
bool advectionCase(int tc) {
  switch (tc) {
    case MHY_CASE_ADVECTION_X: case MHY_CASE_ADVECTION_Y: case MHY_CASE_BI_ADVECTION_X:
    case MHY_CASE_BI_ADVECTION_Y: case MHY_CASE_ADVECTION_VIT_X: case MHY_CASE_ADVECTION_VIT_Y:
    case MHY_CASE_BI_ADVECTION_VIT_X: case MHY_CASE_BI_ADVECTION_VIT_Y:
      return true;
    default:
      return false;
  }
}

struct Reader {
  std::string dir;  // directory of the JSON file (relative "donnees" paths)
  std::string err;

  bool fail(const std::string& m) {
    if (err.empty()) err = m;
    return false;
  }
  bool asInt(const JVal& v, const std::string& key, int& out) {
    if (v.kind != JVal::NUM || v.num != std::floor(v.num)) return fail(key + ": integer expected");
    out = (int)v.num;
    return true;
  }
  bool asNum(const JVal& v, const std::string& key, double& out) {
    if (v.kind != JVal::NUM) return fail(key + ": number expected");
    out = v.num;
    return true;
  }
  // a name from table, or the numeric id itself
  bool asName(const JVal& v, const std::string& key, const std::map<std::string, int>& table, int& out) {
    if (v.kind == JVal::NUM) return asInt(v, key, out);
    if (v.kind != JVal::STR) return fail(key + ": name expected");
    auto it = table.find(v.str);
    if (it == table.end()) return fail(key + ": unknown name " + v.str);
    out = it->second;
    return true;
  }
  // true / false, "oui" / "non", 1 / 0
  bool asFlag(const JVal& v, const std::string& key, int& out) {
    if (v.kind == JVal::BOOL) return out = v.b ? 1 : 0, true;
    if (v.kind == JVal::STR && (v.str == "oui" || v.str == "non")) return out = (v.str == "oui") ? 1 : 0, true;
    if (v.kind == JVal::NUM && (v.num == 0.0 || v.num == 1.0)) return out = (int)v.num, true;
    return fail(key + ": true/false, \"oui\"/\"non\" or 0/1 expected");
  }
  template <class T, class F>
  bool asArray(const JVal& v, const std::string& key, T* out, int cap, F&& one) {
    if (v.kind != JVal::ARR) return fail(key + ": array expected");
    if ((int)v.arr.size() > cap) return fail(key + ": at most " + std::to_string(cap) + " entries");
    for (size_t k = 0; k < v.arr.size(); ++k)
      if (!one(v.arr[k], key + "[" + std::to_string(k) + "]", out[k])) return false;
    return true;
  }

  // Applies the keys of object o (except "problems") to c.
  bool apply(const JVal& o, mhy_config& c) {
    if (o.kind != JVal::OBJ) return fail("object expected");
    // the base file first, whatever its position in the object
    for (const auto& kv : o.obj)
      if (kv.first == "donnees") {
        if (kv.second.kind != JVal::STR) return fail("donnees: path expected");
        std::string p = kv.second.str;
        if (!p.empty() && p[0] != '/' && !dir.empty()) p = dir + "/" + p;
        const int st = mhy_read_donnees(p.c_str(), &c);
        if (st != MHY_OK) return fail("donnees: cannot read " + p);
      }
    bool nbmatGiven = false, caseGiven = false;
    for (const auto& kv : o.obj) {
      const std::string& k = kv.first;
      const JVal& v = kv.second;
      bool ok = true;
      if (k == "donnees" || k == "problems") continue;
      else if (k == "test_case") ok = asName(v, k, caseNames(), c.test_case), caseGiven = true;
      else if (k == "nx") ok = asInt(v, k, c.nx);
      else if (k == "ny") ok = asInt(v, k, c.ny);
      else if (k == "dx") ok = asNum(v, k, c.dx);
      else if (k == "dy") ok = asNum(v, k, c.dy);
      else if (k == "final_time") ok = asNum(v, k, c.final_time);
      else if (k == "output_time") ok = asNum(v, k, c.output_time);
      else if (k == "deltat_init") ok = asNum(v, k, c.deltat_init);
      else if (k == "deltat_min") ok = asNum(v, k, c.deltat_min);
      else if (k == "max_time_iterations") ok = asInt(v, k, c.max_time_iterations);
      else if (k == "nbmat") ok = asInt(v, k, c.nbmat), nbmatGiven = true;
      else if (k == "eos")
        ok = asArray(v, k, c.eos, MHY_NBMATMAX,
                     [&](const JVal& e, const std::string& kk, int& out) { return asName(e, kk, eosNames(), out); });
      else if (k == "gammap")
        ok = asArray(v, k, c.gammap, MHY_NBMATMAX,
                     [&](const JVal& e, const std::string& kk, double& out) { return asNum(e, kk, out); });
      else if (k == "gamma") ok = asNum(v, k, c.gamma);
      else if (k == "schema") ok = asName(v, k, schemaNames(), c.schema);
      else if (k == "space_order") ok = asInt(v, k, c.space_order);
      else if (k == "avec_projection") ok = asFlag(v, k, c.avec_projection);
      else if (k == "projection_conservative") ok = asFlag(v, k, c.projection_conservative);
      else if (k == "threshold") ok = asNum(v, k, c.threshold);
      else if (k == "projection_order") ok = asInt(v, k, c.projection_order);
      else if (k == "limiter") ok = asName(v, k, limiterNames(), c.limiter);
      else if (k == "pp") ok = asFlag(v, k, c.pp);
      else if (k == "pp_mixte") ok = asFlag(v, k, c.pp_mixte);
      else if (k == "limiter_pure") ok = asName(v, k, limiterNames(), c.limiter_pure);
      else if (k == "pp_complet") ok = asFlag(v, k, c.pp_complet);
      else if (k == "pseudo_centree") ok = asFlag(v, k, c.pseudo_centree);
      else if (k == "methode_flux_masse") ok = asName(v, k, dualNames(), c.methode_flux_masse);
      else if (k == "sans_lagrange") ok = asFlag(v, k, c.sans_lagrange);
      else if (k == "abort_on_energy") ok = asFlag(v, k, c.abort_on_energy);
      else if (k == "precision") ok = asName(v, k, precisionNames(), c.precision);
      else if (k == "seed") {
        if (v.kind != JVal::NUM || v.str.find_first_of(".eE-") != std::string::npos)
          ok = fail("seed: non-negative integer expected");
        else
          c.seed = std::strtoull(v.str.c_str(), nullptr, 10);
      } else if (k == "param") {
        std::memset(c.param, 0, sizeof(c.param));
        ok = asArray(v, k, c.param, 8, [&](const JVal& e, const std::string& kk, double& out) {
          return asNum(e, kk, out);
        });
      } else {
        ok = fail("unknown key " + k);
      }
      if (!ok) return false;
    }
    if (caseGiven) {
      // what the positional reader derives from the case (LectureDonnees.cc:34-46)
      if (!nbmatGiven) c.nbmat = materialsOf(c.test_case);
      c.sans_lagrange = advectionCase(c.test_case) ? 1 : c.sans_lagrange;
    }
    return true;
  }
};

// A configuration for keys that give no base file: the reference's defaults
// with the VNR scheme.
void jsonDefault(mhy_config* c) {
  mhy_config_default(c);
  c->schema = MHY_SCHEMA_VNR;
  c->limiter = MHY_LIM_MINMOD;
}

}  // namespace

extern "C" int mhy_read_json(const char* path, mhy_config* cfgs, int capacity, int* count) {
  if (!path || !count || capacity < 0 || (capacity > 0 && !cfgs)) return MHY_ERR_ARG;
  *count = 0;
  std::ifstream in(path);
  if (!in) return MHY_ERR_IO;
  std::stringstream ss;
  ss << in.rdbuf();
  const std::string text = ss.str();
  JVal doc;
  if (!Parser(text).parse(doc) || doc.kind != JVal::OBJ) return MHY_ERR_IO;
  Reader rd;
  const std::string p(path);
  const size_t slash = p.find_last_of('/');
  rd.dir = (slash == std::string::npos) ? std::string() : p.substr(0, slash);
  mhy_config base;
  jsonDefault(&base);
  if (!rd.apply(doc, base)) return MHY_ERR_IO;
  const JVal* problems = nullptr;
  for (const auto& kv : doc.obj)
    if (kv.first == "problems") problems = &kv.second;
  if (!problems) {
    *count = 1;
    if (capacity < 1) return capacity == 0 ? MHY_OK : MHY_ERR_ARG;
    cfgs[0] = base;
    return MHY_OK;
  }
  if (problems->kind != JVal::ARR || problems->arr.empty()) return MHY_ERR_IO;
  const int n = (int)problems->arr.size();
  *count = n;
  if (capacity == 0) return MHY_OK;  // size query
  if (capacity < n) return MHY_ERR_ARG;
  for (int i = 0; i < n; ++i) {
    mhy_config c = base;
    if (!rd.apply(problems->arr[i], c)) return MHY_ERR_IO;
    cfgs[i] = c;
  }
  return MHY_OK;
}
