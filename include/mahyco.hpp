/* C++ drivers of the VNR Lagrange + remap path over the C ABI of mahyco.h.
 *
 * They keep the reference's object graph, class names and constructor
 * signatures, so that the reference's own main (main.cc:13-76) compiles
 * against this header with only its include lines and namespace names
 * changed (#include "mahyco.hpp"; `using namespace nablalib` becomes
 * `using namespace mahyco`; Kokkos::initialize / finalize become
 * mahyco::initialize / finalize):
 *
 *   LectureDonneesClass lecture;                                        // lecture_donnees/LectureDonnees.h:11-20
 *   lecture.LectureDonnees(argv[1], scheme, o, cstmesh, gt, lim, eos, test);
 *   auto nm = CartesianMesh2DGenerator::generate(nx, ny, dx, dy);
 *   auto varlp = new variableslagremaplib::VariablesLagRemap(nm);      // includes/VariablesLagRemap.h:23-84
 *   auto part = new particleslib::SchemaParticules(nm, cstmesh, gt, test);
 *   auto proj = new Remap(o, cstmesh, gt, cl, lim, nm, varlp);          // remap/Remap.h:102-157
 *   auto c = new Vnr(o, cstmesh, gt, test, cl, lim, part, eos, nm, varlp, proj, output);  // vnr_remap/Vnr.h:184-196
 *   c->simulate();
 *
 * The option structs mirror the headers of includes/ (same names, members,
 * defaults).
 * The device context is created by the Vnr constructor, the first object
 * that sees every option; Remap reaches it through the shared
 * VariablesLagRemap.  A shorter form builds the context from one mhy_config:
 *
 *   mhy_config cfg;
 *   mahyco::LectureDonneesClass().LectureDonnees(file, &cfg);
 *   mahyco::VariablesLagRemap varlp(cfg);
 *   mahyco::Remap proj(&varlp);
 *   mahyco::Vnr c(&varlp, &proj, output);
 *   c.simulate();
 *
 * Error behaviour follows the reference: fatal conditions (unreadable input,
 * NaN after the final remap, |e| > 5 when that abort is on, device errors,
 * an unsupported scheme) print a message and exit(1).  Header-only; link
 * libmahyco_hip.so. */
#ifndef MAHYCO_HPP
#define MAHYCO_HPP

#include <array>
#include <cstdio>
#include <cstdlib>
#include <filesystem>
#include <iostream>
#include <string>

#include "mahyco.h"

namespace mahyco {

/* The reference's callers name std::string unqualified (its mesh headers
 * open namespace std); `using namespace mahyco` brings it in. */
using std::string;

/* Runtime lifecycle (the reference's Kokkos::initialize / finalize,
 * main.cc:15,73).  The HIP runtime initialises itself on the first context
 * (device from HIP_VISIBLE_DEVICES, one stream per context), so neither has
 * work to do; they exist so the caller's sequence stays as it is. */
inline void initialize(int& /*argc*/, char* /*argv*/[]) {}
inline void finalize() {}

/* ---- option structs: includes/SchemaLagrange.h, Options.h, ConditionsLimites.h,
 * Limiteurs.h, Eos.h, CasTest.h, CstMesh.h, GestionTemps.h ---- */
namespace schemalagrangelib {
class SchemaLagrangeClass {
 public:
  struct SchemaLagrange {
    int Eucclhyd = MHY_SCHEMA_EUCCLHYD;
    int VNR = MHY_SCHEMA_VNR;
    int CSTS = 2002;
    int MYR = 2003;
    int schema = -1;
  };
  SchemaLagrange* schemas = nullptr;
};
}  // namespace schemalagrangelib

namespace optionschemalib {
class OptionsSchema {
 public:
  struct Options {
    int nbmat = -1;
    double threshold = 1.0E-16;
    int spaceOrder = 2;
    int projectionOrder = 2;
    int projectionConservative = 0;
    int AvecProjection = 1;
    int AvecParticules = 0;
    int Adiabatique = 1;
    int Isotherme = 2;
    int AvecEquilibrage = -1;
    int pseudo_centree = 1;
    int methode_flux_masse = 0;
    int sansLagrange = 0;
  };
  Options* options = nullptr;
};
}  // namespace optionschemalib

namespace conditionslimiteslib {
/* Boundary conditions: the Vnr constructor derives them from the test case
 * (vnr_remap/Init.cc:9-95) inside the context, as the reference's
 * initBoundaryConditions does; the struct keeps the reference's members for
 * callers that read them. */
class ConditionsLimites {
 public:
  struct Cdl {
    std::array<double, 2> ex = {{1.0, 0.0}};
    std::array<double, 2> ey = {{0.0, 1.0}};
    std::array<double, 3 * MHY_NBMATMAX + 4> Uzero = {};
    int symmetry = MHY_BC_SYMMETRY;
    int imposedVelocity = MHY_BC_IMPOSED_VELOCITY;
    int freeSurface = MHY_BC_FREE_SURFACE;
    int periodic = MHY_BC_PERIODIC;
    int leftFluxBC = 0;
    std::array<double, 3 * MHY_NBMATMAX + 4> leftFluxBCValue = Uzero;
    int rightFluxBC = 0;
    std::array<double, 3 * MHY_NBMATMAX + 4> rightFluxBCValue = Uzero;
    int bottomFluxBC = 0;
    std::array<double, 3 * MHY_NBMATMAX + 4> bottomFluxBCValue = Uzero;
    int topFluxBC = 0;
    std::array<double, 3 * MHY_NBMATMAX + 4> topFluxBCValue = Uzero;
    int FluxBC = leftFluxBC + rightFluxBC + bottomFluxBC + topFluxBC;
    int leftBC = 0;
    std::array<double, 2> leftBCValue = ey;
    int rightBC = 0;
    std::array<double, 2> rightBCValue = ey;
    int topBC = 0;
    std::array<double, 2> topBCValue = ex;
    int bottomBC = 0;
    std::array<double, 2> bottomBCValue = ex;
    int leftCellBC = 0;
    int rightCellBC = 0;
    int bottomCellBC = 0;
    int topCellBC = 0;
  };
  Cdl* cdl = nullptr;
};
}  // namespace conditionslimiteslib

namespace limiteurslib {
class LimiteursClass {
 public:
  struct Limiteurs {
    int minmod = MHY_LIM_MINMOD;
    int superBee = MHY_LIM_SUPERBEE;
    int vanLeer = MHY_LIM_VANLEER;
    int minmodG = MHY_LIM_MINMODG;
    int superBeeG = MHY_LIM_SUPERBEEG;
    int vanLeerG = MHY_LIM_VANLEERG;
    int arithmeticG = MHY_LIM_ARITHMETICG;
    int ultrabeeG = MHY_LIM_ULTRABEEG;
    int projectionAvecPlateauPente = 0;
    int projectionLimiterId = -1;
    int projectionLimiterIdPure = -1;
    int projectionLimiteurMixte = 0;
    int projectionPlateauPenteComplet = 0;
  };
  Limiteurs* limiteurs = nullptr;
};
}  // namespace limiteurslib

namespace eoslib {
class EquationDetat {
 public:
  struct Eos {
    int Void = MHY_EOS_VOID;
    int PerfectGas = MHY_EOS_PERFECT_GAS;
    int StiffenedGas = MHY_EOS_STIFFENED_GAS;
    int Murnhagan = MHY_EOS_MURNHAGAN;
    int SolidLinear = MHY_EOS_SOLID_LINEAR;
    std::array<int, MHY_NBMATMAX> Nom = {{PerfectGas, PerfectGas, PerfectGas}};
    std::array<double, MHY_NBMATMAX> gammap = {{1.4, 1.4, 1.4}};
    double gamma = 1.4;
  };
  Eos* eos = nullptr;
};
}  // namespace eoslib

namespace castestlib {
class CasTest {
 public:
  struct Test {
    int UnitTestCase = 0;
    int SedovTestCase = 1;
    int TriplePoint = 2;
    int SodCaseX = 4;
    int SodCaseY = 5;
    int NohTestCase = 6;
    int AdvectionX = 7;
    int AdvectionY = 8;
    int AdvectionVitX = 9;
    int AdvectionVitY = 10;
    int BiUnitTestCase = 11;
    int BiSedovTestCase = 12;
    int BiTriplePoint = 13;
    int BiShockBubble = 14;
    int BiSodCaseX = 15;
    int BiSodCaseY = 16;
    int BiNohTestCase = 17;
    int BiAdvectionX = 18;
    int BiAdvectionY = 19;
    int BiAdvectionVitX = 20;
    int BiAdvectionVitY = 21;
    int Nom = -1;
  };
  Test* test = nullptr;
};
}  // namespace castestlib

namespace cstmeshlib {
class ConstantesMaillagesClass {
 public:
  struct ConstantesMaillages {
    double X_LENGTH = 1.2;
    double Y_LENGTH = X_LENGTH;
    int X_EDGE_ELEMS = 30;
    int Y_EDGE_ELEMS = 30;
    double X_EDGE_LENGTH = X_LENGTH / X_EDGE_ELEMS;
    double Y_EDGE_LENGTH = Y_LENGTH / Y_EDGE_ELEMS;
  };
  ConstantesMaillages* cstmesh = nullptr;
};
}  // namespace cstmeshlib

namespace gesttempslib {
class GestionTempsClass {
 public:
  struct GestTemps {
    double final_time = 1.0;
    double output_time = final_time;
    double cfl = 0.45;
    int max_time_iterations = 500000000;
    double deltat_init = 0.;
    double deltat_min = 1.0E-10;
    double deltat_n = 0.;
    double deltat_nplus1 = 0.;
    double t_n = 0.;
    double t_nplus1 = 0.;
  };
  GestTemps* gt = nullptr;
};
}  // namespace gesttempslib

using SchemaLagrangeT = schemalagrangelib::SchemaLagrangeClass::SchemaLagrange;
using OptionsT = optionschemalib::OptionsSchema::Options;
using CdlT = conditionslimiteslib::ConditionsLimites::Cdl;
using LimiteursT = limiteurslib::LimiteursClass::Limiteurs;
using EosT = eoslib::EquationDetat::Eos;
using TestT = castestlib::CasTest::Test;
using CstMeshT = cstmeshlib::ConstantesMaillagesClass::ConstantesMaillages;
using GestTempsT = gesttempslib::GestionTempsClass::GestTemps;

/* The option structs from one mhy_config: what LectureDonnees fills. */
inline void configToOptions(const mhy_config& c, SchemaLagrangeT* s, OptionsT* o, CstMeshT* cm, GestTempsT* gt,
                            LimiteursT* l, EosT* e, TestT* t) {
  t->Nom = c.test_case;
  o->nbmat = c.nbmat;
  o->sansLagrange = c.sans_lagrange;
  cm->X_EDGE_ELEMS = c.nx, cm->Y_EDGE_ELEMS = c.ny;
  cm->X_EDGE_LENGTH = c.dx, cm->Y_EDGE_LENGTH = c.dy;
  gt->final_time = c.final_time, gt->output_time = c.output_time;
  gt->deltat_init = c.deltat_init, gt->deltat_min = c.deltat_min;
  gt->max_time_iterations = c.max_time_iterations;
  for (int k = 0; k < MHY_NBMATMAX; ++k) e->Nom[k] = c.eos[k], e->gammap[k] = c.gammap[k];
  e->gamma = c.gamma;
  s->schema = c.schema;
  o->spaceOrder = c.space_order;
  o->AvecProjection = c.avec_projection;
  o->projectionConservative = c.projection_conservative;
  o->threshold = c.threshold;
  o->projectionOrder = c.projection_order;
  l->projectionLimiterId = c.limiter;
  l->projectionAvecPlateauPente = c.pp;
  l->projectionLimiteurMixte = c.pp_mixte;
  l->projectionLimiterIdPure = c.limiter_pure;
  l->projectionPlateauPenteComplet = c.pp_complet;
  o->pseudo_centree = c.pseudo_centree;
  o->methode_flux_masse = c.methode_flux_masse;
}

/* One mhy_config from the option structs (the inverse of configToOptions);
 * keys the structs do not carry keep mhy_config_default's values. */
inline mhy_config optionsToConfig(int schema, const OptionsT* o, const CstMeshT* cm, const GestTempsT* gt,
                                  const LimiteursT* l, const EosT* e, const TestT* t) {
  mhy_config c;
  mhy_config_default(&c);
  c.test_case = t->Nom;
  c.nbmat = o->nbmat;
  c.sans_lagrange = o->sansLagrange;
  c.nx = cm->X_EDGE_ELEMS, c.ny = cm->Y_EDGE_ELEMS;
  c.dx = cm->X_EDGE_LENGTH, c.dy = cm->Y_EDGE_LENGTH;
  c.final_time = gt->final_time, c.output_time = gt->output_time;
  c.deltat_init = gt->deltat_init, c.deltat_min = gt->deltat_min;
  c.max_time_iterations = gt->max_time_iterations;
  for (int k = 0; k < MHY_NBMATMAX; ++k) c.eos[k] = e->Nom[k], c.gammap[k] = e->gammap[k];
  c.gamma = e->gamma;
  c.schema = schema;
  c.space_order = o->spaceOrder;
  c.avec_projection = o->AvecProjection;
  c.projection_conservative = o->projectionConservative;
  c.threshold = o->threshold;
  c.projection_order = o->projectionOrder;
  c.limiter = l->projectionLimiterId;
  c.pp = l->projectionAvecPlateauPente;
  c.pp_mixte = l->projectionLimiteurMixte;
  c.limiter_pure = l->projectionLimiterIdPure;
  c.pp_complet = l->projectionPlateauPenteComplet;
  c.pseudo_centree = o->pseudo_centree;
  c.methode_flux_masse = o->methode_flux_masse;
  return c;
}

/* cppnabla/mesh/CartesianMesh2D.h:22-198 and CartesianMesh2DGenerator.cc:26-101.
 * The mesh is a structured nx x ny grid whose topology the kernels compute
 * in closed form; this object only carries its sizes to the drivers. */
class CartesianMesh2D {
 public:
  CartesianMesh2D(int nx, int ny, double dx, double dy) : nx_(nx), ny_(ny), dx_(dx), dy_(dy) {}
  int getNbNodes() const { return (nx_ + 1) * (ny_ + 1); }
  int getNbCells() const { return nx_ * ny_; }
  int getNbFaces() const { return 2 * nx_ * ny_ + nx_ + ny_; }
  int getNbXQuads() const { return nx_; }
  int getNbYQuads() const { return ny_; }
  double getXEdgeLength() const { return dx_; }
  double getYEdgeLength() const { return dy_; }

 private:
  int nx_, ny_;
  double dx_, dy_;
};

class CartesianMesh2DGenerator {
 public:
  static CartesianMesh2D* generate(size_t nbXQuads, size_t nbYQuads, double xSize, double ySize) {
    return new CartesianMesh2D((int)nbXQuads, (int)nbYQuads, xSize, ySize);
  }
};

/* Prints the context's last error and exits(1), as the reference does on
 * fatal conditions (e.g. vnr_remap/PhaseRemapFinal.cc:236-246). */
inline void check(mhy_ctx* ctx, int status, const char* what) {
  if (status == MHY_OK) return;
  const char* msg = ctx ? mhy_last_error(ctx) : "";
  std::cout << "ERREUR: " << what << " (status " << status << ")" << (msg && *msg ? ": " : "") << (msg ? msg : "")
            << std::endl;
  std::exit(1);
}

/* lecture_donnees/LectureDonnees.h:13-20: fills the options from a
 * positional donnees.txt.  The seven option structs of the reference are
 * the fields of one mhy_config. */
class LectureDonneesClass {
 public:
  void LectureDonnees(const std::string& fichier, mhy_config* cfg) {
    const int st = mhy_read_donnees(fichier.c_str(), cfg);
    if (st != MHY_OK) {
      std::cout << "ERREUR: lecture du fichier de donnees " << fichier << " impossible (status " << st << ")"
                << std::endl;
      std::exit(1);
    }
  }
  /* The reference's signature (lecture_donnees/LectureDonnees.h:13-20). */
  void LectureDonnees(string Fichier, SchemaLagrangeT* s, OptionsT* o, CstMeshT* cstmesh, GestTempsT* gt,
                      LimiteursT* l, EosT* eos, TestT* test) {
    mhy_config c;
    LectureDonnees(Fichier, &c);
    configToOptions(c, s, o, cstmesh, gt, l, eos, test);
  }
};

/* The state shared by the Lagrange and remap drivers
 * (includes/VariablesLagRemap.h:30-50 and the Vnr member Views): one device
 * context built from the options, with the mesh, the initial state of
 * Init.cc and setUpTimeLoopN already applied. */
class VariablesLagRemap {
 public:
  explicit VariablesLagRemap(const mhy_config& cfg) { create(cfg); }
  /* The reference's signature (includes/VariablesLagRemap.h:52-84): only the
   * mesh is known here; the Lagrange driver's constructor creates the
   * context once it has every option (create()). */
  explicit VariablesLagRemap(CartesianMesh2D* mesh) : mesh_(mesh) { mhy_config_default(&cfg_); }
  ~VariablesLagRemap() {
    if (ctx_) mhy_destroy(ctx_);
  }
  VariablesLagRemap(const VariablesLagRemap&) = delete;
  VariablesLagRemap& operator=(const VariablesLagRemap&) = delete;

  void create(const mhy_config& cfg) {
    if (ctx_) mhy_destroy(ctx_);
    cfg_ = cfg;
    ctx_ = nullptr;
    const int st = mhy_create(&cfg_, &ctx_);
    check(ctx_, st, "creation du contexte");
  }
  mhy_ctx* ctx() const {
    if (!ctx_) check(nullptr, MHY_ERR_ARG, "contexte non cree (le schema Lagrange ne l'a pas construit)");
    return ctx_;
  }
  const mhy_config& config() const { return cfg_; }
  CartesianMesh2D* mesh() const { return mesh_; }

 private:
  mhy_config cfg_;
  mhy_ctx* ctx_ = nullptr;
  CartesianMesh2D* mesh_ = nullptr;
};

namespace variableslagremaplib {
using mahyco::VariablesLagRemap;
}

/* particle_scheme/SchemaParticules.h:16-122.  Particles are coupled to the
 * Eucclhyd scheme only (eucclhyd_remap/EucclhydRemap.cc:119-126); the VNR
 * driver stores the pointer and never calls it, so the object only keeps
 * its constructor arguments. */
namespace particleslib {
class SchemaParticules {
 public:
  SchemaParticules(CartesianMesh2D* mesh, CstMeshT* cstmesh, GestTempsT* gt, TestT* test)
      : mesh(mesh), cstmesh(cstmesh), gt(gt), test(test) {}
  CartesianMesh2D* mesh;
  CstMeshT* cstmesh;
  GestTempsT* gt;
  TestT* test;
};
}  // namespace particleslib

/* remap/Remap.h:161-172: the public entry points of the directional remap.
 * The four cell stages of a sweep run as one kernel, launched by
 * computeUremap{1,2}; computeGradPhiFace{1,2}, computeGradPhi{1,2} and
 * computeUpwindFaceQuantitiesForProjection{1,2} are kept so the reference's
 * call sequence compiles unchanged and launch nothing.  FacesOfNode builds no
 * table: node/face adjacency is closed-form index arithmetic. */
class Remap {
 public:
  explicit Remap(VariablesLagRemap* varlp) : varlp_(varlp) {}
  /* The reference's signature (remap/Remap.h:102-157).  The options reach
   * the kernels through the context the Lagrange driver creates in varlp. */
  Remap(OptionsT*, CstMeshT*, GestTempsT*, CdlT*, LimiteursT*, CartesianMesh2D*, VariablesLagRemap* varlp)
      : varlp_(varlp) {}

  void computeGradPhiFace1() noexcept {}
  void computeGradPhi1() noexcept {}
  void computeUpwindFaceQuantitiesForProjection1() noexcept {}
  void computeUremap1() noexcept { stage(1, MHY_STAGE_CELLS); }
  void computeDualUremap1() noexcept { stage(1, MHY_STAGE_NODES); }
  void computeGradPhiFace2() noexcept {}
  void computeGradPhi2() noexcept {}
  void computeUpwindFaceQuantitiesForProjection2() noexcept {}
  void computeUremap2() noexcept { stage(2, MHY_STAGE_CELLS); }
  void computeDualUremap2() noexcept { stage(2, MHY_STAGE_NODES); }
  void FacesOfNode() noexcept {}

 private:
  void stage(int sweep, int which) {
    check(varlp_->ctx(), mhy_remap_stage(varlp_->ctx(), sweep, which), "projection");
  }
  VariablesLagRemap* varlp_;
};

/* vnr_remap/Vnr.h:184-373: the VNR driver.  simulate() runs the reference
 * time loop (vnr_remap/VnrRemap.cc:127-253) phase by phase through the C
 * ABI, with its output (dumpVariables, VnrRemap.cc:255-299) into
 * <output>/output ("" gives ./output, "none" disables it, as
 * cppnabla/mesh/FileWriter.cc:20-35). */
class Vnr {
 public:
  Vnr(VariablesLagRemap* varlp, Remap* remap, const std::string& output, bool verbose = true)
      : varlp_(varlp), remap_(remap), verbose_(verbose) {
    setOutput(output);
  }
  /* The reference's signature (vnr_remap/Vnr.h:184-196): builds the device
   * context (mesh, Init.cc initial state, setUpTimeLoopN) from the options
   * into the shared varlp.  The mesh sizes come from the ConstantesMaillages
   * the mesh was generated from. */
  Vnr(OptionsT* o, CstMeshT* cstmesh, GestTempsT* gt, TestT* test, CdlT* /*cdl*/, LimiteursT* lim,
      particleslib::SchemaParticules* /*part*/, EosT* eos, CartesianMesh2D* /*mesh*/, VariablesLagRemap* varlp,
      Remap* remap, string output)
      : varlp_(varlp), remap_(remap), verbose_(true) {
    varlp_->create(optionsToConfig(MHY_SCHEMA_VNR, o, cstmesh, gt, lim, eos, test));
    setOutput(output);
  }

  void simulate() {
    mhy_ctx* x = varlp_->ctx();
    check(x, mhy_output_open(x, outdir_.c_str()), "sortie");
    if (verbose_) std::cout << "Debut de la boucle en temps" << std::endl;
    executeTimeLoopN();
    if (verbose_) std::cout << "Fin de la boucle en temps" << std::endl;
  }

 private:
  void executeTimeLoopN() {
    mhy_ctx* x = varlp_->ctx();
    const mhy_config& c = varlp_->config();
    const bool projection = c.avec_projection == 1;
    int n = 0;
    bool continueLoop = true;
    do {
      n++;
      check(x, mhy_compute_dt(x), "pas de temps");  // computeDeltaT + computeTime
      dumpVariables();
      check(x, mhy_lagrange(x), "phase de Lagrange");  // updateVelocity .. computePressionMoyenne
      if (projection) {
        check(x, mhy_prepare_remap(x), "preparation de la projection");
        remap_->computeGradPhiFace1();
        remap_->computeGradPhi1();
        remap_->computeUpwindFaceQuantitiesForProjection1();
        remap_->computeUremap1();
        remap_->computeDualUremap1();
        remap_->computeGradPhiFace2();
        remap_->computeGradPhi2();
        remap_->computeUpwindFaceQuantitiesForProjection2();
        remap_->computeUremap2();
        remap_->computeDualUremap2();
        check(x, mhy_remap_final(x), "projection finale");  // remapVariables + node mass + EOS
      }
      mhy_time_state ts;
      check(x, mhy_get_time_state(x, &ts), "etat en temps");
      check(x, ts.status, "calcul");
      continueLoop = (n + 1 < c.max_time_iterations && ts.t_nplus1 < c.final_time);
      if (verbose_) std::printf("[%5d] t = %.8e DT = %.8e\n", n, ts.t_n, ts.deltat_nplus1);
      if (continueLoop) check(x, mhy_end_iteration(x), "fin d'iteration");
    } while (continueLoop);
    dumpVariables();  // a last output at the end
  }

  void setOutput(const std::string& output) {
    if (output == "none" || output == "NONE") {
      outdir_ = "none";
    } else {
      // a previous output directory is replaced (FileWriter.cc:29-33)
      outdir_ = output.empty() ? std::string("output") : output + "/output";
      std::error_code ec;
      std::filesystem::remove_all(outdir_, ec);
      std::filesystem::create_directory(outdir_, ec);
    }
  }

  void dumpVariables() {
    mhy_ctx* x = varlp_->ctx();
    nbCalls_++;
    if (outdir_ == "none") return;
    mhy_time_state ts;
    check(x, mhy_get_time_state(x, &ts), "etat en temps");
    if (ts.t_n >= lastDump_ + varlp_->config().output_time || ts.t_n == 0.) {
      check(x, mhy_output_write(x, nbCalls_, ts.t_n), "ecriture des sorties");
      lastDump_ = ts.t_n;
    }
  }

  VariablesLagRemap* varlp_;
  Remap* remap_;
  bool verbose_;
  std::string outdir_;
  int nbCalls_ = 0;
  double lastDump_ = 0.0;
};

/* eucclhyd_remap/Eucclhyd.h:40-371: the Eucclhyd driver.  The
 * reference's signature builds the device context (mesh, Init.cc initial
 * state, setUpTimeLoopN) for this scheme; simulate() runs the time loop
 * (eucclhyd_remap/EucclhydRemap.cc:84-290) with its output (EucclhydRemap
 * and the empty Particules files) into <output>/output. */
class Eucclhyd {
 public:
  Eucclhyd(OptionsT* o, CstMeshT* cstmesh, GestTempsT* gt, TestT* test, CdlT* /*cdl*/, LimiteursT* lim,
           particleslib::SchemaParticules* /*part*/, EosT* eos, CartesianMesh2D* /*mesh*/,
           VariablesLagRemap* varlp, Remap* /*remap*/, string output)
      : varlp_(varlp), output_(output) {
    varlp_->create(optionsToConfig(MHY_SCHEMA_EUCCLHYD, o, cstmesh, gt, lim, eos, test));
  }
  /* Over a context already created for the Eucclhyd scheme. */
  Eucclhyd(VariablesLagRemap* varlp, const std::string& output) : varlp_(varlp), output_(output) {}
  void simulate() {
    mhy_ctx* x = varlp_->ctx();
    const std::string dir = (output_ == "none" || output_ == "NONE") ? std::string("none")
                            : output_.empty()                        ? std::string("output")
                                                                     : output_ + "/output";
    check(x, mhy_simulate(x, dir.c_str(), 0), "calcul");
  }

 private:
  VariablesLagRemap* varlp_;
  std::string output_;
};

}  // namespace mahyco

#endif /* MAHYCO_HPP */
