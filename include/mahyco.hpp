/* C++ drivers of the VNR Lagrange + remap path over the C ABI of mahyco.h.
 *
 * They keep the reference's object graph and entry-point names so a caller
 * of the reference's drivers can switch with the same call sequence
 * (main.cc:15-62):
 *
 *   mhy_config cfg;                                   // the option structs
 *   mahyco::LectureDonneesClass().LectureDonnees(file, &cfg);
 *   mahyco::VariablesLagRemap varlp(cfg);             // shared state (device)
 *   mahyco::Remap proj(&varlp);                       // remap/Remap.h:102-184
 *   mahyco::Vnr c(&varlp, &proj, output);             // vnr_remap/Vnr.h:184-373
 *   c.simulate();
 *
 * Error behaviour follows the reference: fatal conditions (unreadable input,
 * NaN after the final remap, |e| > 5 when that abort is on, device errors)
 * print a message and exit(1).  Header-only; link libmahyco_hip.so. */
#ifndef MAHYCO_HPP
#define MAHYCO_HPP

#include <cstdio>
#include <cstdlib>
#include <filesystem>
#include <iostream>
#include <string>

#include "mahyco.h"

namespace mahyco {

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
};

/* The state shared by the Lagrange and remap drivers
 * (includes/VariablesLagRemap.h:30-50 and the Vnr member Views): one device
 * context built from the options, with the mesh, the initial state of
 * Init.cc and setUpTimeLoopN already applied. */
class VariablesLagRemap {
 public:
  explicit VariablesLagRemap(const mhy_config& cfg) : cfg_(cfg) {
    const int st = mhy_create(&cfg_, &ctx_);
    check(ctx_, st, "creation du contexte");
  }
  ~VariablesLagRemap() {
    if (ctx_) mhy_destroy(ctx_);
  }
  VariablesLagRemap(const VariablesLagRemap&) = delete;
  VariablesLagRemap& operator=(const VariablesLagRemap&) = delete;

  mhy_ctx* ctx() const { return ctx_; }
  const mhy_config& config() const { return cfg_; }

 private:
  mhy_config cfg_;
  mhy_ctx* ctx_ = nullptr;
};

/* remap/Remap.h:161-172: the public entry points of the directional remap.
 * The four cell stages of a sweep run as one kernel, launched by
 * computeUremap{1,2}; computeGradPhiFace{1,2}, computeGradPhi{1,2} and
 * computeUpwindFaceQuantitiesForProjection{1,2} are kept so the reference's
 * call sequence compiles unchanged and launch nothing.  FacesOfNode builds no
 * table: node/face adjacency is closed-form index arithmetic. */
class Remap {
 public:
  explicit Remap(VariablesLagRemap* varlp) : varlp_(varlp) {}

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

}  // namespace mahyco

#endif /* MAHYCO_HPP */
