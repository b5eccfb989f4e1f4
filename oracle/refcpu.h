/*
 * ORACLE -- test infrastructure, not product code.
 *
 * C API of the CPU restatement of MaHyCo's VNR and Eucclhyd Lagrange +
 * directional remap time loops (oracle/refcpu.cc).  Only tests/, __graft_entry__.smoke() and
 * bench.py's cpu_baseline leg load it, as the checker.
 */
#ifndef MAHYCO_ORACLE_REFCPU_H
#define MAHYCO_ORACLE_REFCPU_H

#include "../include/mahyco.h"

#ifdef __cplusplus
extern "C" {
#endif

typedef struct ref_ctx ref_ctx;

/* Positional donnees.txt reader (LectureDonnees.cc:12-226). */
int ref_read_donnees(const char* path, mhy_config* cfg);
void ref_config_default(mhy_config* cfg);

/* Mesh + initial state + setUpTimeLoopN (VnrRemap.cc:330-341; Eucclhyd:
 * EucclhydRemap.cc:273-286), by cfg->schema. */
ref_ctx* ref_create(const mhy_config* cfg, int* status);
void ref_destroy(ref_ctx* r);

/* One loop iteration (VnrRemap.cc:130-222, EucclhydRemap.cc:84-168) with
 * the n/n+1 swap, no output.
 * Returns MHY_OK or an error status. */
int ref_step(ref_ctx* r);
/* Whole reference loop with VTP/PVD output into output_dir (NULL/"none":
 * disabled).  Returns MHY_OK or an error status. */
int ref_simulate(ref_ctx* r, const char* output_dir, int max_steps);

/* Single-formula probes (tests only): the limited gradient of one component
 * (remap/UtilesRemap-Impl.h:99-126), INT2Y (remap/UtilesRemap.cc:177-188),
 * and the PenteBorne swept flux of one component (UtilesRemap-Impl.h:142-316;
 * type 0: back face, 1: front face) under limiter lim. */
double ref_probe_gradient(ref_ctx* r, int lim, double gp, double gm, double phi, double phip, double phim, double h0,
                          double hp, double hm);
double ref_probe_int2y(double X, double x0, double y0, double x1, double y1);
double ref_probe_flux_pp(ref_ctx* r, int lim, double grad, double phi, double phip, double phim, double h0, double hp,
                         double hm, double fnv, double dt, int type);

/* The cell stage of computeGradPhi (remap/PhaseRemap1.cc:134-320) on one
 * cell: arrays of the 13 vector components; gp / gm as the reference passes
 * them to computeAndLimitGradPhi; mix3 / pure3 = mixte / pure flags of the
 * cell, its plus and its minus neighbour; Fm / Fp the oriented normal
 * velocities of the back / front faces, fdual the PB dual velocity.
 * Outputs the limited gradient, the back (Ar) and front (Av) PenteBorne
 * fluxes and the PB dual flux.  Returns the component count (13). */
int ref_probe_cell_pp(ref_ctx* r, const double* gp, const double* gm, const double* phi, const double* phip,
                      const double* phim, const int* mix3, const int* pure3, double h0, double hp, double hm,
                      double Fm, double Fp, double fdual, double dt, double* grad, double* ar, double* av,
                      double* dual);

/* The third-order face value of one component (remap/UtilesRemap.cc:197-320):
 * y, h = means and widths of the cells c-3 .. c+2 around the face, vdt = face
 * normal velocity times dt. */
double ref_probe_flux_o3(ref_ctx* r, const double* y, const double* h, double vdt);

int ref_get_field(ref_ctx* r, int field, double* out, long long n);
int ref_get_time_state(ref_ctx* r, mhy_time_state* ts);
int ref_set_threads(int nthreads);

#ifdef __cplusplus
}
#endif
#endif
