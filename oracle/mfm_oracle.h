/*
 * mfm_oracle.h -- CPU FP64 restatement of the reference patch update.
 *
 * TEST INFRASTRUCTURE ONLY.  Only tests/, __graft_entry__.smoke() and the
 * cpu_baseline leg of bench.py may load this library, and only as the
 * checker (or as the timed CPU baseline).  The product path never calls it.
 *
 * Every function here restates a reference routine in the reference's own
 * operation order (built with -ffp-contract=off, no fast-math); the .cpp
 * cites the reference file:line for each.
 */
#ifndef MFM_ORACLE_H
#define MFM_ORACLE_H
#include "../include/mfm.h"

#ifdef __cplusplus
extern "C" {
#endif

/* PatchMHDAM::updateState on host FABs.  Same contract as
 * mfm_patch_update_host (include/mfm.h).  nthreads: OpenMP threads (<=0: 1). */
int oracle_patch_update(const mfm_params* p, double ch, double cp,
                        mfm_fab U, mfm_box valid, double dt, double dx,
                        double time, mfm_fab* F, mfm_fab* divB,
                        double* new_dt, int min_cell[3], int nthreads);
/* PatchMHDMF::computeDt */
int oracle_compute_dt(const mfm_params* p, mfm_fab U, mfm_box box, double dx,
                      double* dt, int min_cell[3]);
/* One EulerStep (src/PatchMHDAM.cpp:495-709), for stage-level tests. */
int oracle_euler_step(const mfm_params* p, double ch, double cp, mfm_fab U,
                      mfm_box box, double dt, double dx, double time,
                      mfm_fab* F, mfm_fab* divB, int nthreads);
/* Initial data: HELIOINIT / ORSZAGINIT / RIEMANNINIT over box. */
int oracle_init_problem(const mfm_params* p, mfm_fab U, mfm_box box, double dx);
/* Periodic ghost fill for a one-box periodic level. */
int oracle_fill_periodic(const mfm_params* p, mfm_fab U, mfm_box valid);

/* ---- level ghost fill (mfm_oracle_fill.cpp) ---------------------------- */
/* LevelData::exchange of n boxes with ghost depth ng; U[i] holds box i grown
 * by ng.  Domain and periodicity from p. */
int oracle_level_exchange(const mfm_params* p, const mfm_box* boxes, int n,
                          mfm_fab* U, int ng);
/* PiecewiseLinearFillPatchMHDAM::define + fillInterp: fills the ghost cells
 * of U[i] (boxes[i] grown by ng) from the coarse level (cboxes; cold / cnew
 * either may be NULL) at time coefficient alpha.  p is the fine level. */
int oracle_fill_interp(const mfm_params* p, const mfm_box* boxes, int n, mfm_fab* U,
                       const mfm_box* cboxes, int ncb, const mfm_fab* cold,
                       const mfm_fab* cnew, int r, int ng, double dxc, double alpha);
/* interp_divB_0 on one slope triple (dBx/dx, dBy/dy, dBz/dz), in place */
void oracle_interp_divb(double s[3], int dims);
/* Coarse-fine synchronisation.  R: dense register over the coarse problem
 * domain, R[(((d*2 + h) * ncell) + cell) * nc + comp] for the low (h = 0) or
 * high (h = 1) face of a coarse cell in direction d; F: per box, per
 * direction face FABs (box-major). */
int oracle_fr_increment_coarse(const mfm_params* p, const mfm_box* fine, int nf,
                               const mfm_box* crse, int ncb, int r, const mfm_fab* F,
                               double scale, double* R);
int oracle_fr_increment_fine(const mfm_params* p, const mfm_box* fine, int nf,
                             const mfm_box* crse, int ncb, int r, const mfm_fab* F,
                             double scale, double* R);
int oracle_fr_reflux(const mfm_params* p, const mfm_box* fine, int nf, const mfm_box* crse,
                     int ncb, int r, mfm_fab* U, double scale, double dxc, const double* R);
int oracle_average_down(const mfm_params* p, const mfm_box* fine, int nf, const mfm_box* crse,
                        int ncb, int r, mfm_fab* Uc, const mfm_fab* Uf, double dxf);

/* ---- pointwise kernels, for the pinning tests ------------------------- */
/* Primitive -> face flux.  wl/wr: plasma primitive states in the face's
 * rotated order is NOT applied: wl/wr are (rho,u,v,w,p,bx,by,bz[,psi]) in
 * Cartesian order and idir selects the normal; out receives 8 (9) fluxes. */
void oracle_flux_roe8(const mfm_params* p, const double* wl, const double* wr,
                      int idir, int averaging, int lax_friedrichs, double* out);
void oracle_flux_rusanov(const mfm_params* p, const double* wl,
                         const double* wr, int idir, int averaging, double* out);
void oracle_flux_dedner(const mfm_params* p, double ch, const double* wl,
                        const double* wr, int idir, int averaging, double* out);
/* the solver the patch update selects from p->flux and p->dedner */
void oracle_flux_face(const mfm_params* p, double ch, const double* wl,
                      const double* wr, int idir, double* out);
void oracle_flux_vanleer(const mfm_params* p, const double* wl,
                         const double* wr, int idir, double* out);
/* the problem's physical-boundary ghost fill of direction dir on a
   primitive FAB (HELIOGS, LISMINIT and the axis reflection, the Riemann
   tube's copies); dx the cell size */
int oracle_fill_ghost(const mfm_params* p, mfm_fab W, int dir, double dx);
/* SWLISMREINIT (swlism/SWLISMF.ChF:305-403) on the conserved FAB U over
   box b; dx the cell size */
int oracle_swlism_reinit(const mfm_params* p, mfm_fab U, mfm_box b, double dx);
/* BFLUX / BFLUXR (swlism/HeliosphericF.ChF:1363-1774) in face coordinates:
   win, wout = (rho, vn, vt1, vt2, p, Bn, Bt1, Bt2) inside and outside the
   boundary, dxyz the outward direction (BFLUXR only); out = 8 fluxes,
   out[8] = the face's normal field, out[9] = 1 when the sub-fast outflow
   branch formed a characteristic boundary state, then in out[10..17]
   (out[10..17] must hold NaN on entry) */
void oracle_helio_bflux(const mfm_params* p, int radial, const double* win,
                        const double* wout, const double* dxyz, double* out);
/* HELIOBC (swlism/HeliosphericF.ChF:221-925) at one physical boundary face
   of cell (ijk) in direction dir on side lohisign: wleft = W+ of the cell
   below the face, wright = W- of the cell above (full primitive vectors);
   f = the face's flux vector, bn = its normal field */
void oracle_helio_bc_face(const mfm_params* p, int lohisign, int dir, double dx,
                          const int* ijk, const double* wleft,
                          const double* wright, double* f, double* bn);
/* BFLUXGAS (swlism/SWLISMF.ChF:1434-1553): win, wout = (rho, vn, vt1, vt2,
   p) of a neutral population; f5 holds the Riemann flux on entry and the
   boundary flux on return */
void oracle_bfluxgas(const mfm_params* p, const double* win,
                     const double* wout, double* f5);
/* 1-D face values of cell 0 from (w_-1, w_0, w_+1): out[0]=W-, out[1]=W+ */
void oracle_limiter(int lim, double wm, double w0, double wp, double* out);
/* cons -> prim for one cell (CONSTOPRIM / CONSTOPRIM_MF) */
void oracle_cons_to_prim(const mfm_params* p, const double* u, double* w);
/* physical flux (FLUXESHANCOCK, srcIdeal/PatchIdealMHDF.ChF:235-297) */
void oracle_physical_flux(const mfm_params* p, const double* w, int idir,
                          double* out);
/* charge exchange source for one cell (CHARGE_EXCHANGE_4F etc.) */
void oracle_charge_exchange(const mfm_params* p, const double* w, int reg,
                            double dt, double* s);
/* DEFINE_REGIONS_2F for one cell at global index (i,j,k) */
int oracle_region(const mfm_params* p, const double* w, int i, int j, int k,
                  double dx);
/* MAXWAVESPEED + MAXWAVESPEED_E for one cell */
double oracle_max_wave_speed(const mfm_params* p, const double* u);
/* POSTPROCESSING / POSTPROCESSING_MF positivity floors on one cell (in place) */
void oracle_floors(const mfm_params* p, double* u);

#ifdef __cplusplus
}
#endif
#endif
