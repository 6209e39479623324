/*
 * ddns_oracle.h -- FP64 CPU restatement of the reference hot path (the oracle).
 *
 * TEST INFRASTRUCTURE ONLY.  The oracle is the checker of the HIP library: it is
 * loaded by tests/, by __graft_entry__.smoke() and by bench.py's cpu_baseline leg,
 * never by the product path.  The entry points mirror include/ddns.h so that a test
 * drives the GPU library and the oracle with the same call sequence.
 */
#ifndef DDNS_ORACLE_H
#define DDNS_ORACLE_H

#include <stdint.h>
#include "../include/ddns.h"

#ifdef __cplusplus
extern "C" {
#endif

typedef struct orc_sim orc_sim;

int  orc_create(const ddns_params* p, orc_sim** out);
void orc_destroy(orc_sim* s);
int  orc_nek(const orc_sim* s);

uint64_t orc_splitmix64(uint64_t key, uint64_t i);
int  orc_init_flow(orc_sim* s, int kind, int seed, double k0, double urms);
int  orc_set_flow(orc_sim* s, const double* vx, const double* vy, const double* vz,
                  const double* ox, const double* oy, const double* oz);
int  orc_get_flow(orc_sim* s, double* vx, double* vy, double* vz,
                  double* ox, double* oy, double* oz);
int  orc_set_forcing_state(orc_sim* s, const double* b, int iseedf, int iyf, const int* ivf);
int  orc_get_forcing_state(orc_sim* s, double* b, int* iseedf, int* iyf, int* ivf);
int  orc_set_droplets(orc_sim* s, int n, const double* yp, const double* fvp,
                      const double* fdp, const double* pertvel);
int  orc_get_droplets(orc_sim* s, int* n, double* yp, double* fvp, double* fdp,
                      double* pertvel, double* vp, double* dmove);
int  orc_set_new_particle(orc_sim* s, int flag);
int  orc_num_threads(void);
int  orc_set_istep(orc_sim* s, int istep);

int  orc_step(orc_sim* s, int nsteps);
int  orc_flow_only(orc_sim* s, int on);          /* skip the droplet block when on     */
int  orc_get_spectrum(orc_sim* s, double* ek);
int  orc_get_spectra(orc_sim* s, double* ek, double* dk, double* tk);
int  orc_get_monitors(orc_sim* s, int divstat, ddns_monitors* m);
int  orc_get_physical_velocity(orc_sim* s, double* ux, double* uy, double* uz);
int  orc_get_collision_counts(orc_sim* s, ddns_collision_counts* out);
int  orc_get_collision_bins(orc_sim* s, int* nirbin, double* wrbin, double* wxbin,
                            double* wybin, double* wzbin, double* urbin,
                            double* thetabin, double* wrtanbin);
int  orc_get_collision_pairs(orc_sim* s, int64_t* triples, int cap, int* npairs);
/* The last detection's contact records in visiting order, 12 doubles each: index1,
 * index2, va(3), vb(3), unit offset xd(3), Wr (buffer(1:12,ij), partlib.f90:1431-1437). */
int  orc_get_contact_records(orc_sim* s, double* rec, int cap, int* n);
int  orc_get_removal_list(orc_sim* s, int* list, int cap, int* n);
int  orc_collision_detect(orc_sim* s, int n, const double* yp, const double* dmove,
                          const double* fvp0, const double* vnew, const double* vp);
int  orc_overlap_fix(orc_sim* s);
int  orc_get_relocate_state(orc_sim* s, int* iseedr, int* iyr, int* ivr, int* idum2);

/* Derived constants of PART_INIT (partlib.f90:104-150): radius[10] in DNS units,
 * then radmax, Sv_const, tau_CLOUD_const2, d_coll, d_hdi, gravity_dns. */
int  orc_get_scaling(orc_sim* s, double* out /* 16 */);

/* Building blocks, exposed for unit tests. */
int    orc_fft_r2c(int N, double* v);            /* mpifft3DRC on v(N+2,N+2,N)        */
int    orc_fft_c2r(int N, double* v);            /* mpifft3DCR                        */
int    orc_fft1d(int n, double* re_im, int sign);/* unnormalised complex DFT          */
double orc_ranff(int* idum, int* iy, int* iv);   /* accesories.f90:38-68              */
double orc_ranpp(int* idum, int* iy, int* iv);   /* accesories.f90:73-103             */
double orc_ranff2(int* idum, int* iv, int* iy, int* idum2); /* accesories.f90:107-147 */
int    orc_interp_weights(double y, double dx, int* node, double* bg /* 6 */);
int    orc_symmetrize(int N, double* c);         /* flowlib.f90:233-334, intended form */
int    orc_dealias(int N, double* v);            /* flowlib.f90:180-191               */
int    orc_supfor_det(int N, double* vx, double* vy, double* vz, const double* force,
                      double* e_before /* 2 */);
int    orc_supfor_sth(orc_sim* s, double* vx, double* vy, double* vz, double* a_out /*450*/);
int    orc_shell_counts(int N, double* enumv);   /* turb.f90:405-418                  */
int    orc_hdi_solve(orc_sim* s, int n, const double* yp, const double* vold,
                     const double* vp, double* pertvel, int* iters, double* res);
int    orc_init_droplets(orc_sim* s, int seed, int radii_kind, double median_um,
                         double sigma_ln, double rmin_um, double rmax_um);

#ifdef __cplusplus
}
#endif
#endif
