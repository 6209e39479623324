/*
 * ddns.h -- C ABI of the MI355X droplet-laden DNS step library (libddns.so).
 *
 * One handle owns one periodic [0,2pi)^3 domain resident on one GPU: the spectral
 * velocity state, the Adams-Bashforth history, the stochastic-forcing state and the
 * droplet arrays.  The entry points replace the per-timestep block of the reference
 * driver loop (turb.f90:506-691 flow, turb.f90:546-623 droplets) and the routines it
 * calls.  Each declaration names the reference interface it stands in for.
 *
 * Conventions (all entry points):
 *   - plain pointers and sizes; host arrays are caller-owned and copied synchronously;
 *   - arrays use the reference's Fortran column-major layouts, 1-based in the comments;
 *   - return 0 on success, a negative DDNS_ERR_* code on a fatal error (the reference
 *     stops the run there), a positive DDNS_WARN_* code on a warning;
 *   - one host thread per handle; handles are not thread-safe.
 */
#ifndef DDNS_H
#define DDNS_H

#include <stdint.h>

#ifdef __cplusplus
extern "C" {
#endif

/* ---- status codes ------------------------------------------------------------- */
#define DDNS_OK                       0
#define DDNS_ERR_ARG                 -1   /* bad argument / unsupported size           */
#define DDNS_ERR_HIP                 -2   /* HIP runtime failure                       */
#define DDNS_ERR_OVERLAP             -3   /* overlapping droplets with HI on (partlib.f90:2653, 3294) */
#define DDNS_ERR_NEIGHBOUR_OVERFLOW  -4   /* > 100 HI neighbours (partlib.f90:3308)    */
#define DDNS_ERR_DROPLET_LOSS        -5   /* fewer than npart/2 droplets (turb.f90:600) */
#define DDNS_ERR_DOMAIN              -6   /* cell/halo too thin (partlib.f90:181-220)   */
#define DDNS_ERR_CAPACITY            -7   /* a device buffer would overflow            */
#define DDNS_ERR_IO                  -8   /* parameter file unreadable                 */
#define DDNS_ERR_NO_DEVICE           -9   /* no GPU visible                            */
#define DDNS_WARN_GMRES               1   /* HI GMRES did not converge (partlib.f90:2980) */

#define DDNS_FP64        0   /* every kernel in double precision                       */
#define DDNS_FP32_MIXED  1   /* FFT / nonlinear / HI iteration in float; droplet
                                positions, time integration and collision tests double */

#define DDNS_NBIN    180     /* RDF / relative-velocity bins (partlib.f90:36)          */
#define DDNS_MAXSET  10      /* radius sets (global.f90:22, radius(10))                */
#define DDNS_MAXNEI  100     /* HI neighbours per droplet (partlib.f90:3230)           */

/* Initial-spectrum choices for ddns_init_flow. */
#define DDNS_SPEC_KOLMOGOROV  0  /* k^-5/3 with E(1) tied to force(1) (flowlib.f90:113-123) */
#define DDNS_SPEC_K4_GAUSS    1  /* E ~ k^4 exp(-2 (k/k0)^2) scaled to u_rms (flowlib.f90:125-130, commented branch) */

/* Radius distributions for ddns_init_droplets. */
#define DDNS_RADII_SETS       0  /* npart/nset droplets per set, in set order (partlib.f90:275-276) */
#define DDNS_RADII_LOGNORMAL  1  /* per-droplet log-normal radius, clipped                   */

/*
 * Run parameters.  The first block mirrors parameter.d field by field, in the order
 * the reference reads it (turb.f90:103-164); the second block holds what the
 * reference hard-codes, exposed with the reference value as default.
 */
typedef struct ddns_params {
    /* --- parameter.d, turb.f90:103-164 --- */
    int    nx, ny, nz;          /* grid; must be equal powers of two, 16..1024          */
    int    nprocY;              /* read for file compatibility; the domain is not split */
    int    nstep;
    double dt;
    double rnu;
    double time;
    int    new_flow;
    int    idp;
    int    iseed_m0;            /* initial-flow seed                                   */
    int    forcing;             /* 1 deterministic, 2 stochastic                        */
    double force[2];
    int    iseedf;              /* stochastic-forcing seed                              */
    int    ieout;
    int    ispec;
    int    divstat;
    double rho_water;
    double rho_air;
    double etk_DNS, tauk_DNS;
    double st;                  /* shell thickness in percent                           */
    double ediss_CLOUD;
    double viscosity_CLOUD;
    double gravity;
    int    new_particle;
    int    nonoverlap;
    int    hdi_included;
    double hdi_trunc;
    int    npart;
    int    ndrag;
    int    nset;
    int    pad0;
    double radius[DDNS_MAXSET]; /* micrometres, as written in parameter.d               */

    /* --- hard-coded in the reference; defaults reproduce it --- */
    int    precision;           /* DDNS_FP64 or DDNS_FP32_MIXED                          */
    int    nstep_grow;          /* coalescence starts here (turb.f90:528), 40000         */
    double ou_tf;               /* OU time scale, 0.038 (flowlib.f90:402)                */
    double ou_var;              /* OU variance, 447.31 (flowlib.f90:401)                 */
    int    gmres_maxiter;       /* 50 (partlib.f90:2816)                                 */
    int    capacity;            /* droplet slots on the device; 0 means npart            */
    double gmres_tol;           /* 1e-6 (partlib.f90:2815)                               */
    int    relocate_seed;       /* ranff2 seed, -10*rank = 0 on one domain (partlib.f90:89) */
    int    istep0;              /* step counter before the first ddns_step call          */
} ddns_params;

/* Flow statistics of the current u-hat, as POSTPROC computes them
 * (postprocessing.f90:76-272) for monitora.dat, monitorb.dat and monitorc.dat. */
typedef struct ddns_monitors {
    /* monitora.dat (postprocessing.f90:208) */
    double eta, tmse, xintls, tk, et, xl, uprime;
    /* monitorb.dat (postprocessing.f90:209); cfl = pmax dt / dx (the reference uses dt/2
     * at jstep 0, postprocessing.f90:206) */
    double qq, dissp, ree, resol, cfl, vskew, vflat;
    double pmax;                             /* max over the grid of |u|+|v|+|w|   */
    /* monitorc.dat (postprocessing.f90:217-270), filled when divstat is set */
    double vdiv1, vdiv2, vort1, vort2, divvt;
} ddns_monitors;

/* Per-step collision outcome (COLLISION_DET, partlib.f90:1166-1846). */
typedef struct ddns_collision_counts {
    int     icc[3];                          /* type I, II, III collisions          */
    int     ncoll;                           /* entries of the removal list         */
    int     npairs;                          /* detected colliding pairs            */
    int     nset;
    int     colltype[DDNS_MAXSET * DDNS_MAXSET]; /* colltype(i1,i2), column-major     */
    int     nir[DDNS_MAXSET * DDNS_MAXSET];      /* contact counts nir(i1,i2)         */
    int64_t pairs_tested;                    /* candidate pairs whose distance was evaluated */
    int64_t pairs_in_dstat;                  /* pairs closer than 10*rcoll           */
    int     noverlap_sweeps;                 /* OVERLAP_DET sweeps that relocated    */
    int     noverlap_relocated;              /* relocations done by OVERLAP_DET     */
    int     gmres_iters;                     /* HI GMRES iterations this step        */
    int     pad1;
    double  gmres_residual;                  /* final relative residual              */
} ddns_collision_counts;

/* Kernel-level timing of the last ddns_step call, from HIP events (milliseconds). */
typedef struct ddns_timing {
    double flow_ms;           /* spectral step: FFT stages + fused update              */
    double interp_ms;         /* Lagrange gather                                       */
    double cells_ms;          /* cell binning (keys + sort)                            */
    double hdi_ms;            /* HI neighbour build + GMRES                            */
    double advance_ms;        /* drag/gravity advance + position update + wrap         */
    double colldet_ms;        /* collision detection + statistics + relocation         */
    double overlap_ms;        /* overlap sweeps                                        */
    double total_ms;          /* whole call                                            */
    int    nsteps;
    int    pad;
    /* flow kernels, summed over the call: inverse y stage, fused x stage, forward y
       stage, low-mode pre-kernel, z stage (see ddns_profile_flow), kx = 0
       symmetrisation */
    double flow_kernel_ms[6];
} ddns_timing;

typedef struct ddns_handle ddns_handle;

/* ---- parameters ------------------------------------------------------------------ */

/* Fill *p with the reference defaults (parameter.d of the 64^3 case, hard-coded values). */
int ddns_params_default(ddns_params* p);

/* Read a positional parameter.d (label lines alternate with value lines).
 * Replaces the rank-0 read block turb.f90:101-166 (and the HDI=>nonoverlap rule, 171). */
int ddns_read_parameter_file(const char* path, ddns_params* p);

/* ---- lifetime --------------------------------------------------------------------- */

/* Allocate the domain on the current GPU.  Replaces the allocation and setup block
 * turb.f90:282-420 (decomposition sizes, wavenumbers, shell counts) and the scaling
 * part of PART_INIT (partlib.f90:104-220, including its domain-size STOPs). */
int ddns_create(const ddns_params* p, ddns_handle** out);
int ddns_destroy(ddns_handle* h);

/* Text of the last error on this handle (static storage, never NULL). */
const char* ddns_last_error(const ddns_handle* h);

/* Number of shells nek = int(N/2 - 1.5) (turb.f90:303-304). */
int ddns_nek(const ddns_handle* h);

/* ---- flow state --------------------------------------------------------------------- */

/* Spectral velocity u-hat and Adams-Bashforth history o-hat, each three arrays in the
 * single-domain reference layout v(N+2, N+2, N): first index kz (rows N+1..N+2 are the
 * x-padding and are ignored), second index kx real/imag pairs, third index ky
 * (mpifft.f90:91, 105-137; flowlib.f90:12-33).  Any of the six pointers may be NULL. */
int ddns_set_flow(ddns_handle* h, const double* vx, const double* vy, const double* vz,
                  const double* ox, const double* oy, const double* oz);
int ddns_get_flow(ddns_handle* h, double* vx, double* vy, double* vz,
                  double* ox, double* oy, double* oz);

/* Stochastic-forcing state: b1r,b2r,b3r (6,5,5 each, 450 doubles) and the ranff
 * generator state (iseedf, iyf, ivf(32)) (turb.f90:441-447; load_save.f90:58-96). */
int ddns_set_forcing_state(ddns_handle* h, const double* b, int iseedf, int iyf, const int* ivf);
int ddns_get_forcing_state(ddns_handle* h, double* b, int* iseedf, int* iyf, int* ivf);

/* Generate the initial spectral field on the device (initial_flow, flowlib.f90:53-144,
 * with a counter-based SplitMix64 generator keyed by seed in place of IFPORT rand):
 * Gaussian modes, Hermitian kx = 0 plane, projection, shell-by-shell scaling.
 * kind: DDNS_SPEC_*; k0 and urms are used by DDNS_SPEC_K4_GAUSS.  The history is set
 * to the velocity (turb.f90:452-454).  For forcing=2 the reference starts from rest
 * (turb.f90:441-450): pass kind < 0 to get that. */
int ddns_init_flow(ddns_handle* h, int kind, int seed, double k0, double urms);

/* Physical velocity of the last step (after the inverse FFTs, turb.f90:537-539),
 * three arrays u(N,N,N) x-fastest, in double.  Valid after ddns_step. */
int ddns_get_physical_velocity(ddns_handle* h, double* ux, double* uy, double* uz);

/* Energy spectrum E(1..nek) of the current u-hat (postprocessing.f90:151-170). */
int ddns_get_spectrum(ddns_handle* h, double* ek);

/* Allocation-free E(k) (no counterpart in the reference): the values of
 * ddns_get_spectrum written to ek_dev, a device array of nek doubles, by kernels
 * queued on stream (a hipStream_t; NULL means the handle's own stream, which orders
 * the call after the last step).  It allocates nothing, copies nothing and does not
 * wait, so it is for callers that cannot allocate device memory: code under HIP
 * graph capture and loops that forbid hipMalloc.  Its work buffer is a fixed 1 MiB
 * device array per GPU (512 shells x 256 partial rows), held until the process ends;
 * a grid with more than 512 shells is refused with DDNS_ERR_ARG.  Calls that overlap
 * on one GPU share that array and race. */
int ddns_get_spectrum_device(ddns_handle* h, double* ek_dev, void* stream);

/* Energy, dissipation and transfer spectra of the current u-hat, nek shells each (any
 * pointer may be NULL): the spectrum.dat pairs of POSTPROC (postprocessing.f90:151-170)
 * and TRANSPOSTPROC's transfer.dat (postprocessing.f90:275-315, with the dealiased
 * nonlinear term of this u-hat, as the next step forms it, turb.f90:634-655). */
int ddns_get_spectra(ddns_handle* h, double* ek, double* dk, double* tk);

/* POSTPROC's statistics of the current u-hat (postprocessing.f90:76-272); divstat as
 * parameter.d's divstat (turb.f90:131) adds the monitorc.dat sums. */
int ddns_get_monitors(ddns_handle* h, int divstat, ddns_monitors* m);

/* ---- droplets ------------------------------------------------------------------------ */

/* Droplet state in the reference's arrays: yp(8,n) = x,y,z, Vx,Vy,Vz, radius (DNS
 * units), global id; fvp(3,0:2,n); fdp(3,0:3,n); pertvel(3,n) (global.f90:48-54).
 * fvp/fdp/pertvel may be NULL (zeros).  n <= capacity. */
int ddns_set_droplets(ddns_handle* h, int n, const double* yp, const double* fvp,
                      const double* fdp, const double* pertvel);
/* Also returns vp(3,n) (fluid velocity at the droplets) and dmove(3,n) when non-NULL. */
int ddns_get_droplets(ddns_handle* h, int* n, double* yp, double* fvp, double* fdp,
                      double* pertvel, double* vp, double* dmove);

/* PART_INIT positions and radii (partlib.f90:227-281): ranpp draws (seed -seed),
 * positions 2*pi*r, radii per set or log-normal (median_um, sigma_ln, clipped to
 * [rmin_um, rmax_um]), then PERIODIC and, if nonoverlap, OVERLAP_DET (turb.f90:426-429). */
int ddns_init_droplets(ddns_handle* h, int seed, int radii_kind, double median_um,
                       double sigma_ln, double rmin_um, double rmax_um);

/* OVERLAP_DET alone (partlib.f90:2047-2167): relocate until no two droplets overlap. */
int ddns_overlap_fix(ddns_handle* h);

/* ---- the step ------------------------------------------------------------------------ */

/* Advance nsteps time steps: the flow block turb.f90:506-691 and, when npart > 0, the
 * droplet block turb.f90:546-623, with the reference's data flow (the droplets see the
 * physical velocity of the step's start; the flow does not see the droplets), the two
 * blocks overlapped on two streams.  Returns, with all work done, the first nonzero
 * status (a warning does not stop the loop).
 * The two halves of the flow step are captured into HIP graphs by the first step and
 * replayed from then on (DDNS_GRAPH=0 in the environment at ddns_create launches the
 * kernels one by one; the results are the same bits).  Without droplets the call then
 * waits for the device once, at its end, and ddns_timing holds the whole call's time as
 * flow_ms with no per-kernel split; with droplets the droplet block waits each step on
 * device results (GMRES convergence, buffer sizes, relocations, overlap sweeps). */
int ddns_step(ddns_handle* h, int nsteps);

/* Outcome of the last step's COLLISION_DET (and HI solve). */
int ddns_get_collision_counts(ddns_handle* h, ddns_collision_counts* out);

/* Bin increments of the last step (nirbin0 ... wrtan0 of partlib.f90:1207-1222), each
 * nset*nset*DDNS_NBIN, column-major (i1, i2, ir).  Any pointer may be NULL. */
int ddns_get_collision_bins(ddns_handle* h, int* nirbin, double* wrbin, double* wxbin,
                            double* wybin, double* wzbin, double* urbin,
                            double* thetabin, double* wrtanbin);

/* Canonical list of the last step's colliding pairs: triples (min gid, max gid, type)
 * sorted ascending, type 1..3 (partlib.f90:1464-1562).  *npairs gets the full count;
 * at most cap triples are written. */
int ddns_get_collision_pairs(ddns_handle* h, int64_t* triples, int cap, int* npairs);

/* Records of the last step's droplet pairs in contact (rcolla < d <= rcollb), the
 * buffer(1:12, ij) that COLLISION_DET writes to Vp<i1><i2>.dat (partlib.f90:1431-1437,
 * 1775-1797): 12 doubles each, index1, index2, va(3), vb(3), unit offset xd(3), Wr, in
 * the one-rank visiting order (ih1, cell loop, linked list).  *n gets the full count;
 * at most cap records are written (rec may be NULL). */
int ddns_get_contact_records(ddns_handle* h, double* rec, int cap, int* n);

/* Tile shapes chosen at ddns_create by timing the candidates (no counterpart in the
 * reference; the role its FFTW_MEASURE plans play, mpifft.f90:36-56):
 *   shapes[0] = kx columns per y-stage tile, shapes[1] = its default;
 *   shapes[2] = threads of the fused z-stage work-group, shapes[3] = its default;
 *   ms[0..2] = time of one inverse (6 fields) plus forward (3 fields) y stage at half,
 *   once and twice the default width; ms[3..5] = time of one fused z stage at once
 *   and twice the default threads (ms[5] is kept for a third size and is 0), 0 where a
 *   shape does not fit the tile constraints or was not timed (the z stage where the
 *   step runs it as separate kernels, from N = 512).
 * A shape changes how work is grouped, never a column's or a mode's arithmetic, so
 * results are the same bits.  Either pointer may be NULL; shapes takes 4 ints, ms 6
 * doubles. */
int ddns_get_tuning(ddns_handle* h, int* shapes, double* ms);

/* HIP-event timings of the last ddns_step call. */
int ddns_get_timing(ddns_handle* h, ddns_timing* out);

/* ---- kernel-level entry points (parity tests, benchmarks) ----------------------------- */

/* mpifft3DRC (mpifft.f90:81-142): physical v(N+2,N+2,N) -> spectral, same layout, total
 * normalisation 1/N^3, in the handle's precision.  In place on the host array. */
int ddns_fft_r2c(ddns_handle* h, double* v);
/* mpifft3DCR (mpifft.f90:147-198): spectral -> physical, unnormalised. */
int ddns_fft_c2r(ddns_handle* h, double* v);

/* COLLISION_DET on a given droplet state without advancing it (partlib.f90:1166-1846):
 * yp(8,n) positions at the step start, dmove(3,n), fvp0(3,n) old velocities, vnew(3,n)
 * new velocities (yp(4:6) after PART_ADVANCE), vp(3,n) fluid velocity at the droplets.
 * Fills the counts/bins/pairs getters as ddns_step does.  Device-side FP64 tests. */
int ddns_collision_detect(ddns_handle* h, int n, const double* yp, const double* dmove,
                          const double* fvp0, const double* vnew, const double* vp);

/* Kernel timing helper (no counterpart in the reference): run the flow step alone k
 * times and return the mean time (ms) of, in order: inverse y stage, fused x stage,
 * forward y stage, low-mode pre-kernel, z stage (one fused kernel below N = 512; from
 * N = 512 the three forward transforms, the update and the six inverse transforms),
 * kx = 0 symmetrisation, whole flow step.  nkernels <= 7 values are written.
 * The flow state (u-hat, history) advances by k steps; time, the step counter and the
 * droplets do not. */
int ddns_profile_flow(ddns_handle* h, int k, double* ms_per_kernel, int nkernels);

/* HBM calibration helper (no counterpart in the reference): a plain streaming copy
 * between two of the handle's spectral work buffers, run reps times.  Returns the mean
 * time per launch (ms) and the bytes the copy moves per launch (read + write), so
 * that profiler byte counters can be scaled against a kernel of known traffic.  The
 * flow state is kept: the next step rebuilds its work buffers from u-hat. */
int ddns_profile_copy(ddns_handle* h, int reps, double* ms_per_launch, double* bytes_per_launch);

/* ---- batched realisations ---------------------------------------------------------- */

/* R independent realisations advanced together (no counterpart in the reference, which
 * runs one realisation per job: the step turb.f90:492-707 over the state of one
 * parameter.d).  params[r] is realisation r's parameter set; all share the grid (N <= 256),
 * the precision, dt and the viscosity, and each keeps its own forcing kind and amplitudes
 * (forcing, force, OU time scale and variance), seeds, droplets, HI and collision settings.
 * Every flow stage of a step is one launch over all realisations, replayed from HIP
 * graphs; each realisation's droplet block runs as a single handle's does, beside the
 * flow, the blocks spread over worker threads (DDNS_BATCH_THREADS, default 8). */
typedef struct ddns_batch ddns_batch;
int ddns_batch_create(const ddns_params* params, int nreal, ddns_batch** out);
int ddns_batch_destroy(ddns_batch* b);
int ddns_batch_size(const ddns_batch* b);
/* Realisation r as a handle, for ddns_init_flow / set_flow / get_flow / get_spectrum /
 * init_droplets / set_droplets / get_droplets / the collision getters / the forcing state.
 * It belongs to the batch: ddns_step and ddns_destroy refuse it. */
ddns_handle* ddns_batch_member(ddns_batch* b, int r);
/* nsteps steps of every realisation (turb.f90:492-707 each).  Returns the first error of
 * any realisation (ddns_batch_last_error names it), else DDNS_WARN_GMRES when some
 * realisation's GMRES stopped short of its tolerance, else DDNS_OK. */
int ddns_batch_step(ddns_batch* b, int nsteps);
/* Timings of the last ddns_batch_step: flow_ms and flow_kernel_ms for the batched flow
 * launches, total_ms for the call; the droplet phases summed over the realisations (each
 * member's own are in its ddns_get_timing). */
int ddns_batch_get_timing(ddns_batch* b, ddns_timing* out);
/* As ddns_get_tuning, for the tile shapes timed on the batch's fields at creation. */
int ddns_batch_get_tuning(ddns_batch* b, int* shapes, double* ms);
const char* ddns_batch_last_error(const ddns_batch* b);

#ifdef __cplusplus
}
#endif

#endif /* DDNS_H */
