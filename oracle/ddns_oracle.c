/*
 * ddns_oracle.c -- FP64 CPU restatement of the reference per-timestep hot path:
 * the pseudo-spectral Navier-Stokes step of turb.f90 / mpifft.f90 / flowlib.f90 and
 * the Lagrangian droplet step of partlib.f90, on one periodic domain [0,2pi)^3.
 *
 * TEST INFRASTRUCTURE ONLY.  This file is the oracle: tests/, __graft_entry__.smoke()
 * and bench.py's cpu_baseline leg load it as the checker.  The product (libddns.so)
 * neither links nor calls it.
 *
 * Semantics.  The restatement follows the reference statement by statement, in the
 * reference's array layouts (Fortran column-major, REAL = 8 bytes) and evaluation
 * order, compiled with strict IEEE arithmetic (-ffp-contract=off).  Where the
 * reference is decomposition-dependent or undefined, the single-domain choice is:
 *   - symmetrize zeroes the ky = N/2 plane (the behaviour of every even nprocZ >= 2;
 *     with nprocZ = 1 the reference zeroes ky = 0 instead, flowlib.f90:256);
 *   - the x-padding rows of spectral arrays are zero: the inverse x-transform
 *     therefore sees a zero Nyquist coefficient, exactly as the reference does when
 *     its padding is zero (mpifft.f90:189 leaves rows lx+1:lx+2 untouched);
 *   - droplets keep their array slots: PERIODIC only wraps positions (partlib.f90:
 *     1142-1151); migration between ranks does not exist on one domain;
 *   - HDI_GMRES returns early on a small initial residual without leaking its
 *     neighbour arrays (partlib.f90:2866-2868);
 *   - neighbour-array overflow and endless overlap sweeps return an error code.
 * Everything is FP64.  OpenMP is used only where each output element is produced by
 * one thread (FFT lines, element-wise passes, per-droplet gathers): results do not
 * depend on the thread count.
 *
 * Parity status of each function is stated where it is defined ("pinned by ...").
 */
#include "ddns_oracle.h"

#include <math.h>
#include <stdlib.h>
#include <string.h>
#include <stdio.h>

#ifdef _OPENMP
#include <omp.h>
#endif

#define NTAB 32

/* ------------------------------------------------------------------------------- */
/* Random number generators (accesories.f90).  Pinned bit-exactly against the      */
/* reference's own accesories.f90 compiled by oracle/build_ref.sh, and the Park-    */
/* Miller core against the published known answer.                                */
/* ------------------------------------------------------------------------------- */

/* ranff: Park-Miller minimal standard with Bays-Durham shuffle, accesories.f90:38-68.
 * iv is iv(1:32) stored 0-based. */
double orc_ranff(int* idum, int* iy, int* iv)
{
    const int IA = 16807, IM = 2147483647, IQ = 127773, IR = 2836;
    const int NDIV = 1 + (IM - 1) / NTAB;
    const double AM = 1.0 / IM, EPS = 1.2e-7, RNMX = 1.0 - EPS;
    int j, k;
    if (*idum <= 0 || *iy == 0) {                       /* accesories.f90:49-58 */
        *idum = (-*idum > 1) ? -*idum : 1;
        for (j = NTAB + 8; j >= 1; --j) {
            k = *idum / IQ;
            *idum = IA * (*idum - k * IQ) - IR * k;
            if (*idum < 0) *idum += IM;
            if (j <= NTAB) iv[j - 1] = *idum;
        }
        *iy = iv[0];
    }
    k = *idum / IQ;                                     /* accesories.f90:59-65 */
    *idum = IA * (*idum - k * IQ) - IR * k;
    if (*idum < 0) *idum += IM;
    j = 1 + *iy / NDIV;
    *iy = iv[j - 1];
    iv[j - 1] = *idum;
    {
        double r = AM * (double)(*iy);
        return r < RNMX ? r : RNMX;
    }
}

/* ranpp: identical algorithm with its own state, accesories.f90:73-103. */
double orc_ranpp(int* idum, int* iy, int* iv) { return orc_ranff(idum, iy, iv); }

/* ranff2: L'Ecuyer with Bays-Durham shuffle, accesories.f90:107-147. */
double orc_ranff2(int* idum, int* iv, int* iy, int* idum2)
{
    const int IM1 = 2147483563, IM2 = 2147483399, IMM1 = IM1 - 1;
    const int IA1 = 40014, IA2 = 40692, IQ1 = 53668, IQ2 = 52774, IR1 = 12211, IR2 = 3791;
    const int NDIV = 1 + IMM1 / NTAB;
    const double AM = 1.0 / IM1, EPS = 1.2e-7, RNMX = 1.0 - EPS;
    int j, k;
    if (*idum <= 0) {                                   /* accesories.f90:123-133 */
        *idum = (-*idum > 1) ? -*idum : 1;
        *idum2 = *idum;
        for (j = NTAB + 8; j >= 1; --j) {
            k = *idum / IQ1;
            *idum = IA1 * (*idum - k * IQ1) - k * IR1;
            if (*idum < 0) *idum += IM1;
            if (j <= NTAB) iv[j - 1] = *idum;
        }
        *iy = iv[0];
    }
    k = *idum / IQ1;                                    /* accesories.f90:135-145 */
    *idum = IA1 * (*idum - k * IQ1) - k * IR1;
    if (*idum < 0) *idum += IM1;
    k = *idum2 / IQ2;
    *idum2 = IA2 * (*idum2 - k * IQ2) - k * IR2;
    if (*idum2 < 0) *idum2 += IM2;
    j = 1 + *iy / NDIV;
    *iy = iv[j - 1] - *idum2;
    iv[j - 1] = *idum;
    if (*iy < 1) *iy += IMM1;
    {
        double r = AM * (double)(*iy);
        return r < RNMX ? r : RNMX;
    }
}

/* ------------------------------------------------------------------------------- */
/* 1-D complex DFT (radix-2, unnormalised).  Stands in for FFTW 2.1.5 (un-vendored, */
/* Makefile:7-11): pinned by numpy.fft golden vectors and closed forms.             */
/* ------------------------------------------------------------------------------- */

typedef struct { int n; double* c; double* s; int* rev; } fftplan;

static fftplan* plan_cache[32];

static const fftplan* get_plan(int n)
{
    int lg = 0;
    while ((1 << lg) < n) ++lg;
    if ((1 << lg) != n || lg >= 32) return NULL;
    if (plan_cache[lg]) return plan_cache[lg];
    {
        fftplan* p = (fftplan*)calloc(1, sizeof(fftplan));
        int i;
        p->n = n;
        p->c = (double*)malloc(sizeof(double) * (n / 2 + 1));
        p->s = (double*)malloc(sizeof(double) * (n / 2 + 1));
        p->rev = (int*)malloc(sizeof(int) * n);
        for (i = 0; i <= n / 2; ++i) {
            double ang = 2.0 * M_PI * (double)i / (double)n;
            p->c[i] = cos(ang);
            p->s[i] = sin(ang);
        }
        for (i = 0; i < n; ++i) {
            int r = 0, b;
            for (b = 0; b < lg; ++b) if (i & (1 << b)) r |= 1 << (lg - 1 - b);
            p->rev[i] = r;
        }
        /* plans are created before any parallel region touches them */
        plan_cache[lg] = p;
        return p;
    }
}

/* a[2k], a[2k+1] = re, im.  sign = -1 forward (exp(-i..)), +1 backward. */
static void fft_exec(const fftplan* p, double* a, int sign)
{
    int n = p->n, i, len;
    for (i = 0; i < n; ++i) {
        int r = p->rev[i];
        if (r > i) {
            double t0 = a[2 * i], t1 = a[2 * i + 1];
            a[2 * i] = a[2 * r]; a[2 * i + 1] = a[2 * r + 1];
            a[2 * r] = t0; a[2 * r + 1] = t1;
        }
    }
    for (len = 2; len <= n; len <<= 1) {
        int half = len / 2, step = n / len, st, j;
        for (st = 0; st < n; st += len) {
            for (j = 0; j < half; ++j) {
                double wr = p->c[j * step], wi = sign * p->s[j * step];
                double* u = a + 2 * (st + j);
                double* v = a + 2 * (st + j + half);
                double tr = wr * v[0] - wi * v[1];
                double ti = wr * v[1] + wi * v[0];
                v[0] = u[0] - tr; v[1] = u[1] - ti;
                u[0] = u[0] + tr; u[1] = u[1] + ti;
            }
        }
    }
}

int orc_fft1d(int n, double* re_im, int sign)
{
    const fftplan* p = get_plan(n);
    if (!p) return DDNS_ERR_ARG;
    fft_exec(p, re_im, sign);
    return 0;
}

/* ------------------------------------------------------------------------------- */
/* 3-D transforms in the single-domain reference layout v(N+2, N+2, N).             */
/* ------------------------------------------------------------------------------- */

#define IDX(i, j, k, L) ((size_t)(i) + (size_t)(L) * ((size_t)(j) + (size_t)(L) * (size_t)(k)))

/* mpifft3DRC, mpifft.f90:81-142: R2C along x, /N; C2C forward along y, /N; C2C
 * forward along z, /N.  Input v(x, y, z) physical (x, y in 1..N); output the
 * spectral layout v(kz, kx re/im, ky) with zero x-padding rows.  The kx = N/2
 * pair is computed, as the reference does.  Pinned by numpy.fft vectors. */
int orc_fft_r2c(int N, double* v)
{
    const int L = N + 2;
    const fftplan* p = get_plan(N);
    size_t tot = (size_t)L * L * N;
    double* t;
    int k;
    if (!p) return DDNS_ERR_ARG;
    t = (double*)malloc(sizeof(double) * tot);
    if (!t) return DDNS_ERR_ARG;
    /* stage 1: x lines, mpifft.f90:97-103; t holds (jx, y, z) */
#pragma omp parallel for schedule(static)
    for (k = 0; k < N; ++k) {
        double* line = (double*)malloc(sizeof(double) * 2 * N);
        int j, i;
        for (j = 0; j < L; ++j) {
            if (j >= N) { for (i = 0; i < L; ++i) t[IDX(i, j, k, L)] = 0.0; continue; }
            for (i = 0; i < N; ++i) { line[2 * i] = v[IDX(i, j, k, L)]; line[2 * i + 1] = 0.0; }
            fft_exec(p, line, -1);
            for (i = 0; i <= N / 2; ++i) {
                t[IDX(2 * i, j, k, L)] = line[2 * i] / (double)N;
                t[IDX(2 * i + 1, j, k, L)] = line[2 * i + 1] / (double)N;
            }
        }
        free(line);
    }
    /* stage 2: y columns for every kx pair, mpifft.f90:105-121 */
#pragma omp parallel for schedule(static)
    for (k = 0; k < N; ++k) {
        double* col = (double*)malloc(sizeof(double) * 2 * N);
        int m, j;
        for (m = 0; m <= N / 2; ++m) {
            for (j = 0; j < N; ++j) { col[2 * j] = t[IDX(2 * m, j, k, L)]; col[2 * j + 1] = t[IDX(2 * m + 1, j, k, L)]; }
            fft_exec(p, col, -1);
            for (j = 0; j < N; ++j) { t[IDX(2 * m, j, k, L)] = col[2 * j] / (double)N; t[IDX(2 * m + 1, j, k, L)] = col[2 * j + 1] / (double)N; }
        }
        free(col);
    }
    /* stage 3: z columns, mpifft.f90:123-139; write the spectral layout v(kz, jx, ky) */
#pragma omp parallel for schedule(static)
    for (k = 0; k < N; ++k) {       /* k = ky index */
        double* col = (double*)malloc(sizeof(double) * 2 * N);
        int m, z;
        for (m = 0; m <= N / 2; ++m) {
            for (z = 0; z < N; ++z) { col[2 * z] = t[IDX(2 * m, k, z, L)]; col[2 * z + 1] = t[IDX(2 * m + 1, k, z, L)]; }
            fft_exec(p, col, -1);
            for (z = 0; z < N; ++z) {
                v[IDX(z, 2 * m, k, L)] = col[2 * z] / (double)N;
                v[IDX(z, 2 * m + 1, k, L)] = col[2 * z + 1] / (double)N;
            }
            v[IDX(N, 2 * m, k, L)] = 0.0; v[IDX(N + 1, 2 * m, k, L)] = 0.0;
            v[IDX(N, 2 * m + 1, k, L)] = 0.0; v[IDX(N + 1, 2 * m + 1, k, L)] = 0.0;
        }
        free(col);
    }
    free(t);
    return 0;
}

/* mpifft3DCR, mpifft.f90:147-198: C2C backward along z, along y, C2R along x,
 * unnormalised.  The x-Nyquist coefficient fed to the C2R comes from the zero
 * padding rows (mpifft.f90:189 never copies jx = N+1..N+2), so it is zero; the
 * imaginary part of the kx = 0 coefficient is ignored by a C2R.  Output padding
 * and y-extension rows are zero. */
int orc_fft_c2r(int N, double* v)
{
    const int L = N + 2;
    const fftplan* p = get_plan(N);
    size_t tot = (size_t)L * L * N;
    double* t;
    int k;
    if (!p) return DDNS_ERR_ARG;
    t = (double*)calloc(tot, sizeof(double));
    if (!t) return DDNS_ERR_ARG;
    /* z columns: input v(kz, jx, ky) -> t(jx, ky, z) */
#pragma omp parallel for schedule(static)
    for (k = 0; k < N; ++k) {
        double* col = (double*)malloc(sizeof(double) * 2 * N);
        int m, z;
        for (m = 0; m < N / 2; ++m) {
            for (z = 0; z < N; ++z) { col[2 * z] = v[IDX(z, 2 * m, k, L)]; col[2 * z + 1] = v[IDX(z, 2 * m + 1, k, L)]; }
            fft_exec(p, col, +1);
            for (z = 0; z < N; ++z) { t[IDX(2 * m, k, z, L)] = col[2 * z]; t[IDX(2 * m + 1, k, z, L)] = col[2 * z + 1]; }
        }
        free(col);
    }
    /* y columns for each z */
#pragma omp parallel for schedule(static)
    for (k = 0; k < N; ++k) {
        double* col = (double*)malloc(sizeof(double) * 2 * N);
        int m, j;
        for (m = 0; m < N / 2; ++m) {
            for (j = 0; j < N; ++j) { col[2 * j] = t[IDX(2 * m, j, k, L)]; col[2 * j + 1] = t[IDX(2 * m + 1, j, k, L)]; }
            fft_exec(p, col, +1);
            for (j = 0; j < N; ++j) { t[IDX(2 * m, j, k, L)] = col[2 * j]; t[IDX(2 * m + 1, j, k, L)] = col[2 * j + 1]; }
        }
        free(col);
    }
    /* x lines: C2R */
#pragma omp parallel for schedule(static)
    for (k = 0; k < N; ++k) {
        double* line = (double*)malloc(sizeof(double) * 2 * N);
        int j, i;
        for (j = 0; j < L; ++j) {
            if (j >= N) { for (i = 0; i < L; ++i) v[IDX(i, j, k, L)] = 0.0; continue; }
            line[0] = t[IDX(0, j, k, L)];
            line[1] = 0.0;
            for (i = 1; i < N / 2; ++i) {
                line[2 * i] = t[IDX(2 * i, j, k, L)];
                line[2 * i + 1] = t[IDX(2 * i + 1, j, k, L)];
                line[2 * (N - i)] = t[IDX(2 * i, j, k, L)];
                line[2 * (N - i) + 1] = -t[IDX(2 * i + 1, j, k, L)];
            }
            line[N] = 0.0; line[N + 1] = 0.0;
            fft_exec(p, line, +1);
            for (i = 0; i < N; ++i) v[IDX(i, j, k, L)] = line[2 * i];
            v[IDX(N, j, k, L)] = 0.0; v[IDX(N + 1, j, k, L)] = 0.0;
        }
        free(line);
    }
    free(t);
    return 0;
}

/* ------------------------------------------------------------------------------- */
/* Wavenumbers (flowlib.f90:3-49; turb.f90:403-406).  k-arrays are functions of the */
/* index here.  Pinned by the shell-count closed check (dealias keeps enum modes).  */
/* ------------------------------------------------------------------------------- */

static inline double wn_z(int i, int N) { return (i >= N) ? 0.0 : (double)((i <= N / 2) ? i : i - N); }
static inline double wn_x(int j) { return (double)(j / 2); }
static inline double wn_y(int k, int N) { return (double)((k <= N / 2) ? k : k - N); }
static inline double wn_k2(int i, int j, int k, int N)
{
    double a, b, c;
    if (i >= N) return 1.0;                                  /* flowlib.f90:46 */
    if (i == 0 && j <= 1 && k == 0) return 1.e-5;            /* flowlib.f90:36-38 */
    a = wn_x(j); b = wn_y(k, N); c = wn_z(i, N);
    return a * a + b * b + c * c;                            /* flowlib.f90:35 */
}
static inline int wn_ik2(int i, int j, int k, int N)
{
    if (i >= N) return 0;                                    /* turb.f90:406 */
    return (int)(sqrt(wn_k2(i, j, k, N)) + 0.5);             /* turb.f90:405 */
}

/* enum(i): full-space mode count of shell i (turb.f90:408-418). */
int orc_shell_counts(int N, double* enumv)
{
    const int L = N + 2, nek = (int)(N / 2 - 1.5);
    int s;
    for (s = 1; s <= nek; ++s) {
        double co1 = 0.0, co2 = 0.0;
        int i, j, k;
        for (k = 0; k < N; ++k)
            for (j = 0; j < L; ++j)
                for (i = 0; i < N; ++i) {
                    if (wn_ik2(i, j, k, N) != s) continue;
                    if (wn_x(j) > 0.1) co1 += 1.0; else co2 += 1.0;
                }
        enumv[s - 1] = co1 + co2 / 2.;
    }
    return nek;
}

/* dealiasing, flowlib.f90:180-191.  Pinned: the retained modes per shell equal the
 * closed-form count of lattice points, and every removed mode has int(|k|+.5) > nek. */
int orc_dealias(int N, double* v)
{
    const int L = N + 2, nek = (int)(N / 2 - 1.5);
    int k;
#pragma omp parallel for schedule(static)
    for (k = 0; k < N; ++k) {
        int i, j;
        for (j = 0; j < L; ++j)
            for (i = 0; i < L; ++i)
                if (wn_ik2(i, j, k, N) > nek) v[IDX(i, j, k, L)] = 0.0;
    }
    return 0;
}

/* symmetrize, flowlib.f90:233-334, in the form every even nprocZ >= 2 executes:
 * zero mean, zero the Nyquist planes kx = N/2, ky = N/2 (flowlib.f90:256 with
 * idlocalZ = nprocZ/2), kz = N/2, then on the kx = 0 plane replace c(k) by
 * (c(k) + conj c(-k))/2 using the values before averaging (flowlib.f90:268-302,
 * 321-329).  Pinned by: C2R of the result is real, R2C(C2R(c)) returns c. */
int orc_symmetrize(int N, double* c)
{
    const int L = N + 2;
    int i, j, k;
    double* re = (double*)malloc(sizeof(double) * N * N);
    double* im = (double*)malloc(sizeof(double) * N * N);
    c[IDX(0, 0, 0, L)] = 0.0; c[IDX(0, 1, 0, L)] = 0.0;                 /* 249 */
    for (k = 0; k < N; ++k)                                              /* 254 */
        for (i = 0; i < L; ++i) { c[IDX(i, N, k, L)] = 0.0; c[IDX(i, N + 1, k, L)] = 0.0; }
    for (j = 0; j < L; ++j)                                              /* 256 */
        for (i = 0; i < L; ++i) c[IDX(i, j, N / 2, L)] = 0.0;
    for (k = 0; k < N; ++k)                                              /* 258 */
        for (j = 0; j < L; ++j) c[IDX(N / 2, j, k, L)] = 0.0;
    for (k = 0; k < N; ++k)
        for (i = 0; i < N; ++i) { re[i + N * k] = c[IDX(i, 0, k, L)]; im[i + N * k] = c[IDX(i, 1, k, L)]; }
    /* ky != 0 planes: all kz (flowlib.f90:276-284, 297-302) */
    for (k = 1; k < N; ++k) {
        int kk = N - k;
        if (k == N / 2) continue;
        for (i = 0; i < N; ++i) {
            int ii = (i == 0) ? 0 : N - i;
            c[IDX(i, 0, k, L)] = 0.5 * (re[i + N * k] + re[ii + N * kk]);
            c[IDX(i, 1, k, L)] = 0.5 * (im[i + N * k] - im[ii + N * kk]);
        }
    }
    /* ky = 0 line, kz = 1..N/2-1 and mirrors (flowlib.f90:321-329) */
    for (i = 1; i < N / 2; ++i) {
        int ii = N - i;
        double r = 0.5 * (re[i] + re[ii]);
        double m = 0.5 * (im[i] - im[ii]);
        c[IDX(i, 0, 0, L)] = r;  c[IDX(i, 1, 0, L)] = m;
        c[IDX(ii, 0, 0, L)] = r; c[IDX(ii, 1, 0, L)] = -m;
    }
    free(re); free(im);
    return 0;
}

/* supforDET, flowlib.f90:195-229: shell energies with the kx = 0 plane halved,
 * summed in array order, then scale shells 1 and 2 to force(1:2).  Pinned by the
 * invariant E_s(after) = force(s) and by E_s(before) from a direct numpy sum. */
int orc_supfor_det(int N, double* vx, double* vy, double* vz, const double* force,
                   double* e_before)
{
    const int L = N + 2;
    int is;
    for (is = 1; is <= 2; ++is) {
        double ek = 0.0, ff;
        int i, j, k;
        for (k = 0; k < N; ++k)
            for (j = 0; j < L; ++j)
                for (i = 0; i < L; ++i) {
                    size_t o = IDX(i, j, k, L);
                    double t;
                    if (wn_ik2(i, j, k, N) != is) continue;
                    t = vx[o] * vx[o] + vy[o] * vy[o] + vz[o] * vz[o];
                    if (j <= 1) t = 0.5 * t;
                    ek += t;
                }
        if (e_before) e_before[is - 1] = ek;
        ff = sqrt(force[is - 1] / ek);
        for (k = 0; k < N; ++k)
            for (j = 0; j < L; ++j)
                for (i = 0; i < L; ++i) {
                    size_t o = IDX(i, j, k, L);
                    if (wn_ik2(i, j, k, N) != is) continue;
                    vx[o] = vx[o] * ff; vy[o] = vy[o] * ff; vz[o] = vz[o] * ff;
                }
    }
    return 0;
}

/* ------------------------------------------------------------------------------- */
/* Simulation state                                                                 */
/* ------------------------------------------------------------------------------- */

struct orc_sim {
    ddns_params p;
    int N, L, nek;
    size_t fsz;
    double *vx, *vy, *vz, *ox, *oy, *oz;       /* spectral state */
    double *wx, *wy, *wz, *vxc, *vyc, *vzc;    /* work / physical fields */
    int have_phys;
    int flow_only;
    int istep;
    double time;
    /* stochastic forcing, flowlib.f90:378-562 */
    double b1r[150], b2r[150], b3r[150];
    int iseedf, iyf, ivf[NTAB];
    /* droplet scaling, partlib.f90:104-150 */
    double radius[DDNS_MAXSET], radmax, Sv_const, tau_CLOUD_const2, gravity_dns;
    double d_coll, d_hdi, pi, pi2;
    int ncd, nx_coll, nx_hdi, nx_max;
    /* droplets */
    int cap, nps, new_particle;
    double *yp, *vp, *dmove, *fvp, *fdp, *pertvel;
    /* halo list (HEADLIST) */
    int ypn_rows, ypn_cap, nps2;
    double *ypn, *vpn;
    int *list, *head;
    int ih_p[8];
    int *pindex;                 /* pindex[d*cap + m] = local index (1-based) */
    int hy0, hny;                /* head y/z index offset and extent */
    /* relocation generator state, partlib.f90:89-92 */
    int iseedr, ivr[NTAB], iyr, idum2;
    /* HI */
    int *nei_num;
    double *nei_data;            /* nei_data[(ih*100 + m)*7 + c] */
    /* per-step collision outcome */
    ddns_collision_counts cc;
    int nbins;
    int* nirbin0;
    double *wrbin0, *wxbin0, *wybin0, *wzbin0, *urbin0, *thetabin0, *wrtan0;
    int64_t* pairs;              /* 3 per pair */
    int npairs, pairs_cap;
    double* contact;             /* 12 per pair in contact: buffer(1:12,ij) of partlib.f90:1431-1437 */
    int ncontact, contact_cap;
    int* removepart;
    int ncoll, remove_cap;
    double* collist;             /* 20 per record, partlib.f90:1557-1562 */
};

static void* xcalloc(size_t n, size_t sz) { return calloc(n ? n : 1, sz); }

int orc_nek(const orc_sim* s) { return s->nek; }

static void compute_scaling(orc_sim* s)
{
    const ddns_params* p = &s->p;
    double etk_CLOUD, tauk_CLOUD, vk_CLOUD, vk_DNS, vk_ratio, etk_ratio, tauk_ratio;
    double tau_CLOUD_const, hx;
    int il;
    s->pi = 4. * atan(1.);                                  /* turb.f90:296-297 */
    s->pi2 = s->pi * 2.0;
    hx = s->pi2;
    etk_CLOUD = sqrt(sqrt(p->viscosity_CLOUD * p->viscosity_CLOUD * p->viscosity_CLOUD / p->ediss_CLOUD));
    tauk_CLOUD = sqrt(p->viscosity_CLOUD / p->ediss_CLOUD);
    vk_CLOUD = etk_CLOUD / tauk_CLOUD;
    vk_DNS = p->etk_DNS / p->tauk_DNS;
    vk_ratio = vk_CLOUD / vk_DNS;
    etk_ratio = etk_CLOUD / p->etk_DNS;
    tauk_ratio = tauk_CLOUD / p->tauk_DNS;
    s->radmax = 0.0;
    for (il = 0; il < DDNS_MAXSET; ++il) s->radius[il] = 0.0;
    for (il = 0; il < p->nset; ++il) {                      /* partlib.f90:114-115 */
        s->radius[il] = p->radius[il] * (p->etk_DNS / etk_CLOUD) / 10000.0;
        if (il == 0 || s->radius[il] > s->radmax) s->radmax = s->radius[il];
    }
    tau_CLOUD_const = 2. * p->rho_water * (etk_ratio * etk_ratio) / (9. * p->rho_air * p->viscosity_CLOUD);
    s->Sv_const = tau_CLOUD_const * p->gravity / vk_ratio;  /* partlib.f90:117-120 */
    s->gravity_dns = p->gravity * tauk_ratio / vk_ratio;
    s->tau_CLOUD_const2 = tau_CLOUD_const / tauk_ratio;
    s->ncd = s->N;                                          /* partlib.f90:33-35 */
    s->d_coll = 20.0 * s->radmax;                           /* partlib.f90:130-150 */
    s->nx_coll = 1 + (int)(s->d_coll * (double)s->ncd / hx);
    s->nx_max = s->nx_coll;
    s->d_hdi = 0.0;
    s->nx_hdi = 0;
    if (p->hdi_included) {
        s->d_hdi = p->hdi_trunc * s->radmax;
        s->nx_hdi = 1 + (int)(s->d_hdi * (double)s->ncd / hx);
        if (s->nx_hdi > s->nx_max) s->nx_max = s->nx_hdi;
    }
}

int orc_get_scaling(orc_sim* s, double* out)
{
    int i;
    for (i = 0; i < 10; ++i) out[i] = s->radius[i];
    out[10] = s->radmax; out[11] = s->Sv_const; out[12] = s->tau_CLOUD_const2;
    out[13] = s->d_coll; out[14] = s->d_hdi; out[15] = s->gravity_dns;
    return 0;
}

int orc_create(const ddns_params* p, orc_sim** out)
{
    orc_sim* s;
    int N = p->nx, i;
    if (!out || N < 16 || N != p->ny || N != p->nz || (N & (N - 1))) return DDNS_ERR_ARG;
    s = (orc_sim*)calloc(1, sizeof(orc_sim));
    s->p = *p;
    if (s->p.hdi_included) s->p.nonoverlap = 1;             /* turb.f90:171 */
    s->N = N; s->L = N + 2; s->nek = (int)(N / 2 - 1.5);
    s->fsz = (size_t)s->L * s->L * N;
    s->vx = (double*)xcalloc(s->fsz, 8); s->vy = (double*)xcalloc(s->fsz, 8); s->vz = (double*)xcalloc(s->fsz, 8);
    s->ox = (double*)xcalloc(s->fsz, 8); s->oy = (double*)xcalloc(s->fsz, 8); s->oz = (double*)xcalloc(s->fsz, 8);
    s->wx = (double*)xcalloc(s->fsz, 8); s->wy = (double*)xcalloc(s->fsz, 8); s->wz = (double*)xcalloc(s->fsz, 8);
    s->vxc = (double*)xcalloc(s->fsz, 8); s->vyc = (double*)xcalloc(s->fsz, 8); s->vzc = (double*)xcalloc(s->fsz, 8);
    if (!s->vx || !s->wz || !s->vzc) { orc_destroy(s); return DDNS_ERR_ARG; }
    s->time = p->time;
    s->istep = p->istep0;
    /* stochastic forcing start, turb.f90:441-447 */
    s->iseedf = -p->iseedf; s->iyf = 0;
    for (i = 0; i < NTAB; ++i) s->ivf[i] = 0;
    compute_scaling(s);
    /* domain checks, partlib.f90:171-220 (hy = hz = 2 pi on one domain) */
    if (p->npart > 0) {
        if (s->pi2 <= s->d_coll || s->pi2 <= 2.0 * s->radmax ||
            (s->p.hdi_included && s->pi2 <= s->d_hdi) || 2 * s->nx_max + 1 >= s->ncd) {
            orc_destroy(s);
            return DDNS_ERR_DOMAIN;
        }
    }
    s->cap = p->capacity > 0 ? p->capacity : p->npart;
    if (s->cap < 1) s->cap = 1;
    s->yp = (double*)xcalloc((size_t)8 * s->cap, 8);
    s->vp = (double*)xcalloc((size_t)3 * s->cap, 8);
    s->dmove = (double*)xcalloc((size_t)3 * s->cap, 8);
    s->fvp = (double*)xcalloc((size_t)9 * s->cap, 8);
    s->fdp = (double*)xcalloc((size_t)12 * s->cap, 8);
    s->pertvel = (double*)xcalloc((size_t)3 * s->cap, 8);
    s->pindex = (int*)xcalloc((size_t)8 * s->cap, sizeof(int));
    s->new_particle = p->new_particle;
    s->ypn_rows = s->p.hdi_included ? 21 : 15;
    s->iseedr = p->relocate_seed;                           /* partlib.f90:89-92 */
    s->idum2 = 0; s->iyr = 0;
    for (i = 0; i < NTAB; ++i) s->ivr[i] = 0;
    s->nbins = (p->nset > 0 ? p->nset : 1) * (p->nset > 0 ? p->nset : 1) * DDNS_NBIN;
    s->nirbin0 = (int*)xcalloc(s->nbins, sizeof(int));
    s->wrbin0 = (double*)xcalloc(s->nbins, 8); s->wxbin0 = (double*)xcalloc(s->nbins, 8);
    s->wybin0 = (double*)xcalloc(s->nbins, 8); s->wzbin0 = (double*)xcalloc(s->nbins, 8);
    s->urbin0 = (double*)xcalloc(s->nbins, 8); s->thetabin0 = (double*)xcalloc(s->nbins, 8);
    s->wrtan0 = (double*)xcalloc(s->nbins, 8);
    s->pairs_cap = 1024;
    s->pairs = (int64_t*)xcalloc((size_t)3 * s->pairs_cap, sizeof(int64_t));
    s->collist = (double*)xcalloc((size_t)20 * s->pairs_cap, 8);
    s->contact_cap = 1024;
    s->contact = (double*)xcalloc((size_t)12 * s->contact_cap, 8);
    s->remove_cap = 1024;
    s->removepart = (int*)xcalloc(s->remove_cap, sizeof(int));
    /* the head array spans ny_max halo cells on each side (partlib.f90:159) */
    s->hy0 = 1 - s->nx_max;
    s->hny = s->ncd + 2 * s->nx_max;
    s->head = (int*)xcalloc((size_t)s->ncd * s->hny * s->hny, sizeof(int));
    *out = s;
    return 0;
}

void orc_destroy(orc_sim* s)
{
    if (!s) return;
    free(s->vx); free(s->vy); free(s->vz); free(s->ox); free(s->oy); free(s->oz);
    free(s->wx); free(s->wy); free(s->wz); free(s->vxc); free(s->vyc); free(s->vzc);
    free(s->yp); free(s->vp); free(s->dmove); free(s->fvp); free(s->fdp); free(s->pertvel);
    free(s->pindex); free(s->ypn); free(s->vpn); free(s->list); free(s->head);
    free(s->nei_num); free(s->nei_data);
    free(s->nirbin0); free(s->wrbin0); free(s->wxbin0); free(s->wybin0); free(s->wzbin0);
    free(s->urbin0); free(s->thetabin0); free(s->wrtan0);
    free(s->pairs); free(s->collist); free(s->removepart); free(s->contact);
    free(s);
}

static void zero_pads_spectral(int N, double* v)
{
    const int L = N + 2;
    int j, k;
    for (k = 0; k < N; ++k)
        for (j = 0; j < L; ++j) { v[IDX(N, j, k, L)] = 0.0; v[IDX(N + 1, j, k, L)] = 0.0; }
}

int orc_set_flow(orc_sim* s, const double* vx, const double* vy, const double* vz,
                 const double* ox, const double* oy, const double* oz)
{
    const double* src[6] = {vx, vy, vz, ox, oy, oz};
    double* dst[6] = {s->vx, s->vy, s->vz, s->ox, s->oy, s->oz};
    int f;
    for (f = 0; f < 6; ++f) {
        if (!src[f]) continue;
        memcpy(dst[f], src[f], s->fsz * 8);
        zero_pads_spectral(s->N, dst[f]);
    }
    s->have_phys = 0;
    return 0;
}

/* SplitMix64 (Steele, Lea and Flood, OOPSLA 2014) as a counter-based generator: output
 * i of the stream with key s is the finaliser of s + (i + 1) * 0x9E3779B97F4A7C15.  With
 * key 0 the outputs are the published stream 0xe220a8397b1dcdaf, 0x6e789e6aa1b965f4,
 * 0x06c45d188009454f, ...  Replaces IFPORT rand in initial_flow (flowlib.f90:158-172). */
uint64_t orc_splitmix64(uint64_t key, uint64_t i)
{
    uint64_t z = key + (i + 1u) * 0x9E3779B97F4A7C15ull;
    z = (z ^ (z >> 30)) * 0xBF58476D1CE4E5B9ull;
    z = (z ^ (z >> 27)) * 0x94D049BB133111EBull;
    return z ^ (z >> 31);
}

static void init_gauss(uint64_t key, int N, int c, int kz, int ky, int kx, double* re, double* im)
{
    const uint64_t m = (((uint64_t)c * N + kz) * N + ky) * (uint64_t)(N / 2) + kx;
    double t1 = (double)(orc_splitmix64(key, 2 * m) >> 11) * (1.0 / 9007199254740992.0);
    double t2 = (double)(orc_splitmix64(key, 2 * m + 1) >> 11) * (1.0 / 9007199254740992.0);
    double a, ph;
    if (t1 <= 1.e-10) t1 = 1.e-10;
    if (t2 <= 1.e-10) t2 = 1.e-10;
    a = sqrt(-2.0 * log(t1));
    ph = 2.0 * 3.14159265358979323846 * t2;
    *re = a * cos(ph);
    *im = a * sin(ph);
}

/* initial_flow (flowlib.f90:53-144) with the counter-based generator: Gaussian modes
 * (one pair of uniforms per component and mode), the kx = 0 plane made Hermitian as
 * (g(k) + conj g(-k)) / 2, the Nyquist planes and the mean zero, the projection (94-101),
 * the shell factors (107-137: the k^4 exp(-2 (k/k0)^2) spectrum at total energy
 * 3/2 urms^2, or initial_flow's own force(1) / k^(-5/3) scaling) and the dealiasing
 * (139-141); the history starts equal to the velocity (turb.f90:452-454).  kind < 0
 * gives the flow at rest. */
int orc_init_flow(orc_sim* s, int kind, int seed, double k0, double urms)
{
    const int N = s->N, L = s->L, M = N / 2, nek = s->nek;
    const uint64_t key = (uint64_t)(int64_t)seed;
    double* v[3] = {s->vx, s->vy, s->vz};
    double *e_t = (double*)xcalloc((size_t)nek + 1, 8), *cnt = (double*)xcalloc((size_t)nek + 1, 8);
    double *fac = (double*)xcalloc((size_t)nek + 1, 8);
    const double pi = 4. * atan(1.);
    int c, kz, ky, kx, i;
    for (c = 0; c < 3; ++c) memset(v[c], 0, s->fsz * 8);
    if (kind >= 0) {
        for (kz = 0; kz < N; ++kz)
            for (ky = 0; ky < N; ++ky)
                for (kx = 0; kx < M; ++kx) {
                    const int zero = ky == N / 2 || kz == N / 2 || (kx == 0 && ky == 0 && kz == 0);
                    const double k[3] = {(double)kx, (double)(ky <= N / 2 ? ky : ky - N), (double)(kz <= N / 2 ? kz : kz - N)};
                    double gr[3], gi[3], k2, pr, pim, e = 0.0;
                    int ik2;
                    for (c = 0; c < 3; ++c) {
                        init_gauss(key, N, c, kz, ky, kx, &gr[c], &gi[c]);
                        if (kx == 0) {
                            double hr, hi;
                            init_gauss(key, N, c, (N - kz) % N, (N - ky) % N, 0, &hr, &hi);
                            gr[c] = 0.5 * (gr[c] + hr);
                            gi[c] = 0.5 * (gi[c] - hi);
                        }
                        if (zero) gr[c] = gi[c] = 0.0;
                    }
                    k2 = k[0] * k[0] + k[1] * k[1] + k[2] * k[2];
                    if (kx == 0 && ky == 0 && kz == 0) k2 = 1.e-5;
                    pr = (k[0] * gr[0] + k[1] * gr[1] + k[2] * gr[2]) / k2;
                    pim = (k[0] * gi[0] + k[1] * gi[1] + k[2] * gi[2]) / k2;
                    for (c = 0; c < 3; ++c) {
                        const double x = gr[c] - k[c] * pr, y = gi[c] - k[c] * pim;
                        v[c][IDX(kz, 2 * kx, ky, L)] = x;
                        v[c][IDX(kz, 2 * kx + 1, ky, L)] = y;
                        e += x * x + y * y;
                    }
                    ik2 = (int)(sqrt(k2) + 0.5);
                    if (ik2 < 1 || ik2 > nek) continue;
                    e_t[ik2] += kx == 0 ? 0.5 * e : e;
                    cnt[ik2] += kx == 0 ? 0.5 : 1.0;
                }
        if (kind == DDNS_SPEC_K4_GAUSS) {
            double sum = 0.0;
            for (i = 1; i <= nek; ++i) {
                const double q = (double)i / k0;
                fac[i] = q * q * q * q * exp(-2.0 * q * q);
                sum += fac[i];
            }
            for (i = 1; i <= nek; ++i) {
                const double tg = fac[i] * 1.5 * urms * urms / sum;
                fac[i] = e_t[i] > 0 ? sqrt(tg / e_t[i]) : 0.0;
            }
        } else {
            const double ccc = s->p.force[0] * 4. / 3. * pi * (1.5 * 1.5 * 1.5 - 0.5 * 0.5 * 0.5) / cnt[1];
            for (i = 1; i <= nek; ++i) {
                const double vol = 4. / 3. * pi * ((i + 0.5) * (i + 0.5) * (i + 0.5) - (i - 0.5) * (i - 0.5) * (i - 0.5));
                fac[i] = e_t[i] > 0 ? sqrt(ccc / e_t[i] * cnt[i] / vol) / pow((double)i, 0.833333) : 0.0;
            }
        }
        for (kz = 0; kz < N; ++kz)
            for (ky = 0; ky < N; ++ky)
                for (kx = 0; kx < M; ++kx) {
                    const double k1 = kx, k2y = ky <= N / 2 ? ky : ky - N, k3 = kz <= N / 2 ? kz : kz - N;
                    double k2 = k1 * k1 + k2y * k2y + k3 * k3, f;
                    int ik2;
                    if (k2 == 0.0) k2 = 1.e-5;
                    ik2 = (int)(sqrt(k2) + 0.5);
                    if (ik2 < 1) continue;
                    f = ik2 <= nek ? fac[ik2] : 0.0;
                    for (c = 0; c < 3; ++c) {
                        v[c][IDX(kz, 2 * kx, ky, L)] *= f;
                        v[c][IDX(kz, 2 * kx + 1, ky, L)] *= f;
                    }
                }
    }
    memcpy(s->ox, s->vx, s->fsz * 8);
    memcpy(s->oy, s->vy, s->fsz * 8);
    memcpy(s->oz, s->vz, s->fsz * 8);
    free(e_t); free(cnt); free(fac);
    s->have_phys = 0;
    return 0;
}

int orc_get_flow(orc_sim* s, double* vx, double* vy, double* vz, double* ox, double* oy, double* oz)
{
    double* dst[6] = {vx, vy, vz, ox, oy, oz};
    const double* src[6] = {s->vx, s->vy, s->vz, s->ox, s->oy, s->oz};
    int f;
    for (f = 0; f < 6; ++f) if (dst[f]) memcpy(dst[f], src[f], s->fsz * 8);
    return 0;
}

int orc_set_forcing_state(orc_sim* s, const double* b, int iseedf, int iyf, const int* ivf)
{
    int i;
    if (b) { memcpy(s->b1r, b, 150 * 8); memcpy(s->b2r, b + 150, 150 * 8); memcpy(s->b3r, b + 300, 150 * 8); }
    s->iseedf = iseedf; s->iyf = iyf;
    for (i = 0; i < NTAB; ++i) s->ivf[i] = ivf ? ivf[i] : 0;
    return 0;
}

int orc_get_forcing_state(orc_sim* s, double* b, int* iseedf, int* iyf, int* ivf)
{
    int i;
    if (b) { memcpy(b, s->b1r, 150 * 8); memcpy(b + 150, s->b2r, 150 * 8); memcpy(b + 300, s->b3r, 150 * 8); }
    if (iseedf) *iseedf = s->iseedf;
    if (iyf) *iyf = s->iyf;
    if (ivf) for (i = 0; i < NTAB; ++i) ivf[i] = s->ivf[i];
    return 0;
}

int orc_set_droplets(orc_sim* s, int n, const double* yp, const double* fvp,
                     const double* fdp, const double* pertvel)
{
    if (n < 0 || n > s->cap) return DDNS_ERR_CAPACITY;
    s->nps = n;
    memset(s->yp, 0, (size_t)8 * s->cap * 8);
    memset(s->fvp, 0, (size_t)9 * s->cap * 8);
    memset(s->fdp, 0, (size_t)12 * s->cap * 8);
    memset(s->pertvel, 0, (size_t)3 * s->cap * 8);
    if (yp) memcpy(s->yp, yp, (size_t)8 * n * 8);
    if (fvp) memcpy(s->fvp, fvp, (size_t)9 * n * 8);
    if (fdp) memcpy(s->fdp, fdp, (size_t)12 * n * 8);
    if (pertvel) memcpy(s->pertvel, pertvel, (size_t)3 * n * 8);
    return 0;
}

int orc_get_droplets(orc_sim* s, int* n, double* yp, double* fvp, double* fdp,
                     double* pertvel, double* vp, double* dmove)
{
    int m = s->nps;
    if (n) *n = m;
    if (yp) memcpy(yp, s->yp, (size_t)8 * m * 8);
    if (fvp) memcpy(fvp, s->fvp, (size_t)9 * m * 8);
    if (fdp) memcpy(fdp, s->fdp, (size_t)12 * m * 8);
    if (pertvel) memcpy(pertvel, s->pertvel, (size_t)3 * m * 8);
    if (vp) memcpy(vp, s->vp, (size_t)3 * m * 8);
    if (dmove) memcpy(dmove, s->dmove, (size_t)3 * m * 8);
    return 0;
}

int orc_set_new_particle(orc_sim* s, int flag) { s->new_particle = flag; return 0; }

/* OpenMP threads the oracle's parallel loops use (1 without OpenMP). */
int orc_num_threads(void)
{
#ifdef _OPENMP
    return omp_get_max_threads();
#else
    return 1;
#endif
}
int orc_set_istep(orc_sim* s, int istep) { s->istep = istep; return 0; }
int orc_flow_only(orc_sim* s, int on) { s->flow_only = on; return 0; }

int orc_get_relocate_state(orc_sim* s, int* iseedr, int* iyr, int* ivr, int* idum2)
{
    int i;
    *iseedr = s->iseedr; *iyr = s->iyr; *idum2 = s->idum2;
    for (i = 0; i < NTAB; ++i) ivr[i] = s->ivr[i];
    return 0;
}

/* ------------------------------------------------------------------------------- */
/* supforSTH, flowlib.f90:378-562.  Pinned: ranff bit-exact against the reference's */
/* accesories.f90; forcing solenoidal, Hermitian on kx = 0, zero outside |k|<sqrt8. */
/* a(i,j,k) is a1r(6,5,5) column-major: i = kx re/im, j = ky slot, k = kz slot.     */
/* ------------------------------------------------------------------------------- */

#define A3(i, j, k) (((i) - 1) + 6 * (((j) - 1) + 5 * ((k) - 1)))

int orc_supfor_sth(orc_sim* s, double* ux, double* uy, double* uz, double* a_out)
{
    const int N = s->N, L = s->L;
    const double dt = s->p.dt, pi2 = s->pi2;
    const double smallx = 1.0e-18, xkf = sqrt(8.0);
    const double var = s->p.ou_var, tf = s->p.ou_tf;
    double fr[4][150];
    double a[3][150];
    double kxr[6], kyr[5], kzr[5];
    int mc, i, j, k;
    /* SETWAVERAND for the 6 x 5 x 5 forced block, flowlib.f90:353-372 */
    for (i = 1; i <= 6; i += 2) { kxr[i - 1] = (double)(i / 2); kxr[i] = kxr[i - 1]; }
    for (j = 1; j <= 5; ++j) {
        int jj = (j > 3) ? j + N - 5 : j;
        kyr[j - 1] = (jj < N / 2 + 2) ? (double)(jj - 1) : (double)(-(N + 1 - jj));
        kzr[j - 1] = kyr[j - 1];
    }
    /* white noise, flowlib.f90:405-419 */
    for (mc = 1; mc <= 3; ++mc)
        for (k = 1; k <= 5; ++k)
            for (j = 1; j <= 5; ++j)
                for (i = 1; i <= 6; i += 2) {
                    double cont1 = orc_ranff(&s->iseedf, &s->iyf, s->ivf);
                    double rmag, rph;
                    if (cont1 < 1.e-6) cont1 = 1.e-6;
                    rmag = sqrt(-4.0 * var * dt * log(cont1) / tf);
                    rph = orc_ranff(&s->iseedf, &s->iyf, s->ivf);
                    fr[mc][A3(i, j, k)] = rmag * cos(pi2 * rph);
                    fr[mc][A3(i + 1, j, k)] = rmag * sin(pi2 * rph);
                }
    /* Hermitian fix of the noise on kx = 0, flowlib.f90:424-440 */
    for (mc = 1; mc <= 3; ++mc) {
        for (k = 1; k <= 5; ++k) {
            int kk = (k == 1) ? 1 : 7 - k;
            for (j = 2; j <= 3; ++j) {
                int jj = 7 - j;
                fr[mc][A3(1, jj, kk)] = fr[mc][A3(1, j, k)];
                fr[mc][A3(2, jj, kk)] = -fr[mc][A3(2, j, k)];
            }
        }
        for (k = 2; k <= 3; ++k) {
            int kk = 7 - k;
            fr[mc][A3(1, 1, kk)] = fr[mc][A3(1, 1, k)];
            fr[mc][A3(2, 1, kk)] = -fr[mc][A3(2, 1, k)];
        }
        fr[mc][A3(1, 1, 1)] = 0.0; fr[mc][A3(2, 1, 1)] = 0.0;
    }
    /* OU step and solenoidal projection, flowlib.f90:447-478 */
    for (k = 1; k <= 5; ++k)
        for (j = 1; j <= 5; ++j)
            for (i = 1; i <= 6; ++i) {
                int o = A3(i, j, k);
                double kx_ = kxr[i - 1], ky_ = kyr[j - 1], kz_ = kzr[k - 1];
                double radm = sqrt(kx_ * kx_ + ky_ * ky_ + kz_ * kz_);
                if (radm < xkf) {
                    double t1, t2, t3, divb;
                    s->b1r[o] = s->b1r[o] * (1.0 - dt / tf) + fr[1][o];
                    s->b2r[o] = s->b2r[o] * (1.0 - dt / tf) + fr[2][o];
                    s->b3r[o] = s->b3r[o] * (1.0 - dt / tf) + fr[3][o];
                    t1 = kx_ * s->b1r[o];
                    t2 = ky_ * s->b2r[o];
                    t3 = kz_ * s->b3r[o];
                    divb = (t1 + t2 + t3) / ((radm + smallx) * (radm + smallx));
                    if (radm == 0.) divb = 0.0;
                    a[0][o] = s->b1r[o] - kx_ * divb;
                    a[1][o] = s->b2r[o] - ky_ * divb;
                    a[2][o] = s->b3r[o] - kz_ * divb;
                } else {
                    a[0][o] = 0.0; a[1][o] = 0.0; a[2][o] = 0.0;
                }
            }
    for (mc = 0; mc < 3; ++mc) { a[mc][A3(1, 1, 1)] = 0.0; a[mc][A3(2, 1, 1)] = 0.0; }   /* 481-483 */
    for (k = 1; k <= 5; ++k) {                                                              /* 488-500 */
        int kk = (k == 1) ? 1 : 7 - k;
        for (j = 2; j <= 3; ++j) {
            int jj = 7 - j;
            for (mc = 0; mc < 3; ++mc) {
                a[mc][A3(1, jj, kk)] = a[mc][A3(1, j, k)];
                a[mc][A3(2, jj, kk)] = -a[mc][A3(2, j, k)];
            }
        }
    }
    for (k = 2; k <= 3; ++k) {                                                              /* 502-510 */
        int kk = 7 - k;
        for (mc = 0; mc < 3; ++mc) {
            a[mc][A3(1, 1, kk)] = a[mc][A3(1, 1, k)];
            a[mc][A3(2, 1, kk)] = -a[mc][A3(2, 1, k)];
        }
    }
    if (a_out) { memcpy(a_out, a[0], 150 * 8); memcpy(a_out + 150, a[1], 150 * 8); memcpy(a_out + 300, a[2], 150 * 8); }
    /* Euler add, flowlib.f90:529-559: u(kz slot, kx re/im, ky slot) */
    if (ux) {
        for (k = 1; k <= 5; ++k) {
            int kk = (k > 3) ? k + N - 5 : k;
            for (j = 1; j <= 5; ++j) {
                int jj = (j > 3) ? j + N - 5 : j;
                for (i = 1; i <= 6; ++i) {
                    size_t o = IDX(kk - 1, i - 1, jj - 1, L);
                    ux[o] = ux[o] + dt * a[0][A3(i, j, k)];
                    uy[o] = uy[o] + dt * a[1][A3(i, j, k)];
                    uz[o] = uz[o] + dt * a[2][A3(i, j, k)];
                }
            }
        }
    }
    return 0;
}

/* ------------------------------------------------------------------------------- */
/* Flow step, turb.f90:506-691 (the part before the droplet block and after it).    */
/* Pinned end to end by: Beltrami (ABC) decay closed form, Taylor-Green first       */
/* nonlinear term, solenoidality, supforDET energies, Parseval of E(k).             */
/* ------------------------------------------------------------------------------- */

/* curl w = i k x v and the copy vc = v, then the six inverse FFTs (turb.f90:506-539). */
static void flow_first_half(orc_sim* s)
{
    const int N = s->N, L = s->L;
    int k;
#pragma omp parallel for schedule(static)
    for (k = 0; k < N; ++k) {
        int i, j;
        for (j = 0; j < L; ++j)
            for (i = 0; i < L; ++i) {
                size_t o = IDX(i, j, k, L);
                double kx_ = (i >= N) ? 0.0 : wn_x(j);
                double ky_ = (i >= N) ? 0.0 : wn_y(k, N);
                double kz_ = wn_z(i, N);
                s->wx[o] = ky_ * s->vz[o] - kz_ * s->vy[o];   /* turb.f90:506-508 */
                s->wy[o] = kz_ * s->vx[o] - kx_ * s->vz[o];
                s->wz[o] = kx_ * s->vy[o] - ky_ * s->vx[o];
            }
        /* multiply by i over rows 1:lx (turb.f90:509-517) */
        for (j = 0; j < L; j += 2)
            for (i = 0; i < N; ++i) {
                size_t o0 = IDX(i, j, k, L), o1 = IDX(i, j + 1, k, L);
                double t;
                t = s->wx[o0]; s->wx[o0] = -s->wx[o1]; s->wx[o1] = t;
                t = s->wy[o0]; s->wy[o0] = -s->wy[o1]; s->wy[o1] = t;
                t = s->wz[o0]; s->wz[o0] = -s->wz[o1]; s->wz[o1] = t;
            }
    }
    memcpy(s->vxc, s->vx, s->fsz * 8);                      /* turb.f90:519-521 */
    memcpy(s->vyc, s->vy, s->fsz * 8);
    memcpy(s->vzc, s->vz, s->fsz * 8);
    orc_fft_c2r(N, s->wx); orc_fft_c2r(N, s->wy); orc_fft_c2r(N, s->wz);   /* 524-526 */
    orc_fft_c2r(N, s->vx); orc_fft_c2r(N, s->vy); orc_fft_c2r(N, s->vz);   /* 537-539 */
    s->have_phys = 1;
}

static void flow_second_half(orc_sim* s)
{
    const int N = s->N, L = s->L;
    const double dt = s->p.dt, rnu = s->p.rnu;
    const double dtr = 1.5 * dt, dt_h = 0.5 * dt;           /* turb.f90:483-484 */
    int k;
    /* u x w in physical space, turb.f90:634-637; then restore v = vc (639-641) */
#pragma omp parallel for schedule(static)
    for (k = 0; k < N; ++k) {
        int i, j;
        for (j = 0; j < L; ++j)
            for (i = 0; i < L; ++i) {
                size_t o = IDX(i, j, k, L);
                double ux = s->vx[o], uy = s->vy[o], uz = s->vz[o];
                double wx = s->wx[o], wy = s->wy[o], wz = s->wz[o];
                double tmp = uy * wz - uz * wy;
                s->wy[o] = ux * wy - uy * wx;
                s->wz[o] = uz * wx - ux * wz;
                s->wx[o] = tmp;
            }
    }
    /* keep the physical velocity for the droplet step / getter */
    {
        double* t;
        t = s->vx; s->vx = s->vxc; s->vxc = t;
        t = s->vy; s->vy = s->vyc; s->vyc = t;
        t = s->vz; s->vz = s->vzc; s->vzc = t;
    }
    orc_fft_r2c(N, s->wx); orc_fft_r2c(N, s->wz); orc_fft_r2c(N, s->wy);   /* 644-646 */
    orc_dealias(N, s->wx); orc_dealias(N, s->wy); orc_dealias(N, s->wz);   /* 648-650 */
    /* AB2 / Crank-Nicolson explicit part, turb.f90:662-664 */
#pragma omp parallel for schedule(static)
    for (k = 0; k < N; ++k) {
        int i, j;
        for (j = 0; j < L; ++j)
            for (i = 0; i < L; ++i) {
                size_t o = IDX(i, j, k, L);
                double k2 = wn_k2(i, j, k, N);
                s->vx[o] = s->vx[o] * (1. - rnu * dt_h * k2) + dtr * s->wx[o] - dt_h * s->ox[o];
                s->vy[o] = s->vy[o] * (1. - rnu * dt_h * k2) + dtr * s->wz[o] - dt_h * s->oy[o];
                s->vz[o] = s->vz[o] * (1. - rnu * dt_h * k2) + dtr * s->wy[o] - dt_h * s->oz[o];
            }
    }
    if (s->p.forcing == 2) orc_supfor_sth(s, s->vx, s->vy, s->vz, NULL);   /* 666-667 */
    /* projection and implicit viscous part, turb.f90:669-676 */
#pragma omp parallel for schedule(static)
    for (k = 0; k < N; ++k) {
        int i, j;
        for (j = 0; j < L; ++j)
            for (i = 0; i < L; ++i) {
                size_t o = IDX(i, j, k, L);
                double k2 = wn_k2(i, j, k, N);
                double kx_ = (i >= N) ? 0.0 : wn_x(j);
                double ky_ = (i >= N) ? 0.0 : wn_y(k, N);
                double kz_ = wn_z(i, N);
                double tmp = (kx_ * s->vx[o] + ky_ * s->vy[o] + kz_ * s->vz[o]) / k2;
                double vx = s->vx[o] - kx_ * tmp;
                double vy = s->vy[o] - ky_ * tmp;
                double vz = s->vz[o] - kz_ * tmp;
                s->vx[o] = vx / (1. + rnu * dt_h * k2);
                s->vy[o] = vy / (1. + rnu * dt_h * k2);
                s->vz[o] = vz / (1. + rnu * dt_h * k2);
            }
    }
    if (s->p.forcing == 1) orc_supfor_det(N, s->vx, s->vy, s->vz, s->p.force, NULL);   /* 678 */
    orc_symmetrize(N, s->vx); orc_symmetrize(N, s->vy); orc_symmetrize(N, s->vz);      /* 681-683 */
    orc_dealias(N, s->vx); orc_dealias(N, s->vy); orc_dealias(N, s->vz);               /* 685-687 */
    memcpy(s->ox, s->wx, s->fsz * 8);                        /* turb.f90:689-691 */
    memcpy(s->oy, s->wz, s->fsz * 8);
    memcpy(s->oz, s->wy, s->fsz * 8);
}

int orc_get_physical_velocity(orc_sim* s, double* ux, double* uy, double* uz)
{
    const int N = s->N, L = s->L;
    double* src[3] = {s->vxc, s->vyc, s->vzc};
    double* dst[3] = {ux, uy, uz};
    int c, i, j, k;
    if (!s->have_phys) return DDNS_ERR_ARG;
    for (c = 0; c < 3; ++c)
        if (dst[c])
            for (k = 0; k < N; ++k)
                for (j = 0; j < N; ++j)
                    for (i = 0; i < N; ++i)
                        dst[c][(size_t)i + (size_t)N * ((size_t)j + (size_t)N * k)] = src[c][IDX(i, j, k, L)];
    return 0;
}

/* Energy spectrum, postprocessing.f90:151-170 (summed in array order).  Pinned by
 * Parseval: sum_i E(i) plus the truncated shells equals <|u|^2>/2 in physical space. */
int orc_get_spectrum(orc_sim* s, double* ek)
{
    const int N = s->N, L = s->L;
    int sh, i, j, k;
    for (sh = 0; sh < s->nek; ++sh) ek[sh] = 0.0;
    for (sh = 1; sh <= s->nek; ++sh) {
        double e = 0.0;
        for (k = 0; k < N; ++k)
            for (j = 0; j < L; ++j)
                for (i = 0; i < N; ++i) {
                    size_t o = IDX(i, j, k, L);
                    double t;
                    if (wn_ik2(i, j, k, N) != sh) continue;
                    t = s->vx[o] * s->vx[o] + s->vy[o] * s->vy[o] + s->vz[o] * s->vz[o];
                    if (j <= 1) t = 0.5 * t;
                    e += t;
                }
        ek[sh - 1] = e;
    }
    return 0;
}

/* ------------------------------------------------------------------------------- */
/* Flow statistics: POSTPROC (postprocessing.f90:76-272) and TRANSPOSTPROC         */
/* (275-315) of the current state, on scratch copies.                               */
/* Pinned by: single Fourier modes (skewness 0, flatness 3/2, E, D, CFL in closed   */
/* form), sum_k T(k) = <u . (u x w)> = 0 for a dealiased field, a solenoidal field   */
/* with zero divergence statistics.                                                  */
/* ------------------------------------------------------------------------------- */

/* multiply a spectral field by i over the rows 1:lx (postprocessing.f90:100-111) */
static void mul_i(int N, int L, double* v)
{
    int i, j, k;
    for (k = 0; k < N; ++k)
        for (j = 0; j < L; j += 2)
            for (i = 0; i < N; ++i) {
                size_t o0 = IDX(i, j, k, L), o1 = IDX(i, j + 1, k, L);
                double t = v[o0];
                v[o0] = -v[o1];
                v[o1] = t;
            }
}

/* the three wavenumbers of entry (i, j, k) of the spectral layout */
static void wn3(int i, int j, int k, int N, double* kx_, double* ky_, double* kz_)
{
    *kx_ = (i >= N) ? 0.0 : wn_x(j);
    *ky_ = (i >= N) ? 0.0 : wn_y(k, N);
    *kz_ = wn_z(i, N);
}

/* physical velocity and vorticity of the state (turb.f90:506-539) into u[3], w[3] */
static void phys_u_w(orc_sim* s, double** u, double** w)
{
    const int N = s->N, L = s->L;
    int i, j, k, c;
    const double* v[3] = {s->vx, s->vy, s->vz};
    for (c = 0; c < 3; ++c) memcpy(u[c], v[c], s->fsz * 8);
    for (k = 0; k < N; ++k)
        for (j = 0; j < L; ++j)
            for (i = 0; i < L; ++i) {
                size_t o = IDX(i, j, k, L);
                double kx_, ky_, kz_;
                wn3(i, j, k, N, &kx_, &ky_, &kz_);
                w[0][o] = ky_ * s->vz[o] - kz_ * s->vy[o];
                w[1][o] = kz_ * s->vx[o] - kx_ * s->vz[o];
                w[2][o] = kx_ * s->vy[o] - ky_ * s->vx[o];
            }
    for (c = 0; c < 3; ++c) {
        mul_i(N, L, w[c]);
        orc_fft_c2r(N, w[c]);
        orc_fft_c2r(N, u[c]);
    }
}

int orc_get_spectra(orc_sim* s, double* ek, double* dk, double* tk)
{
    const int N = s->N, L = s->L, nek = s->nek;
    const double rnu = s->p.rnu;
    double* F[3] = {NULL, NULL, NULL};
    int sh, i, j, k, c;
    if (tk) {
        /* the dealiased nonlinear term of this state (turb.f90:634-650) */
        double *u[3], *w[3];
        for (c = 0; c < 3; ++c) { u[c] = (double*)xcalloc(s->fsz, 8); w[c] = (double*)xcalloc(s->fsz, 8); }
        phys_u_w(s, u, w);
        for (k = 0; k < N; ++k)
            for (j = 0; j < L; ++j)
                for (i = 0; i < L; ++i) {
                    size_t o = IDX(i, j, k, L);
                    double ux = u[0][o], uy = u[1][o], uz = u[2][o];
                    double wx = w[0][o], wy = w[1][o], wz = w[2][o];
                    w[0][o] = uy * wz - uz * wy;
                    w[1][o] = uz * wx - ux * wz;
                    w[2][o] = ux * wy - uy * wx;
                }
        for (c = 0; c < 3; ++c) {
            orc_fft_r2c(N, w[c]);
            orc_dealias(N, w[c]);
            F[c] = w[c];
            free(u[c]);
        }
    }
    for (sh = 1; sh <= nek; ++sh) {                            /* 143-170, 285-313 */
        double e = 0.0, d = 0.0, t = 0.0;
        for (k = 0; k < N; ++k)
            for (j = 0; j < L; ++j)
                for (i = 0; i < N; ++i) {
                    size_t o = IDX(i, j, k, L);
                    double ee;
                    if (wn_ik2(i, j, k, N) != sh) continue;
                    ee = s->vx[o] * s->vx[o] + s->vy[o] * s->vy[o] + s->vz[o] * s->vz[o];
                    if (j <= 1) ee = 0.5 * ee;
                    e += ee;
                    d += 2.0 * rnu * ee * wn_k2(i, j, k, N);
                    if (F[0]) {
                        double tt = s->vx[o] * F[0][o] + s->vy[o] * F[1][o] + s->vz[o] * F[2][o];
                        if (j <= 1) tt = 0.5 * tt;
                        t += tt;
                    }
                }
        if (ek) ek[sh - 1] = e;
        if (dk) dk[sh - 1] = d;
        if (tk) tk[sh - 1] = t;
    }
    for (c = 0; c < 3; ++c) free(F[c]);
    return 0;
}

int orc_get_monitors(orc_sim* s, int divstat, ddns_monitors* m)
{
    const int N = s->N, L = s->L, nek = s->nek;
    const double rnu = s->p.rnu, n3 = (double)N * N * N;
    const double pi2 = 8.0 * atan(1.0);
    double *u[3], *w[3], *d;
    double a2 = 0.0, a3 = 0.0, a4 = 0.0, pmax = 0.0;
    double qq = 0.0, dissp = 0.0, xintls = 0.0;
    double* ek = (double*)xcalloc(nek, 8);
    double* dk = (double*)xcalloc(nek, 8);
    const double* v[3] = {s->vx, s->vy, s->vz};
    int i, j, k, c;
    memset(m, 0, sizeof(*m));
    for (c = 0; c < 3; ++c) { u[c] = (double*)xcalloc(s->fsz, 8); w[c] = (double*)xcalloc(s->fsz, 8); }
    d = (double*)xcalloc(s->fsz, 8);
    /* longitudinal derivatives d_c = d u_c / d x_c (97-111) and their moments (113-137) */
    for (c = 0; c < 3; ++c) {
        for (k = 0; k < N; ++k)
            for (j = 0; j < L; ++j)
                for (i = 0; i < L; ++i) {
                    size_t o = IDX(i, j, k, L);
                    double kk[3];
                    wn3(i, j, k, N, &kk[0], &kk[1], &kk[2]);
                    d[o] = kk[c] * v[c][o];
                }
        mul_i(N, L, d);
        orc_fft_c2r(N, d);
        for (k = 0; k < N; ++k)
            for (j = 0; j < N; ++j)
                for (i = 0; i < N; ++i) {
                    double x = d[IDX(i, j, k, L)], x2 = x * x;
                    a2 += x2; a3 += x2 * x; a4 += x2 * x2;
                }
    }
    a2 /= 3.0; a3 /= 3.0; a4 /= 3.0;
    m->vskew = a3 / n3 / pow(a2 / n3, 1.5);
    m->vflat = a4 / n3 / pow(a2 / n3, 2.0);
    orc_get_spectra(s, ek, dk, NULL);                          /* 143-170 */
    for (i = 1; i <= nek; ++i) {
        qq += ek[i - 1];
        dissp += dk[i - 1];
        xintls += ek[i - 1] / (double)i;
    }
    m->qq = qq;
    m->dissp = dissp;
    m->eta = pow(rnu * rnu * rnu / dissp, 0.25);               /* 190-202 */
    m->tk = sqrt(rnu / dissp);
    m->uprime = sqrt(2.0 * qq / 3.0);
    m->tmse = sqrt(15.0 * rnu * m->uprime * m->uprime / dissp);
    m->ree = m->uprime * m->tmse / rnu;
    m->xl = m->uprime * m->uprime * m->uprime / dissp;
    m->et = m->xl / m->uprime;
    m->resol = ((double)(N / 2) - 1.5) * m->eta;               /* akmax, turb.f90:303 */
    m->xintls = xintls * (pi2 / 2.0) / 2.0 / (m->uprime * m->uprime);
    phys_u_w(s, u, w);
    for (k = 0; k < N; ++k)                                    /* 204-207 */
        for (j = 0; j < N; ++j)
            for (i = 0; i < N; ++i) {
                size_t o = IDX(i, j, k, L);
                double t = fabs(u[0][o]) + fabs(u[1][o]) + fabs(u[2][o]);
                if (t > pmax) pmax = t;
            }
    m->pmax = pmax;
    m->cfl = pmax * s->p.dt / (pi2 / (double)N);
    if (divstat) {                                             /* 217-266 */
        double s1 = 0.0, s2 = 0.0, g1 = 0.0, g2 = 0.0, gd = 0.0;
        for (k = 0; k < N; ++k)
            for (j = 0; j < L; ++j)
                for (i = 0; i < L; ++i) {
                    size_t o = IDX(i, j, k, L);
                    double kx_, ky_, kz_;
                    wn3(i, j, k, N, &kx_, &ky_, &kz_);
                    d[o] = kx_ * s->vx[o] + ky_ * s->vy[o] + kz_ * s->vz[o];
                }
        mul_i(N, L, d);
        orc_fft_c2r(N, d);
        for (k = 0; k < N; ++k)
            for (j = 0; j < N; ++j)
                for (i = 0; i < N; ++i) {
                    size_t o = IDX(i, j, k, L);
                    double dv = d[o];
                    double g = sqrt(w[0][o] * w[0][o] + w[1][o] * w[1][o] + w[2][o] * w[2][o]);
                    s1 += dv; s2 += dv * dv;
                    g1 += g; g2 += g * g; gd += g * dv;
                }
        m->vdiv1 = s1 / n3;
        m->vdiv2 = sqrt((s2 - m->vdiv1 * m->vdiv1) / n3);
        m->vort1 = g1 / n3;
        m->vort2 = sqrt((g2 - m->vort1 * m->vort1) / n3);
        m->divvt = (gd - m->vdiv1 * m->vort1) / (2.0 * n3);
    }
    for (c = 0; c < 3; ++c) { free(u[c]); free(w[c]); }
    free(d); free(ek); free(dk);
    return 0;
}

/* ------------------------------------------------------------------------------- */
/* Droplets                                                                          */
/* ------------------------------------------------------------------------------- */

#define YP(c, ih) s->yp[(size_t)((ih) - 1) * 8 + ((c) - 1)]
#define VP(c, ih) s->vp[(size_t)((ih) - 1) * 3 + ((c) - 1)]
#define DM(c, ih) s->dmove[(size_t)((ih) - 1) * 3 + ((c) - 1)]
#define FVP(c, t, ih) s->fvp[(size_t)((ih) - 1) * 9 + (size_t)(t) * 3 + ((c) - 1)]
#define FDP(c, t, ih) s->fdp[(size_t)((ih) - 1) * 12 + (size_t)(t) * 3 + ((c) - 1)]
#define PV(c, ih) s->pertvel[(size_t)((ih) - 1) * 3 + ((c) - 1)]
#define YPN(r, ih) s->ypn[(size_t)((ih) - 1) * s->ypn_rows + ((r) - 1)]
#define VPN(c, ih) s->vpn[(size_t)((ih) - 1) * 3 + ((c) - 1)]

/* Lagrange weights, PREPARE_INTERP partlib.f90:324-338.  Pinned: weights sum to 1,
 * reproduce polynomials of degree <= 5, equal e_3 at z = 0. */
int orc_interp_weights(double y, double dx, int* node, double* bg)
{
    int nd = (int)floor(y / dx);
    double z = (y - (double)nd * dx) / dx;
    double z2 = z * z, z3 = z2 * z, z4 = z3 * z, z5 = z4 * z;
    *node = nd;
    bg[0] = (6.0 * z - 5.0 * z2 - 5.0 * z3 + 5.0 * z4 - z5) / 120.0;
    bg[1] = (-12.0 * z + 16.0 * z2 - z3 - 4.0 * z4 + z5) / 24.0;
    bg[2] = (12.0 - 4.0 * z - 15.0 * z2 + 5.0 * z3 + 3.0 * z4 - z5) / 12.0;
    bg[3] = (12.0 * z + 8.0 * z2 - 7.0 * z3 - 2.0 * z4 + z5) / 12.0;
    bg[4] = (-6.0 * z - z2 + 7.0 * z3 + z4 - z5) / 24.0;
    bg[5] = (4.0 * z - 5.0 * z3 + z5) / 120.0;
    return 0;
}

/* INTERPOLATE x3, partlib.f90:351-520 (halo vext replaced by periodic index wrap,
 * which is what vext holds on one domain, 444-452). */
static void interpolate(orc_sim* s, const double* ux, const double* uy, const double* uz)
{
    const int N = s->N, L = s->L;
    const double dx = s->pi2 / (double)N;
    const double* u[3] = {ux, uy, uz};
    int ih;
#pragma omp parallel for schedule(static)
    for (ih = 1; ih <= s->nps; ++ih) {
        int nd[3], ix[6], iy[6], iz[6], m, c, a, b, d;
        double bg[3][6];
        for (c = 0; c < 3; ++c) orc_interp_weights(YP(c + 1, ih), dx, &nd[c], bg[c]);
        for (m = 0; m < 6; ++m) {
            ix[m] = ((nd[0] - 2 + m) % N + N) % N;
            iy[m] = ((nd[1] - 2 + m) % N + N) % N;
            iz[m] = ((nd[2] - 2 + m) % N + N) % N;
        }
        for (c = 0; c < 3; ++c) {
            double acc = 0.0;
            for (a = 0; a < 6; ++a)
                for (b = 0; b < 6; ++b)
                    for (d = 0; d < 6; ++d)
                        acc = acc + u[c][IDX(ix[a], iy[b], iz[d], L)] * bg[0][a] * bg[1][b] * bg[2][d];
            VP(c + 1, ih) = acc;
        }
    }
}

/* PART_HISTORY_FVP, partlib.f90:553-595 */
static void part_history_fvp(orc_sim* s)
{
    int ih, c;
    for (ih = 1; ih <= s->nps; ++ih) {
        if (s->new_particle) {
            double a = YP(7, ih);
            YP(4, ih) = s->Sv_const * (a * a);
            YP(5, ih) = 0.0;
            YP(6, ih) = 0.0;
            for (c = 1; c <= 3; ++c) { FVP(c, 0, ih) = YP(3 + c, ih); FVP(c, 1, ih) = FVP(c, 0, ih); FVP(c, 2, ih) = FVP(c, 0, ih); }
        } else {
            for (c = 1; c <= 3; ++c) FVP(c, 0, ih) = YP(3 + c, ih);
        }
    }
}

/* PART_HISTORY_FDP, partlib.f90:599-625 */
static void part_history_fdp(orc_sim* s)
{
    int ih, c;
    for (ih = 1; ih <= s->nps; ++ih)
        for (c = 1; c <= 3; ++c) {
            double f = s->p.hdi_included ? VP(c, ih) + PV(c, ih) : VP(c, ih);
            FDP(c, 0, ih) = f;
            if (s->new_particle) { FDP(c, 1, ih) = f; FDP(c, 2, ih) = f; FDP(c, 3, ih) = f; }
        }
    s->new_particle = 0;
}

/* PART_ADVANCE, partlib.f90:629-726.  Pinned: uniform steady flow is a fixed point
 * V = U + W e_x; Stokes relaxation matches exp(-t/tau_p) as dt -> 0. */
static void part_advance(orc_sim* s)
{
    const double dt = s->p.dt;
    int ih, ic;
    for (ih = 1; ih <= s->nps; ++ih) {
        double FSN = 1.;
        double a = YP(7, ih);
        if (s->p.ndrag == 1) {
            double Vrel, d1, d2, d3;
            d1 = YP(4, ih) - VP(1, ih); d2 = YP(5, ih) - VP(2, ih); d3 = YP(6, ih) - VP(3, ih);
            if (s->p.hdi_included) {
                d1 = YP(4, ih) - VP(1, ih) - PV(1, ih);
                d2 = YP(5, ih) - VP(2, ih) - PV(2, ih);
                d3 = YP(6, ih) - VP(3, ih) - PV(3, ih);
            }
            Vrel = sqrt(d1 * d1 + d2 * d2 + d3 * d3);
            FSN = 1. + 0.15 * pow(Vrel * 2. * a / s->p.rnu, 0.687);
        }
        for (ic = 1; ic <= 3; ++ic) {
            double unitv = (ic == 1) ? 1.0 : 0.0;
            double taup = s->tau_CLOUD_const2 * (a * a);
            double dtau = dt / (24.0 * taup) * FSN;
            double tempv = 19.0 * FVP(ic, 0, ih) - 5.0 * FVP(ic, 1, ih) + FVP(ic, 2, ih);
            double tempd = 55.0 * FDP(ic, 0, ih) - 59.0 * FDP(ic, 1, ih) + 37.0 * FDP(ic, 2, ih) - 9.0 * FDP(ic, 3, ih);
            double wp_vel = s->Sv_const * (a * a);
            YP(3 + ic, ih) = (YP(3 + ic, ih) + dt * unitv * wp_vel / taup * FSN - dtau * tempv + dtau * tempd) / (1.0 + 9.0 * dtau);
        }
    }
    {
        double dtc = dt / 24.0;                               /* partlib.f90:681-684 */
        for (ih = 1; ih <= s->nps; ++ih)
            for (ic = 1; ic <= 3; ++ic)
                DM(ic, ih) = dtc * (9.0 * YP(3 + ic, ih) + 19.0 * FVP(ic, 0, ih) - 5.0 * FVP(ic, 1, ih) + FVP(ic, 2, ih));
    }
}

/* UPDATE_PART_POSITION, partlib.f90:730-757 */
static void update_part_position(orc_sim* s)
{
    int ih, c;
    for (ih = 1; ih <= s->nps; ++ih) {
        for (c = 1; c <= 3; ++c) {
            FVP(c, 2, ih) = FVP(c, 1, ih); FVP(c, 1, ih) = FVP(c, 0, ih);
            FDP(c, 3, ih) = FDP(c, 2, ih); FDP(c, 2, ih) = FDP(c, 1, ih); FDP(c, 1, ih) = FDP(c, 0, ih);
        }
        for (c = 1; c <= 3; ++c) YP(c, ih) = YP(c, ih) + DM(c, ih);
    }
}

/* PERIODIC wrap, partlib.f90:1142-1151 */
static void periodic(orc_sim* s)
{
    const double hx = s->pi2, pi2 = s->pi2;
    int ih;
    for (ih = 1; ih <= s->nps; ++ih) {
        if (YP(1, ih) >= hx) YP(1, ih) = YP(1, ih) - hx;
        if (YP(1, ih) < 0.0) YP(1, ih) = YP(1, ih) + hx;
        if (YP(2, ih) >= pi2) YP(2, ih) = YP(2, ih) - pi2;
        if (YP(2, ih) < 0.0) YP(2, ih) = YP(2, ih) + pi2;
        if (YP(3, ih) >= pi2) YP(3, ih) = YP(3, ih) - pi2;
        if (YP(3, ih) < 0.0) YP(3, ih) = YP(3, ih) + pi2;
    }
}

/* HEADLIST, partlib.f90:2171-2444, on one domain: the eight halo "neighbours" are
 * the domain itself, so each direction appends the images of the droplets within
 * dist of the matching face or edge, shifted by -+2pi (2231-2347), in direction
 * order 1..8 (2359-2389); then the linked cell list over locals and images
 * (2402-2434).  Pinned through COLLISION_DET / OVERLAP_DET / NEI_LIST_PRECOMP
 * against brute-force minimum-image pair searches. */
static int headlist(orc_sim* s, double dist)
{
    const double pi2 = s->pi2, hx = s->pi2, hy = s->pi2;
    const int nps = s->nps, ncd = s->ncd, R = s->ypn_rows;
    int cnt[8] = {0, 0, 0, 0, 0, 0, 0, 0};
    int ih, d, total, pos;
    /* pass 1: membership */
    for (ih = 1; ih <= nps; ++ih) {
        double y = YP(2, ih), z = YP(3, ih);
        double d_e = fabs(1.0 * hy - y), d_w = fabs(0.0 * hy - y);
        double d_n = fabs(1.0 * hy - z), d_s = fabs(0.0 * hy - z);
        int in[8];
        in[0] = d_e <= dist;
        in[1] = sqrt(d_e * d_e + d_n * d_n) <= dist;
        in[2] = d_n <= dist;
        in[3] = sqrt(d_w * d_w + d_n * d_n) <= dist;
        in[4] = d_w <= dist;
        in[5] = sqrt(d_w * d_w + d_s * d_s) <= dist;
        in[6] = d_s <= dist;
        in[7] = sqrt(d_e * d_e + d_s * d_s) <= dist;
        for (d = 0; d < 8; ++d) if (in[d]) s->pindex[(size_t)d * s->cap + cnt[d]++] = ih;
    }
    total = nps;
    for (d = 0; d < 8; ++d) { s->ih_p[d] = cnt[d]; total += cnt[d]; }
    if (total > s->ypn_cap) {
        free(s->ypn); free(s->vpn); free(s->list);
        s->ypn_cap = total + total / 4 + 16;
        s->ypn = (double*)xcalloc((size_t)R * s->ypn_cap, 8);
        s->vpn = (double*)xcalloc((size_t)3 * s->ypn_cap, 8);
        s->list = (int*)xcalloc((size_t)s->ypn_cap + 1, sizeof(int));
    }
    for (ih = 1; ih <= nps; ++ih) {                          /* partlib.f90:2202-2206 */
        int r;
        for (r = 1; r <= 8; ++r) YPN(r, ih) = YP(r, ih);
        for (r = 1; r <= 3; ++r) { YPN(8 + r, ih) = DM(r, ih); YPN(11 + r, ih) = FVP(r, 0, ih); VPN(r, ih) = VP(r, ih); }
        YPN(15, ih) = YP(8, ih);
    }
    pos = nps;
    for (d = 0; d < 8; ++d) {
        int m;
        for (m = 0; m < cnt[d]; ++m) {
            int src = s->pindex[(size_t)d * s->cap + m], r;
            ++pos;
            for (r = 1; r <= 8; ++r) YPN(r, pos) = YP(r, src);
            for (r = 1; r <= 3; ++r) { YPN(8 + r, pos) = DM(r, src); YPN(11 + r, pos) = FVP(r, 0, src); VPN(r, pos) = VP(r, src); }
            YPN(15, pos) = (double)src;
            switch (d + 1) {                                   /* shifts, 2231-2347 */
            case 1: YPN(2, pos) = YP(2, src) - pi2; break;
            case 2: YPN(2, pos) = YP(2, src) - pi2; YPN(3, pos) = YP(3, src) - pi2; break;
            case 3: YPN(3, pos) = YP(3, src) - pi2; break;
            case 4: YPN(2, pos) = YP(2, src) + pi2; YPN(3, pos) = YP(3, src) - pi2; break;
            case 5: YPN(2, pos) = YP(2, src) + pi2; break;
            case 6: YPN(2, pos) = YP(2, src) + pi2; YPN(3, pos) = YP(3, src) + pi2; break;
            case 7: YPN(3, pos) = YP(3, src) + pi2; break;
            case 8: YPN(2, pos) = YP(2, src) - pi2; YPN(3, pos) = YP(3, src) + pi2; break;
            }
        }
    }
    s->nps2 = total;
    /* linked cell list, 2402-2434 */
    memset(s->head, 0, sizeof(int) * (size_t)ncd * s->hny * s->hny);
    for (ih = 1; ih <= total; ++ih) {
        int ix, iy, iz;
        double dist_r, dist_b;
        ix = 1 + (int)(YPN(1, ih) / hx * (double)ncd);
        dist_r = YPN(2, ih) - 0.0 * hy;
        if (dist_r > 0) iy = 1 + (int)(dist_r / hy * (double)ncd);
        else iy = -(int)(fabs(dist_r) / hy * (double)ncd);
        dist_b = YPN(3, ih) - 0.0 * hy;
        if (dist_b > 0) iz = 1 + (int)(dist_b / hy * (double)ncd);
        else iz = -(int)(fabs(dist_b) / hy * (double)ncd);
        if (ix > ncd) ix = ncd;
        if (iy > ncd + s->nx_max) iy = ncd + s->nx_max;
        if (iy < 1 - s->nx_max) iy = 1 - s->nx_max;
        if (iz > ncd + s->nx_max) iz = ncd + s->nx_max;
        if (iz < 1 - s->nx_max) iz = 1 - s->nx_max;
        {
            size_t c = (size_t)(ix - 1) + (size_t)ncd * ((size_t)(iy - s->hy0) + (size_t)s->hny * (size_t)(iz - s->hy0));
            s->list[ih] = s->head[c];
            s->head[c] = ih;
        }
    }
    return 0;
}

static inline int head_at(const orc_sim* s, int ix, int iy, int iz)
{
    return s->head[(size_t)(ix - 1) + (size_t)s->ncd * ((size_t)(iy - s->hy0) + (size_t)s->hny * (size_t)(iz - s->hy0))];
}

/* COMPLETE_YPN, partlib.f90:761-838: refresh dmove and the new velocity of locals
 * and images after PART_ADVANCE. */
static void complete_ypn(orc_sim* s)
{
    int ih, d, m, c, pos = s->nps;
    for (ih = 1; ih <= s->nps; ++ih)
        for (c = 1; c <= 3; ++c) { YPN(8 + c, ih) = DM(c, ih); YPN(3 + c, ih) = YP(3 + c, ih); }
    for (d = 0; d < 8; ++d)
        for (m = 0; m < s->ih_p[d]; ++m) {
            int src = s->pindex[(size_t)d * s->cap + m];
            ++pos;
            for (c = 1; c <= 3; ++c) { YPN(8 + c, pos) = DM(c, src); YPN(3 + c, pos) = YP(3 + c, src); }
        }
}

/* FINDINDEX, partlib.f90:1850-1885 */
static void findindex(const orc_sim* s, double rad1, double rad2, int* index1, int* index2)
{
    double diff1 = 1000000., diff2 = 1000000.;
    int il;
    *index1 = 1; *index2 = 1;
    for (il = 1; il <= s->p.nset; ++il) {
        double d1 = fabs(s->radius[il - 1] - rad1), d2 = fabs(s->radius[il - 1] - rad2);
        if (d1 <= diff1) { diff1 = d1; *index1 = il; }
        if (d2 <= diff2) { diff2 = d2; *index2 = il; }
    }
    if (*index2 < *index1) { int t = *index2; *index2 = *index1; *index1 = t; }
}

static void push_removal(orc_sim* s, int ind)
{
    if (s->ncoll >= s->remove_cap) {
        s->remove_cap *= 2;
        s->removepart = (int*)realloc(s->removepart, sizeof(int) * s->remove_cap);
    }
    s->removepart[s->ncoll++] = ind;
}

static void push_pair(orc_sim* s, int ih1, int ih2, int type)
{
    int64_t g1 = (int64_t)YP(8, ih1), g2 = (int64_t)YPN(8, ih2);
    int r;
    if (s->npairs >= s->pairs_cap) {
        s->pairs_cap *= 2;
        s->pairs = (int64_t*)realloc(s->pairs, sizeof(int64_t) * 3 * s->pairs_cap);
        s->collist = (double*)realloc(s->collist, sizeof(double) * 20 * s->pairs_cap);
    }
    s->pairs[3 * s->npairs + 0] = g1 < g2 ? g1 : g2;
    s->pairs[3 * s->npairs + 1] = g1 < g2 ? g2 : g1;
    s->pairs[3 * s->npairs + 2] = type;
    for (r = 1; r <= 8; ++r) {                               /* partlib.f90:1557-1562 */
        s->collist[20 * s->npairs + r - 1] = YP(r, ih1);
        s->collist[20 * s->npairs + 10 + r - 1] = YPN(r, ih2);
    }
    s->collist[20 * s->npairs + 8] = (double)ih1;
    s->collist[20 * s->npairs + 9] = 0.0;
    s->collist[20 * s->npairs + 18] = (double)ih2;
    s->collist[20 * s->npairs + 19] = 0.0;
    s->npairs++;
}

static int cmp_int(const void* a, const void* b) { int x = *(const int*)a, y = *(const int*)b; return (x > y) - (x < y); }
static int cmp_triple(const void* a, const void* b)
{
    const int64_t* x = (const int64_t*)a; const int64_t* y = (const int64_t*)b;
    int i;
    for (i = 0; i < 3; ++i) if (x[i] != y[i]) return (x[i] > y[i]) - (x[i] < y[i]);
    return 0;
}

/* COLLISION_DET, partlib.f90:1166-1846 (the per-rank part, 1229-1585, plus the
 * removal list 1629-1631).  Pinned: the detected pairs equal a brute-force
 * minimum-image search with the same tests; two-body trajectories with known
 * contact produce types I, II, III; uniform droplets give RDF = 1. */
static void collision_det(orc_sim* s)
{
    const double hx = s->pi2, hy = s->pi2, dt = s->p.dt;
    const int ncd = s->ncd, nps = s->nps, nset = s->p.nset > 0 ? s->p.nset : 1;
    const int nbin = DDNS_NBIN;
    int nps4, nps5, nps6, nps7, nps8;
    double d_stat;
    int imax, ih1, i;
    memset(&s->cc, 0, sizeof(s->cc));
    s->cc.nset = nset;
    memset(s->nirbin0, 0, sizeof(int) * s->nbins);
    memset(s->wrbin0, 0, 8 * s->nbins); memset(s->wxbin0, 0, 8 * s->nbins);
    memset(s->wybin0, 0, 8 * s->nbins); memset(s->wzbin0, 0, 8 * s->nbins);
    memset(s->urbin0, 0, 8 * s->nbins); memset(s->thetabin0, 0, 8 * s->nbins);
    memset(s->wrtan0, 0, 8 * s->nbins);
    s->npairs = 0;
    s->ncoll = 0;
    s->ncontact = 0;
    nps4 = nps + s->ih_p[0] + s->ih_p[1] + s->ih_p[2] + s->ih_p[3];   /* 1256-1260 */
    nps5 = nps4 + s->ih_p[4];
    nps6 = nps5 + s->ih_p[5];
    nps7 = nps6 + s->ih_p[6];
    nps8 = nps7 + s->ih_p[7];
    d_stat = 20.0 * s->radmax;                                         /* 1264-1267 */
    imax = 1 + (int)(d_stat * (double)ncd / hx);
#define BIN(a, b, r) (((a) - 1) + nset * (((b) - 1) + nset * (r)))
    for (ih1 = 1; ih1 <= nps; ++ih1) {
        double xa[3], xan[3], va[3];
        int ix, iy, iz, ii, jj, kk, ncollip1 = 0;
        for (i = 0; i < 3; ++i) { xa[i] = YP(i + 1, ih1); xan[i] = xa[i] + DM(i + 1, ih1); va[i] = FVP(i + 1, 0, ih1); }
        ix = 1 + (int)(xa[0] / hx * (double)ncd);
        iy = 1 + (int)((xa[1] - 0.0 * hy) / hy * (double)ncd);
        iz = 1 + (int)((xa[2] - 0.0 * hy) / hy * (double)ncd);
        if (ix > ncd) ix = ncd;
        if (iy > ncd) iy = ncd;
        if (iz > ncd) iz = ncd;
        for (ii = -imax; ii <= imax; ++ii)
            for (jj = -imax; jj <= imax; ++jj)
                for (kk = -imax; kk <= imax; ++kk) {
                    int iy2 = iy + jj, iz2 = iz + kk, ix2 = ix + ii, ih2;
                    double sx = 0.0;
                    if (iy2 < 1 - s->nx_coll || iy2 > ncd + s->nx_coll) continue;   /* 1302-1306 */
                    if (iz2 < 1 - s->nx_coll || iz2 > ncd + s->nx_coll) continue;
                    if (iz2 < 1) continue;                                            /* 1309-1310 */
                    if (iz2 <= ncd && iy2 < 1) continue;
                    if (ix2 < 1) { ix2 = ncd + ix2; sx = -1.; }                       /* 1315-1321 */
                    else if (ix2 > ncd) { ix2 = ix2 % ncd; sx = 1.; }
                    ih2 = head_at(s, ix2, iy2, iz2);
                    if (ih2 <= ih1 || ih2 == 0) continue;
                    while (1) {
                        double xb[3], xbn[3], vb[3], xd[3], dnij0, rcoll, rcolla, rcollb;
                        double Wr = 0.0;
                        int index1 = 1, index2 = 1, c;
                        xb[0] = YPN(1, ih2) + sx * hx;                                 /* 1330-1337 */
                        xb[1] = YPN(2, ih2);
                        xb[2] = YPN(3, ih2);
                        for (c = 0; c < 3; ++c) { xbn[c] = xb[c] + YPN(9 + c, ih2); vb[c] = YPN(12 + c, ih2); }
                        for (c = 0; c < 3; ++c) xd[c] = xa[c] - xb[c];
                        dnij0 = sqrt(xd[0] * xd[0] + xd[1] * xd[1] + xd[2] * xd[2]);
                        s->cc.pairs_tested++;
                        rcoll = YP(7, ih1) + YPN(7, ih2);                              /* 1345-1353 */
                        rcolla = s->p.hdi_included ? rcoll : rcoll * (1. - s->p.st / 100.);
                        rcollb = rcoll * (1. + s->p.st / 100.);
                        if (dnij0 > rcolla && dnij0 <= 10 * rcoll) {                   /* 1358 */
                            s->cc.pairs_in_dstat++;
                            if (dnij0 > rcoll) {
                                double Vr1, Vr2, aaa, bbb, vangle, dvr[3], urr;
                                int ir;
                                for (c = 0; c < 3; ++c) xd[c] = xd[c] / dnij0;
                                Vr1 = va[0] * xd[0] + va[1] * xd[1] + va[2] * xd[2];
                                Vr2 = vb[0] * xd[0] + vb[1] * xd[1] + vb[2] * xd[2];
                                Wr = Vr1 - Vr2;
                                ir = (int)((dnij0 - rcoll) / ((9. / (double)nbin) * rcoll));
                                if (ir > nbin - 1) ir = nbin - 1;
                                if (ir < 0) ir = 0;
                                findindex(s, YP(7, ih1), YPN(7, ih2), &index1, &index2);
                                s->nirbin0[BIN(index1, index2, ir)] += 1;
                                if (index2 == index1) {
                                    s->wxbin0[BIN(index1, index2, ir)] += va[0] - vb[0];
                                    s->wybin0[BIN(index1, index2, ir)] += va[1] - vb[1];
                                    s->wzbin0[BIN(index1, index2, ir)] += va[2] - vb[2];
                                }
                                if (YP(7, ih1) > YPN(7, ih2)) {
                                    s->wxbin0[BIN(index1, index2, ir)] += va[0] - vb[0];
                                    s->wybin0[BIN(index1, index2, ir)] += va[1] - vb[1];
                                    s->wzbin0[BIN(index1, index2, ir)] += va[2] - vb[2];
                                } else if (YP(7, ih1) < YPN(7, ih2)) {
                                    s->wxbin0[BIN(index1, index2, ir)] += vb[0] - va[0];
                                    s->wybin0[BIN(index1, index2, ir)] += vb[1] - va[1];
                                    s->wzbin0[BIN(index1, index2, ir)] += vb[2] - va[2];
                                }
                                s->wrbin0[BIN(index1, index2, ir)] += Wr;
                                for (c = 0; c < 3; ++c) dvr[c] = VPN(c + 1, ih1) - VPN(c + 1, ih2);
                                urr = dvr[0] * xd[0] + dvr[1] * xd[1] + dvr[2] * xd[2];
                                s->urbin0[BIN(index1, index2, ir)] += urr;
                                aaa = (va[0] - vb[0]) * xd[0] + (va[1] - vb[1]) * xd[1] + (va[2] - vb[2]) * xd[2];
                                bbb = sqrt((va[0] - vb[0]) * (va[0] - vb[0]) + (va[1] - vb[1]) * (va[1] - vb[1]) + (va[2] - vb[2]) * (va[2] - vb[2]));
                                if (bbb != 0.) {
                                    double ccc;
                                    vangle = acos(fabs(aaa / bbb));
                                    s->thetabin0[BIN(index1, index2, ir)] += vangle;
                                    ccc = (bbb * sin(vangle)) * (bbb * sin(vangle)) / dnij0;
                                    s->wrtan0[BIN(index1, index2, ir)] += ccc;
                                }
                            }
                            if (dnij0 <= rcollb) {                                       /* 1411-1439 */
                                double* r;
                                if (dnij0 <= rcoll) {
                                    /* the offset is normalised here, once, as in the branch above */
                                    double Vr1, Vr2;
                                    for (c = 0; c < 3; ++c) xd[c] = xd[c] / dnij0;
                                    Vr1 = va[0] * xd[0] + va[1] * xd[1] + va[2] * xd[2];
                                    Vr2 = vb[0] * xd[0] + vb[1] * xd[1] + vb[2] * xd[2];
                                    Wr = Vr1 - Vr2;
                                    findindex(s, YP(7, ih1), YPN(7, ih2), &index1, &index2);
                                }
                                s->cc.nir[(index1 - 1) + DDNS_MAXSET * (index2 - 1)] += 1;
                                if (s->ncontact >= s->contact_cap) {
                                    s->contact_cap *= 2;
                                    s->contact = (double*)realloc(s->contact, sizeof(double) * 12 * s->contact_cap);
                                }
                                r = s->contact + 12 * (size_t)s->ncontact++;             /* 1431-1437 */
                                r[0] = (double)index1;
                                r[1] = (double)index2;
                                for (c = 0; c < 3; ++c) { r[2 + c] = va[c]; r[5 + c] = vb[c]; r[8 + c] = xd[c]; }
                                r[11] = Wr;
                            }
                        }
                        /* detection over adjacent cells only, 1445 */
                        if (!(abs(ii) > 1 || abs(jj) > 1 || abs(kk) > 1)) {
                            double dmove_a, dmove_b, tmove, dnij1;
                            int type = 0;
                            dmove_a = sqrt(DM(1, ih1) * DM(1, ih1) + DM(2, ih1) * DM(2, ih1) + DM(3, ih1) * DM(3, ih1));
                            dmove_b = sqrt(YPN(9, ih2) * YPN(9, ih2) + YPN(10, ih2) * YPN(10, ih2) + YPN(11, ih2) * YPN(11, ih2));
                            tmove = rcoll + dmove_a + dmove_b;
                            if (!(dnij0 > tmove)) {
                                for (c = 0; c < 3; ++c) xd[c] = xan[c] - xbn[c];
                                dnij1 = sqrt(xd[0] * xd[0] + xd[1] * xd[1] + xd[2] * xd[2]);
                                if (!(dnij0 < rcoll && dnij1 > rcoll)) {
                                    if (dnij0 > rcoll && dnij1 <= rcoll) {
                                        type = 1;
                                    } else {
                                        double ax1[3], van[3], bx1[3], vbn[3];
                                        int icheck;
                                        for (c = 0; c < 3; ++c) {
                                            ax1[c] = DM(c + 1, ih1); van[c] = YP(4 + c, ih1);
                                            bx1[c] = YPN(9 + c, ih2); vbn[c] = YPN(4 + c, ih2);
                                        }
                                        for (icheck = 1; icheck <= 19 && !type; ++icheck) {
                                            double tfac = 0.05 * (double)icheck;
                                            double tfac1 = dt * tfac, tfac2 = tfac * tfac, tfac3 = tfac2 * tfac;
                                            double xta[3], xtb[3], dnij;
                                            for (c = 0; c < 3; ++c) {
                                                xta[c] = xa[c] + va[c] * tfac1 + (3. * ax1[c] - (2. * va[c] + van[c]) * dt) * tfac2 + ((van[c] + va[c]) * dt - 2. * ax1[c]) * tfac3;
                                                xtb[c] = xb[c] + vb[c] * tfac1 + (3. * bx1[c] - (2. * vb[c] + vbn[c]) * dt) * tfac2 + ((vbn[c] + vb[c]) * dt - 2. * bx1[c]) * tfac3;
                                                xd[c] = xta[c] - xtb[c];
                                            }
                                            dnij = sqrt(xd[0] * xd[0] + xd[1] * xd[1] + xd[2] * xd[2]);
                                            if (dnij0 > rcoll && dnij < rcoll) type = 2;
                                            else if (dnij1 <= rcoll && dnij > rcoll) type = 3;
                                        }
                                    }
                                }
                            }
                            if (type) {                                                     /* 1518-1562 */
                                s->cc.icc[type - 1]++;
                                ncollip1++;
                                findindex(s, YP(7, ih1), YPN(7, ih2), &index1, &index2);
                                s->cc.colltype[(index1 - 1) + DDNS_MAXSET * (index2 - 1)] += 1;
                                if (ih2 <= nps) push_removal(s, ih2);
                                else if (ih2 > nps4 && ih2 <= nps8) push_removal(s, (int)YPN(15, ih2));
                                push_pair(s, ih1, ih2, type);
                            }
                        }
                        ih2 = s->list[ih2];                                                 /* 1566 */
                        if (ih2 <= ih1 || ih2 == 0) break;
                    }
                }
        if (ncollip1 != 0) push_removal(s, ih1);                                            /* 1578-1583 */
    }
#undef BIN
    (void)nps5; (void)nps6; (void)nps7;
    qsort(s->removepart, s->ncoll, sizeof(int), cmp_int);                                   /* 1631 */
    qsort(s->pairs, s->npairs, 3 * sizeof(int64_t), cmp_triple);
    s->cc.ncoll = s->ncoll;
    s->cc.npairs = s->npairs;
}

/* RELOCATE, partlib.f90:1984-2012 */
static void relocate(orc_sim* s)
{
    int i;
    for (i = 0; i < s->ncoll; ++i) {
        int ind = s->removepart[i];
        double r1 = orc_ranff2(&s->iseedr, s->ivr, &s->iyr, &s->idum2);
        double r2 = orc_ranff2(&s->iseedr, s->ivr, &s->iyr, &s->idum2);
        double r3 = orc_ranff2(&s->iseedr, s->ivr, &s->iyr, &s->idum2);
        YP(1, ind) = r1 * s->pi2;
        YP(2, ind) = r2 * s->pi2 + 0.0 * s->pi2;
        YP(3, ind) = r3 * s->pi2 + 0.0 * s->pi2;
    }
}

/* COALESCE, partlib.f90:3642-3825: merge each recorded pair into its first droplet
 * (volume conserved, terminal velocity) and delete the second by moving the last
 * droplet into its slot.  Pinned by sum(a^3) invariance. */
static void coalesce(orc_sim* s)
{
    int i, jj = 0, c;
    for (i = 0; i < s->npairs; ++i) {
        const double* cl = s->collist + 20 * i;
        int p1ind = (int)cl[8], p2ind = (int)cl[18];
        if (p1ind <= s->nps) {
            double a = pow(pow(cl[6], 3.0) + pow(cl[16], 3.0), 1.0 / 3.0);
            YP(7, p1ind) = a;
            YP(4, p1ind) = s->Sv_const * (a * a);
            YP(5, p1ind) = 0.0; YP(6, p1ind) = 0.0;
        }
        if (p2ind <= s->nps) {
            int last;
            jj++;
            last = s->nps - jj + 1;
            for (c = 1; c <= 8; ++c) YP(c, p2ind) = YP(c, last);
            for (c = 1; c <= 3; ++c) { VP(c, p2ind) = VP(c, last); DM(c, p2ind) = DM(c, last); }
            for (c = 0; c < 9; ++c) s->fvp[(size_t)(p2ind - 1) * 9 + c] = s->fvp[(size_t)(last - 1) * 9 + c];
            for (c = 0; c < 12; ++c) s->fdp[(size_t)(p2ind - 1) * 12 + c] = s->fdp[(size_t)(last - 1) * 12 + c];
            for (c = 1; c <= 8; ++c) YP(c, last) = 0.0;
            for (c = 1; c <= 3; ++c) { VP(c, last) = 0.0; DM(c, last) = 0.0; }
            for (c = 0; c < 9; ++c) s->fvp[(size_t)(last - 1) * 9 + c] = 0.0;
            for (c = 0; c < 12; ++c) s->fdp[(size_t)(last - 1) * 12 + c] = 0.0;
        }
    }
    s->nps -= jj;
}

/* OVERLAP_DET, partlib.f90:2047-2167: sweep with HEADLIST(2 radmax), relocate the
 * lower-index droplet of each overlapping pair, repeat until a sweep finds none.
 * Pinned: no pair overlaps afterwards; the generator consumption equals 3 x the
 * number of overlapping pairs found per sweep. */
int orc_overlap_fix(orc_sim* s)
{
    const double hx = s->pi2, hy = s->pi2;
    const int ncd = s->ncd;
    double d_overl = 2.0 * s->radmax;
    int imax = 1 + (int)(d_overl * (double)ncd / hx);
    int sweeps = 0, total = 0;
    if (s->nps == 0) return 0;
    for (;;) {
        int noverl = 0, ih1;
        headlist(s, d_overl);
        for (ih1 = 1; ih1 <= s->nps; ++ih1) {
            int ix, iy, iz, ii, jj, kk;
            ix = 1 + (int)(YPN(1, ih1) / hx * (double)ncd);
            iy = 1 + (int)((YPN(2, ih1) - 0.0 * hy) / hy * (double)ncd);
            iz = 1 + (int)((YPN(3, ih1) - 0.0 * hy) / hy * (double)ncd);
            if (ix > ncd) ix = ncd;
            if (iy > ncd) iy = ncd;
            if (iz > ncd) iz = ncd;
            for (ii = -imax; ii <= imax; ++ii)
                for (jj = -imax; jj <= imax; ++jj)
                    for (kk = -imax; kk <= imax; ++kk) {
                        int iy2, iz2, ix2, ih2;
                        double sx = 0.0;
                        if (abs(ii) > 1 || abs(jj) > 1 || abs(kk) > 1) continue;      /* 2091 */
                        iy2 = iy + jj;
                        if (iy2 < 1 - imax || iy2 > ncd + imax) continue;
                        iz2 = iz + kk;
                        if (iz2 < 1 - imax || iz2 > ncd + imax) continue;
                        if (iz2 < 1) continue;
                        if (iz2 <= ncd && iy2 < 1) continue;
                        ix2 = ix + ii;
                        if (ix2 < 1) { ix2 = ncd + ix2; sx = -1.; }
                        else if (ix2 > ncd) { ix2 = ix2 % ncd; sx = 1.; }
                        ih2 = head_at(s, ix2, iy2, iz2);
                        if (ih2 <= ih1 || ih2 == 0) continue;
                        while (1) {
                            double xb[3], xd[3], dnij, rcoll;
                            xb[0] = YPN(1, ih2) + sx * hx;
                            xb[1] = YPN(2, ih2);
                            xb[2] = YPN(3, ih2);
                            xd[0] = YPN(1, ih1) - xb[0]; xd[1] = YPN(2, ih1) - xb[1]; xd[2] = YPN(3, ih1) - xb[2];
                            dnij = sqrt(xd[0] * xd[0] + xd[1] * xd[1] + xd[2] * xd[2]);
                            rcoll = YPN(7, ih1) + YPN(7, ih2);
                            if (dnij <= rcoll) {
                                double ar1, ar2, ar3;
                                noverl++;
                                ar1 = orc_ranff2(&s->iseedr, s->ivr, &s->iyr, &s->idum2);
                                ar2 = orc_ranff2(&s->iseedr, s->ivr, &s->iyr, &s->idum2);
                                ar3 = orc_ranff2(&s->iseedr, s->ivr, &s->iyr, &s->idum2);
                                YP(1, ih1) = ar1 * hx;
                                YP(2, ih1) = ar2 * hy + 0.0 * hy;
                                YP(3, ih1) = ar3 * hy + 0.0 * hy;
                            }
                            ih2 = s->list[ih2];
                            if (ih2 <= ih1 || ih2 == 0) break;
                        }
                    }
        }
        total += noverl;
        if (noverl == 0) break;
        if (++sweeps > 10000) return DDNS_ERR_OVERLAP;
    }
    s->cc.noverlap_sweeps = sweeps;
    s->cc.noverlap_relocated = total;
    return 0;
}

/* ------------------------------------------------------------------------------- */
/* Hydrodynamic interactions.  NEI_LIST_PRECOMP partlib.f90:3218-3371,              */
/* AFUN_PRECOMP 3375-3510, HDI_GMRES 2795-3005, EUC_NORM 2756-2791.  Pinned: two-   */
/* sphere systems against a dense numpy solve of (I+M)u = b with the Stokes-sphere  */
/* coefficients; random sparse systems against numpy.linalg.solve.                  */
/* ------------------------------------------------------------------------------- */

#define NEI(c, m, ih) s->nei_data[(((size_t)(ih) - 1) * DDNS_MAXNEI + (size_t)((m) - 1)) * 7 + ((c) - 1)]

static int nei_list_precomp(orc_sim* s)
{
    const double hx = s->pi2, hy = s->pi2;
    const int ncd = s->ncd, nps = s->nps, nmax = s->nx_max;
    int ih1, overflow = 0;
    free(s->nei_num); free(s->nei_data);
    s->nei_num = (int*)xcalloc((size_t)nps + 1, sizeof(int));
    s->nei_data = (double*)xcalloc((size_t)(nps + 1) * DDNS_MAXNEI * 7, 8);
    for (ih1 = 1; ih1 <= nps; ++ih1) {
        int ix, iy, iz, i, j, k, iy_min, iy_max, iz_min, iz_max;
        ix = 1 + (int)(YPN(1, ih1) / hx * (double)ncd);
        iy = 1 + (int)((YPN(2, ih1) - 0.0 * hy) / hy * (double)ncd);
        iz = 1 + (int)((YPN(3, ih1) - 0.0 * hy) / hy * (double)ncd);
        if (ix > ncd) ix = ncd;
        if (iy > ncd) iy = ncd;
        if (iz > ncd) iz = ncd;
        iy_max = iy + nmax; iy_min = iy - nmax; iz_max = iz + nmax; iz_min = iz - nmax;
        if (iy_max > ncd + nmax) iy_max = ncd + nmax;
        if (iy_min < 1 - nmax) iy_min = 1 - nmax;
        if (iz_max > ncd + nmax) iz_max = ncd + nmax;
        if (iz_min < 1 - nmax) iz_min = 1 - nmax;
        for (i = ix - nmax; i <= ix + nmax; ++i)
            for (j = iy_min; j <= iy_max; ++j)
                for (k = iz_min; k <= iz_max; ++k) {
                    int ii = i, ih2;
                    double sx = 0.0;
                    if (i > ncd) { ii = i % ncd; sx = 1.0; }
                    else if (i < 1) { ii = ncd + i; sx = -1.0; }
                    ih2 = head_at(s, ii, j, k);
                    if (ih2 == 0) continue;
                    while (ih2 > ih1) {
                        double xb[3], xd[3], dnij, hdirad;
                        xb[0] = YPN(1, ih2) + sx * hx;
                        xb[1] = YPN(2, ih2);
                        xb[2] = YPN(3, ih2);
                        xd[0] = YPN(1, ih1) - xb[0]; xd[1] = YPN(2, ih1) - xb[1]; xd[2] = YPN(3, ih1) - xb[2];
                        dnij = sqrt(xd[0] * xd[0] + xd[1] * xd[1] + xd[2] * xd[2]);
                        if (dnij < YPN(7, ih1) + YPN(7, ih2)) return DDNS_ERR_OVERLAP;     /* 3294-3299 */
                        hdirad = s->p.hdi_trunc * YPN(7, ih2);
                        if (dnij < hdirad) {
                            int inum = ++s->nei_num[ih1 - 1];
                            double fact1, fact2, gamma, beta;
                            if (inum > DDNS_MAXNEI) { overflow = 1; s->nei_num[ih1 - 1] = DDNS_MAXNEI; }
                            else {
                                NEI(1, inum, ih1) = (double)ih2;
                                fact1 = YPN(7, ih2) / dnij;
                                fact2 = fact1 * fact1 * fact1;
                                fact1 = 0.75 * fact1;
                                gamma = (fact1 - 0.75 * fact2) / (dnij * dnij);
                                beta = fact1 + 0.25 * fact2;
                                NEI(2, inum, ih1) = gamma * (xd[0] * xd[0]) + beta;
                                NEI(3, inum, ih1) = gamma * xd[0] * xd[1];
                                NEI(4, inum, ih1) = gamma * xd[0] * xd[2];
                                NEI(5, inum, ih1) = gamma * (xd[1] * xd[1]) + beta;
                                NEI(6, inum, ih1) = gamma * xd[1] * xd[2];
                                NEI(7, inum, ih1) = gamma * (xd[2] * xd[2]) + beta;
                            }
                        }
                        hdirad = s->p.hdi_trunc * YPN(7, ih1);
                        xd[0] = -xd[0]; xd[1] = -xd[1]; xd[2] = -xd[2];
                        if (ih2 <= nps && dnij < hdirad) {
                            int inum = ++s->nei_num[ih2 - 1];
                            double fact1, fact2, gamma, beta;
                            if (inum > DDNS_MAXNEI) { overflow = 1; s->nei_num[ih2 - 1] = DDNS_MAXNEI; }
                            else {
                                NEI(1, inum, ih2) = (double)ih1;
                                fact1 = YPN(7, ih1) / dnij;
                                fact2 = fact1 * fact1 * fact1;
                                fact1 = 0.75 * fact1;
                                gamma = (fact1 - 0.75 * fact2) / (dnij * dnij);
                                beta = fact1 + 0.25 * fact2;
                                NEI(2, inum, ih2) = gamma * (xd[0] * xd[0]) + beta;
                                NEI(3, inum, ih2) = gamma * xd[0] * xd[1];
                                NEI(4, inum, ih2) = gamma * xd[0] * xd[2];
                                NEI(5, inum, ih2) = gamma * (xd[1] * xd[1]) + beta;
                                NEI(6, inum, ih2) = gamma * xd[1] * xd[2];
                                NEI(7, inum, ih2) = gamma * (xd[2] * xd[2]) + beta;
                            }
                        }
                        ih2 = s->list[ih2];
                    }
                }
    }
    return overflow ? DDNS_ERR_NEIGHBOUR_OVERFLOW : 0;
}

/* AFUN_PRECOMP: Ax = x + sum_j M_ij x_j; on icall = 1 also b = sum_j M_ij (V_j - U_j).
 * x is in the reference's blocked order x(1:n), x(n+1:2n), x(2n+1:3n) (3407-3409);
 * images take the x of the droplet they copy (3413-3458). */
static void afun_precomp(orc_sim* s, const double* x, double* Ax, double* b, int icall)
{
    const int nps = s->nps;
    int ih1, d, m, pos;
    for (ih1 = 1; ih1 <= nps; ++ih1) { YPN(16, ih1) = x[ih1 - 1]; YPN(17, ih1) = x[nps + ih1 - 1]; YPN(18, ih1) = x[2 * nps + ih1 - 1]; }
    pos = nps;
    for (d = 0; d < 8; ++d)
        for (m = 0; m < s->ih_p[d]; ++m) {
            int src = s->pindex[(size_t)d * s->cap + m];
            ++pos;
            YPN(16, pos) = x[src - 1]; YPN(17, pos) = x[nps + src - 1]; YPN(18, pos) = x[2 * nps + src - 1];
        }
    for (ih1 = 1; ih1 <= nps; ++ih1) {
        double r1 = 0.0, r2 = 0.0, r3 = 0.0, b1 = 0.0, b2 = 0.0, b3 = 0.0;
        int ijk;
        for (ijk = 1; ijk <= s->nei_num[ih1 - 1]; ++ijk) {
            int ind = (int)NEI(1, ijk, ih1);
            double rv1 = YPN(16, ind), rv2 = YPN(17, ind), rv3 = YPN(18, ind);
            r1 = r1 + NEI(2, ijk, ih1) * rv1 + NEI(3, ijk, ih1) * rv2 + NEI(4, ijk, ih1) * rv3;
            r2 = r2 + NEI(3, ijk, ih1) * rv1 + NEI(5, ijk, ih1) * rv2 + NEI(6, ijk, ih1) * rv3;
            r3 = r3 + NEI(4, ijk, ih1) * rv1 + NEI(6, ijk, ih1) * rv2 + NEI(7, ijk, ih1) * rv3;
            if (icall == 1) {
                double q1 = YPN(4, ind) - VPN(1, ind), q2 = YPN(5, ind) - VPN(2, ind), q3 = YPN(6, ind) - VPN(3, ind);
                b1 = b1 + NEI(2, ijk, ih1) * q1 + NEI(3, ijk, ih1) * q2 + NEI(4, ijk, ih1) * q3;
                b2 = b2 + NEI(3, ijk, ih1) * q1 + NEI(5, ijk, ih1) * q2 + NEI(6, ijk, ih1) * q3;
                b3 = b3 + NEI(4, ijk, ih1) * q1 + NEI(6, ijk, ih1) * q2 + NEI(7, ijk, ih1) * q3;
            }
        }
        Ax[ih1 - 1] = r1 + x[ih1 - 1];
        Ax[nps + ih1 - 1] = r2 + x[nps + ih1 - 1];
        Ax[2 * nps + ih1 - 1] = r3 + x[2 * nps + ih1 - 1];
        if (b) { b[ih1 - 1] = b1; b[nps + ih1 - 1] = b2; b[2 * nps + ih1 - 1] = b3; }
    }
}

static double euc_norm(const double* v, int n)
{
    double x_c = 0.0;
    int i;
    for (i = 0; i < n; ++i) x_c = x_c + v[i] * v[i];
    return sqrt(x_c);
}

/* HDI_GMRES, partlib.f90:2795-3005 */
static int hdi_gmres(orc_sim* s, int* iters_out, double* res_out)
{
    const int nps = s->nps, psize = 3 * nps, maxiter = s->p.gmres_maxiter > 0 ? s->p.gmres_maxiter : 50;
    const int M1 = maxiter + 1;
    double maxerr = s->p.gmres_tol > 0 ? s->p.gmres_tol : 1.0e-6;
    double *x0, *Q, *v, *Ax, *bvec, *H, *Rn, *Qn, *bn, *y, *v1, *v2, *resv;
    double norm_b, norm_rhs, res = 0.0;
    int iter, i, j, status = 0, done = 0;
#define HH(r, c) H[((r) - 1) + (size_t)M1 * ((c) - 1)]
#define RN(r, c) Rn[((r) - 1) + (size_t)M1 * ((c) - 1)]
#define QN(r, c) Qn[((r) - 1) + (size_t)M1 * ((c) - 1)]
#define QQ(r, c) Q[((r) - 1) + (size_t)psize * ((c) - 1)]
    *iters_out = 0; *res_out = 0.0;
    if (nps == 0) return 0;
    x0 = (double*)xcalloc(psize, 8); v = (double*)xcalloc(psize, 8); Ax = (double*)xcalloc(psize, 8);
    bvec = (double*)xcalloc(psize, 8); Q = (double*)xcalloc((size_t)psize * M1, 8);
    H = (double*)xcalloc((size_t)M1 * maxiter, 8); Rn = (double*)xcalloc((size_t)M1 * M1, 8);
    Qn = (double*)xcalloc((size_t)M1 * M1, 8); bn = (double*)xcalloc(M1, 8); y = (double*)xcalloc(maxiter, 8);
    v1 = (double*)xcalloc(M1, 8); v2 = (double*)xcalloc(M1, 8); resv = (double*)xcalloc(M1, 8);
    for (i = 1; i <= nps; ++i) { x0[i - 1] = PV(1, i); x0[nps + i - 1] = PV(2, i); x0[2 * nps + i - 1] = PV(3, i); }
    afun_precomp(s, x0, Ax, bvec, 1);
    for (i = 0; i < psize; ++i) QQ(i + 1, 1) = bvec[i] - Ax[i];
    norm_b = euc_norm(bvec, psize);
    norm_rhs = euc_norm(Q, psize);
    if (norm_rhs < 1e-5) goto cleanup;                                   /* 2866-2868 */
    for (i = 0; i < psize; ++i) QQ(i + 1, 1) = QQ(i + 1, 1) / norm_rhs;
    for (iter = 1; iter <= maxiter; ++iter) {
        afun_precomp(s, &QQ(1, iter), Ax, NULL, 2);
        memcpy(v, Ax, 8 * (size_t)psize);
        for (j = 1; j <= iter; ++j) {                                    /* 2886-2894 */
            double dot = 0.0;
            for (i = 0; i < psize; ++i) dot = dot + QQ(i + 1, j) * v[i];
            HH(j, iter) = dot;
            for (i = 0; i < psize; ++i) v[i] = v[i] - HH(j, iter) * QQ(i + 1, j);
        }
        HH(iter + 1, iter) = euc_norm(v, psize);
        for (i = 0; i < psize; ++i) QQ(i + 1, iter + 1) = v[i] / HH(iter + 1, iter);
        if (iter == 1) {                                                  /* 2904-2914 */
            double nrm = sqrt(HH(1, 1) * HH(1, 1) + HH(2, 1) * HH(2, 1));
            double q11 = HH(1, 1) / nrm, q21 = HH(2, 1) / nrm;
            QN(1, 1) = q11; QN(2, 1) = q21;
            QN(1, 2) = q21; QN(2, 2) = -q11;
            RN(1, 1) = sqrt(HH(1, 1) * HH(1, 1) + HH(2, 1) * HH(2, 1));
            bn[0] = q11 * norm_rhs;
            y[0] = bn[0] / RN(1, 1);
        } else {                                                          /* 2916-2944 */
            double aa, bb, cc, dd, ss;
            int r, c;
            for (r = 1; r <= iter; ++r) {
                double acc = 0.0;
                for (c = 1; c <= iter; ++c) acc = acc + QN(r, c) * HH(c, iter);
                RN(r, iter) = acc;
            }
            aa = RN(iter, iter); bb = HH(iter + 1, iter);
            dd = sqrt(aa * aa + bb * bb);
            cc = aa / dd; ss = bb / dd;
            RN(iter, iter) = dd;
            for (c = 1; c <= iter; ++c) { v1[c - 1] = cc * QN(iter, c); v2[c - 1] = -ss * QN(iter, c); }
            v1[iter] = ss; v2[iter] = cc;
            for (c = 1; c <= iter + 1; ++c) { QN(iter, c) = v1[c - 1]; QN(iter + 1, c) = v2[c - 1]; }
            bn[iter - 2] = QN(iter - 1, 1) * norm_rhs;
            bn[iter - 1] = QN(iter, 1) * norm_rhs;
            y[iter - 1] = bn[iter - 1] / RN(iter, iter);
            for (r = iter - 1; r >= 1; --r) {
                double dot = 0.0;
                for (c = r + 1; c <= iter; ++c) dot = dot + y[c - 1] * RN(r, c);
                y[r - 1] = (bn[r - 1] - dot) / RN(r, r);
            }
        }
        /* x = x0 + Q y (2951-2956) */
        for (i = 0; i < psize; ++i) {
            double acc = 0.0;
            for (j = 1; j <= iter; ++j) acc = acc + QQ(i + 1, j) * y[j - 1];
            v[i] = x0[i] + acc;
        }
        for (i = 1; i <= nps; ++i) { PV(1, i) = v[i - 1]; PV(2, i) = v[nps + i - 1]; PV(3, i) = v[2 * nps + i - 1]; }
        {                                                                 /* 2959-2963 */
            int r, c;
            for (r = 1; r <= iter + 1; ++r) {
                double acc = 0.0;
                for (c = 1; c <= iter; ++c) acc = acc + HH(r, c) * y[c - 1];
                resv[r - 1] = acc;
            }
            resv[0] = resv[0] - norm_rhs;
            res = euc_norm(resv, iter + 1) / norm_b;
        }
        *iters_out = iter;
        if (res < maxerr) { done = 1; break; }                            /* 2969-2971 */
        if (iter == maxiter / 2) maxerr = maxerr * 10;
        if (iter >= maxiter - 2) maxerr = maxerr * 10;
    }
    if (!done && res > maxerr) status = DDNS_WARN_GMRES;
cleanup:
    *res_out = res;
    free(x0); free(v); free(Ax); free(bvec); free(Q); free(H); free(Rn); free(Qn);
    free(bn); free(y); free(v1); free(v2); free(resv);
#undef HH
#undef RN
#undef QN
#undef QQ
    return status;
}

/* Stand-alone HI solve for tests: droplets yp(8,n) at rest positions, vold(3,n) their
 * velocities (ypn(4:6) at NEI_LIST time), vp(3,n) fluid velocity; pertvel in/out. */
int orc_hdi_solve(orc_sim* s, int n, const double* yp, const double* vold, const double* vp,
                  double* pertvel, int* iters, double* res)
{
    int ih, c, st;
    if (n > s->cap) return DDNS_ERR_CAPACITY;
    s->nps = n;
    for (ih = 1; ih <= n; ++ih) {
        for (c = 1; c <= 8; ++c) YP(c, ih) = yp[(size_t)(ih - 1) * 8 + c - 1];
        for (c = 1; c <= 3; ++c) {
            YP(3 + c, ih) = vold[(size_t)(ih - 1) * 3 + c - 1];
            VP(c, ih) = vp[(size_t)(ih - 1) * 3 + c - 1];
            PV(c, ih) = pertvel[(size_t)(ih - 1) * 3 + c - 1];
            FVP(c, 0, ih) = YP(3 + c, ih);
            DM(c, ih) = 0.0;
        }
    }
    headlist(s, s->d_hdi);
    st = nei_list_precomp(s);
    if (st < 0) return st;
    st = hdi_gmres(s, iters, res);
    for (ih = 1; ih <= n; ++ih)
        for (c = 1; c <= 3; ++c) pertvel[(size_t)(ih - 1) * 3 + c - 1] = PV(c, ih);
    return st;
}

/* ------------------------------------------------------------------------------- */
/* Droplet initialisation, PART_INIT partlib.f90:227-281                            */
/* ------------------------------------------------------------------------------- */

int orc_init_droplets(orc_sim* s, int seed, int radii_kind, double median_um,
                      double sigma_ln, double rmin_um, double rmax_um)
{
    const int npart = s->p.npart;
    int iseedp = -seed, iyp = 0, ivp[NTAB], ih;
    double scale;
    {
        double etk_CLOUD = sqrt(sqrt(s->p.viscosity_CLOUD * s->p.viscosity_CLOUD * s->p.viscosity_CLOUD / s->p.ediss_CLOUD));
        scale = (s->p.etk_DNS / etk_CLOUD) / 10000.0;
    }
    if (npart > s->cap) return DDNS_ERR_CAPACITY;
    memset(ivp, 0, sizeof(ivp));
    memset(s->yp, 0, (size_t)8 * s->cap * 8);
    for (ih = 1; ih <= npart; ++ih) {
        double l1 = orc_ranpp(&iseedp, &iyp, ivp);
        double l2 = orc_ranpp(&iseedp, &iyp, ivp);
        double l3 = orc_ranpp(&iseedp, &iyp, ivp);
        YP(1, ih) = s->pi2 * l1;
        YP(2, ih) = s->pi2 * l2;
        YP(3, ih) = s->pi2 * l3;
        YP(8, ih) = (double)ih;
    }
    if (radii_kind == DDNS_RADII_LOGNORMAL) {
        /* radius = median * exp(sigma * g), g standard normal by Box-Muller on a
         * second ranpp stream seeded -(seed + 1), clipped to [rmin, rmax] */
        int is2 = -(seed + 1), iy2 = 0, iv2[NTAB];
        memset(iv2, 0, sizeof(iv2));
        for (ih = 1; ih <= npart; ih += 2) {
            double u1 = orc_ranpp(&is2, &iy2, iv2), u2 = orc_ranpp(&is2, &iy2, iv2);
            double rr = sqrt(-2.0 * log(u1)), th = 2.0 * M_PI * u2;
            double g[2] = {rr * cos(th), rr * sin(th)};
            int q;
            for (q = 0; q < 2 && ih + q <= npart; ++q) {
                double r = median_um * exp(sigma_ln * g[q]);
                if (r < rmin_um) r = rmin_um;
                if (r > rmax_um) r = rmax_um;
                YP(7, ih + q) = r * scale;
            }
        }
    } else {
        for (ih = 1; ih <= npart; ++ih) {
            int i1 = (int)floor((double)(ih - 1) * (double)s->p.nset / (double)npart) + 1;
            YP(7, ih) = s->radius[i1 - 1];
        }
    }
    s->nps = npart;
    memset(s->fvp, 0, (size_t)9 * s->cap * 8);
    memset(s->fdp, 0, (size_t)12 * s->cap * 8);
    memset(s->pertvel, 0, (size_t)3 * s->cap * 8);
    periodic(s);                                                         /* turb.f90:427 */
    if (s->p.nonoverlap) return orc_overlap_fix(s);                      /* turb.f90:429 */
    return 0;
}

/* ------------------------------------------------------------------------------- */
/* The step, turb.f90:492-707                                                        */
/* ------------------------------------------------------------------------------- */

static int droplet_step(orc_sim* s)
{
    int status = 0, st;
    double dist;
    interpolate(s, s->vx, s->vy, s->vz);     /* physical u at this point, 553-556 */
    part_history_fvp(s);                                                 /* 564 */
    dist = s->p.hdi_included ? s->d_hdi : s->d_coll;                     /* 566-567 */
    headlist(s, dist);                                                   /* 569 */
    s->cc.gmres_iters = 0; s->cc.gmres_residual = 0.0;
    if (s->p.hdi_included) {                                             /* 571-572 */
        int it; double res;
        st = nei_list_precomp(s);
        if (st < 0) return st;
        st = hdi_gmres(s, &it, &res);
        if (st != 0) status = st;
        s->cc.gmres_iters = it; s->cc.gmres_residual = res;
    }
    part_history_fdp(s);                                                 /* 574 */
    part_advance(s);                                                     /* 576 */
    complete_ypn(s);                                                     /* 580 */
    {
        int it = s->cc.gmres_iters; double res = s->cc.gmres_residual;
        collision_det(s);                                                /* 584 */
        s->cc.gmres_iters = it; s->cc.gmres_residual = res;
    }
    update_part_position(s);                                             /* 586 */
    if (s->istep < s->p.nstep_grow && s->p.nonoverlap && s->ncoll > 0) relocate(s);   /* 590 */
    if (s->istep >= s->p.nstep_grow && s->p.nonoverlap && s->ncoll > 0) coalesce(s);  /* 591 */
    if (s->nps < s->p.npart / 2) return DDNS_ERR_DROPLET_LOSS;          /* 600-604 */
    periodic(s);                                                         /* 608 */
    if (s->p.nonoverlap) {                                               /* 610 */
        int sw = 0, rl = 0;
        ddns_collision_counts keep = s->cc;
        st = orc_overlap_fix(s);
        sw = s->cc.noverlap_sweeps; rl = s->cc.noverlap_relocated;
        s->cc = keep;
        s->cc.noverlap_sweeps = sw; s->cc.noverlap_relocated = rl;
        if (st < 0) return st;
    }
    return status;
}

int orc_step(orc_sim* s, int nsteps)
{
    int n, status = 0;
    for (n = 0; n < nsteps; ++n) {
        int st;
        s->istep += 1;                                                   /* 492 */
        flow_first_half(s);
        if (!s->flow_only && s->p.npart > 0 && s->nps > 0) {
            st = droplet_step(s);
            if (st < 0) return st;
            if (st > 0 && status == 0) status = st;
        }
        flow_second_half(s);
        s->time += s->p.dt;                                              /* 693 */
    }
    return status;
}

/* COLLISION_DET on a supplied state (kernel-level parity). */
int orc_collision_detect(orc_sim* s, int n, const double* yp, const double* dmove,
                         const double* fvp0, const double* vnew, const double* vp)
{
    int ih, c;
    double dist;
    if (n > s->cap) return DDNS_ERR_CAPACITY;
    s->nps = n;
    for (ih = 1; ih <= n; ++ih) {
        for (c = 1; c <= 8; ++c) YP(c, ih) = yp[(size_t)(ih - 1) * 8 + c - 1];
        for (c = 1; c <= 3; ++c) {
            DM(c, ih) = dmove[(size_t)(ih - 1) * 3 + c - 1];
            FVP(c, 0, ih) = fvp0[(size_t)(ih - 1) * 3 + c - 1];
            VP(c, ih) = vp[(size_t)(ih - 1) * 3 + c - 1];
        }
    }
    dist = s->p.hdi_included ? s->d_hdi : s->d_coll;
    headlist(s, dist);
    for (ih = 1; ih <= n; ++ih)
        for (c = 1; c <= 3; ++c) YP(3 + c, ih) = vnew[(size_t)(ih - 1) * 3 + c - 1];
    complete_ypn(s);
    collision_det(s);
    return 0;
}

int orc_get_collision_counts(orc_sim* s, ddns_collision_counts* out) { *out = s->cc; return 0; }

int orc_get_collision_bins(orc_sim* s, int* nirbin, double* wrbin, double* wxbin,
                           double* wybin, double* wzbin, double* urbin,
                           double* thetabin, double* wrtanbin)
{
    size_t n = (size_t)s->nbins;
    if (nirbin) memcpy(nirbin, s->nirbin0, n * sizeof(int));
    if (wrbin) memcpy(wrbin, s->wrbin0, n * 8);
    if (wxbin) memcpy(wxbin, s->wxbin0, n * 8);
    if (wybin) memcpy(wybin, s->wybin0, n * 8);
    if (wzbin) memcpy(wzbin, s->wzbin0, n * 8);
    if (urbin) memcpy(urbin, s->urbin0, n * 8);
    if (thetabin) memcpy(thetabin, s->thetabin0, n * 8);
    if (wrtanbin) memcpy(wrtanbin, s->wrtan0, n * 8);
    return 0;
}

int orc_get_collision_pairs(orc_sim* s, int64_t* triples, int cap, int* npairs)
{
    int m = s->npairs < cap ? s->npairs : cap;
    *npairs = s->npairs;
    if (triples && m > 0) memcpy(triples, s->pairs, sizeof(int64_t) * 3 * (size_t)m);
    return 0;
}

int orc_get_contact_records(orc_sim* s, double* rec, int cap, int* n)
{
    int m = s->ncontact < cap ? s->ncontact : cap;
    *n = s->ncontact;
    if (rec && m > 0) memcpy(rec, s->contact, sizeof(double) * 12 * (size_t)m);
    return 0;
}

int orc_get_removal_list(orc_sim* s, int* list, int cap, int* n)
{
    int m = s->ncoll < cap ? s->ncoll : cap;
    *n = s->ncoll;
    if (list && m > 0) memcpy(list, s->removepart, sizeof(int) * (size_t)m);
    return 0;
}
