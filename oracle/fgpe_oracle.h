/*
 * fgpe_oracle.h -- CPU FP64 restatement of the fGPE time-stepping hot path.
 *
 * TEST INFRASTRUCTURE ONLY.  This library is the checker for the HIP path:
 * only tests/, __graft_entry__.smoke() and bench.py's cpu_baseline leg may
 * load it.  The product library (libfgpe) never links or calls it.
 *
 * Every function restates one piece of the reference (file:line cited at the
 * definition).  Arrays of complex numbers are interleaved (re, im) doubles,
 * laid out as ID(x,y,z) = z*NX*NY + y*NX + x (hpp/header.hpp:20).
 */
#ifndef FGPE_ORACLE_H
#define FGPE_ORACLE_H

#ifdef __cplusplus
extern "C" {
#endif

typedef struct orc_grid {
    int nx, ny, nz;          /* points per axis; N = nx*ny*nz                 */
    const double *kx, *ky, *kz; /* k axes in cycles per length (InitializeKspace) */
    double dv;               /* volume element DX*DY*DZ                       */
} orc_grid;

typedef struct orc_ham {
    int    ncomp;            /* 1 or 2                                          */
    double a[3];             /* 1 comp: a[0]; 2 comp: a11, a22, a12 (code units) */
    int    quench_on;        /* 1: a(t) from the d_quench_as ramp              */
    double quench[6];        /* a0, a1, a2, t1, t2, t3                          */
    const double *vtilde;    /* real k-space kernel Vt(k) (N doubles) or NULL   */
    double vpx[2];           /* moving-frame factor per component (0: off)      */
    int    user_id;          /* test user pointwise term: 0 none, 1 or 2 (the
                                library's example terms, user_term.hpp)      */
    double user[8];          /* their parameters                                */
} orc_ham;

/* ---- building blocks ---------------------------------------------------- */
void   orc_kaxis(int n, double dk, double kmax, int active, double *k);
void   orc_twiddle(int n, int m, double *c, double *s);
void   orc_fft1d(int n, double *data, int sign);
void   orc_fft3d(int nx, int ny, int nz, double *data, int sign);
void   orc_laplace(const orc_grid *g, const double *psi, double *out);
void   orc_vdd(const orc_grid *g, double a_dd, const double d[3], double *vt);
void   orc_vsoftcore(const orc_grid *g, int dim, double R, double U, double *vt);
void   orc_convolve(const orc_grid *g, const double *vt, const double *dens, double *phi);
double orc_quench_a(double t, const double prms[6]);

/* ---- Hamiltonian application ------------------------------------------- */
/* psi, V, hmu, hen: ncomp consecutive complex arrays of N points each.
   hen may be NULL (the "MU" kernels, alg_calcHpsiMU: no moving-frame term);
   with hen the full form (alg_calcHpsi: with it). */
void   orc_hpsi(const orc_grid *g, const orc_ham *h, double t,
                const double *psi, const double *V, double *hmu, double *hen);
/* The same with the moving-frame term chosen explicitly. */
void   orc_hpsi_vpx(const orc_grid *g, const orc_ham *h, double t, const double *psi,
                    const double *V, double *hmu, double *hen, int vpx);

/* ---- reductions --------------------------------------------------------- */
double orc_norm(const orc_grid *g, const double *psi);
void   orc_obs2(const orc_grid *g, double norm, const double *psi,
                const double *o1psi, const double *o2psi, double *o1, double *o2);

/* ---- time loops --------------------------------------------------------- */
/* Real-time RK4 (coef = -i dt) or imaginary-time RK4 (coef = -dt, renorm) */
void   orc_rk4(const orc_grid *g, const orc_ham *h, double *psi, const double *V,
               double dt, double t0, int nsteps, int imaginary, double npart);
/* Heavy-ball imaginary time of the one- and two-component solvers.  Returns
   the number of iterations run; hist (3*ncomp*niter doubles or NULL) gets
   norm, mu, E per component per iteration. */
int    orc_imag_heavy_ball(const orc_grid *g, const orc_ham *h, double *psi,
                           const double *V, double dt, double beta, double npart,
                           double eps, int niter, double *hist);

int    orc_num_threads(void);
void   orc_set_threads(int n);

/* ---- initial state and trap (fgpe_oracle_init.c) ------------------------ */
typedef struct orc_geom {
    double xmax[3];          /* cell centres -xmax + i dx + dx/2 per axis     */
    double dx[3];
    int n[3];
} orc_geom;

#include <stdint.h>
void   orc_philox4x32_10(const uint32_t ctr[4], const uint32_t key[2], uint32_t out[4]);
void   orc_noise(long long i, uint32_t stream, uint64_t seed, double *u1, double *u2);
/* kind 0 harmonic, 1 tube, 2 void */
void   orc_psi_init(const orc_geom *g, int kind, const double ax[3], double rr, double ci,
                    uint64_t seed, double *psi);
void   orc_psi_add_noise(const orc_geom *g, double rr, double ci, uint64_t seed, uint32_t stream,
                         double *psi);
void   orc_psi_vortex(const orc_geom *g, double ncirc, double *psi);
/* lattice 0 1D, 1 2D triangular, 2 2D square, 3 HEX, 4 FCC, 5 BCC, 6 SC */
double orc_force(int lattice, double x, double y, double z, double amp, double k);
void   orc_psi_force(const orc_geom *g, int lattice, double amp, double k, const double off[3],
                     double npart, double *psi);
double orc_psi_normalize(const orc_geom *g, double npart, double *psi);
void   orc_vext_init(const orc_geom *g, int kind, const double ax[3], double *v);
void   orc_vext_box(const orc_geom *g, double V, double B, const double w[3], double *v);
void   orc_vext_add_cos(const orc_geom *g, double A, double n, double *v);

#ifdef __cplusplus
}
#endif
#endif
