/*
 * fgpe_oracle.c -- CPU FP64 restatement of the reference's time-stepping hot
 * path, written from the reference's CUDA sources (cited per function).
 *
 * TEST INFRASTRUCTURE ONLY: the checker for the HIP path (tests/, smoke(),
 * bench.py cpu_baseline).  Nothing in the product links it.
 *
 * Pinning (each part is checked by tests/test_oracle.py against something
 * other than its own formula):
 *   orc_twiddle        mpmath cos/sin at 50 digits
 *   orc_fft1d/3d       brute-force O(N^2) DFT, numpy.fft golden vectors
 *                      (tests/golden), Parseval, F^-1 F = N I
 *   orc_kaxis          integer closed form of src/baseHamiltonian.cu:61-68
 *   orc_laplace        plane-wave eigenvalue 19.739208802*|m/L|^2
 *   orc_vdd            closed form (4 pi a_dd)(3 cos^2 - 1), Vt(0) = 0
 *   orc_vsoftcore      closed form and k->0 limits; 2D J1 against
 *                      scipy.special.j1 golden values
 *   orc_convolve       single-mode density gives eps*Vt(q)*cos(q.x) exactly
 *   orc_hpsi           plane wave with g != 0 (scalar closed form)
 *   orc_rk4            per-mode amplification 1+z+z^2/2+z^3/6+z^4/24 (g=0),
 *                      scalar RK4 of a uniform state with g != 0
 *   orc_imag_heavy_ball  scalar recursion of a uniform state, harmonic
 *                      ground-state energy
 *   orc_norm/obs2      Gaussian sums against long-double sums
 */
#include "fgpe_oracle.h"

#include <math.h>
#include <stdlib.h>
#include <string.h>
#ifdef _OPENMP
#include <omp.h>
#endif

#define ORC_PI 3.141592653589793   /* hpp/header.hpp:28 */

void orc_set_threads(int n)
{
#ifdef _OPENMP
    if (n > 0) omp_set_num_threads(n);
#else
    (void)n;
#endif
}

int orc_num_threads(void)
{
#ifdef _OPENMP
    return omp_get_max_threads();
#else
    return 1;
#endif
}

/* k axis in cycles per length: src/baseHamiltonian.cu:61-76.  k[i] = i*dk for
   i <= n/2 (so the Nyquist bin is +kmax), -kmax + (i-n/2)*dk above; an axis
   that is not active (1D/2D grids) is all zero. */
void orc_kaxis(int n, double dk, double kmax, int active, double *k)
{
    for (int i = 0; i <= n / 2 && i < n; ++i) k[i] = i * dk;
    for (int i = n / 2 + 1; i < n; ++i) k[i] = -kmax + (i - n / 2) * dk;
    if (!active)
        for (int i = 0; i < n; ++i) k[i] = 0.0;
}

/* cos and sin of 2*pi*m/n, evaluated in long double so that the double
   result is correctly rounded except in rare ties. */
void orc_twiddle(int n, int m, double *c, double *s)
{
    m %= n;
    if (m < 0) m += n;
    long double pil = 3.141592653589793238462643383279502884L;
    long double th = 2.0L * pil * (long double)m / (long double)n;
    *c = (double)cosl(th);
    *s = (double)sinl(th);
}

/* ---------------------------------------------------------------------- */
/* 1D FFT: recursive decimation-in-time mixed radix (Cooley-Tukey), with a
   twiddle table of the full length.  Unnormalised; sign = -1 is the forward
   transform exp(-2 pi i jk/n) (cufftExecZ2Z CUFFT_FORWARD as called at
   src/baseHamiltonian.cu:89), sign = +1 the inverse (:95). */
typedef struct { int n; double *wr, *wi; } orc_plan;

static int smallest_factor(int n)
{
    if (n % 4 == 0) return 4;
    if (n % 2 == 0) return 2;
    for (int p = 3; p * p <= n; p += 2)
        if (n % p == 0) return p;
    return n;
}

static void fft_rec(const orc_plan *pl, int n, const double *x, int stride,
                    double *out, double sgn, double *tmp)
{
    if (n == 1) { out[0] = x[0]; out[1] = x[1]; return; }
    int p = smallest_factor(n);
    int m = n / p;
    int N = pl->n;
    for (int q = 0; q < p; ++q)
        fft_rec(pl, m, x + 2 * q * stride, stride * p, out + 2 * q * m, sgn, tmp);
    /* combine: X[k + s m] = sum_q W_n^{q k} Y_q[k] W_p^{q s} */
    for (int k = 0; k < m; ++k) {
        for (int q = 0; q < p; ++q) {
            int idx = (int)(((long long)q * k % n) * (N / n));
            double wr = pl->wr[idx], wi = sgn * pl->wi[idx];
            double yr = out[2 * (q * m + k)], yi = out[2 * (q * m + k) + 1];
            tmp[2 * q] = yr * wr - yi * wi;
            tmp[2 * q + 1] = yr * wi + yi * wr;
        }
        for (int s = 0; s < p; ++s) {
            double ar = 0.0, ai = 0.0;
            for (int q = 0; q < p; ++q) {
                int idx = (int)(((long long)q * s % p) * (N / p));
                double wr = pl->wr[idx], wi = sgn * pl->wi[idx];
                ar += tmp[2 * q] * wr - tmp[2 * q + 1] * wi;
                ai += tmp[2 * q] * wi + tmp[2 * q + 1] * wr;
            }
            tmp[2 * p + 2 * s] = ar;
            tmp[2 * p + 2 * s + 1] = ai;
        }
        for (int s = 0; s < p; ++s) {
            out[2 * (k + s * m)] = tmp[2 * p + 2 * s];
            out[2 * (k + s * m) + 1] = tmp[2 * p + 2 * s + 1];
        }
    }
}

static orc_plan plan_make(int n)
{
    orc_plan pl;
    pl.n = n;
    pl.wr = (double *)malloc(sizeof(double) * n);
    pl.wi = (double *)malloc(sizeof(double) * n);
    /* W_n^m = exp(-2 pi i m / n); the sign is applied per call */
    for (int m = 0; m < n; ++m) {
        double c, s;
        orc_twiddle(n, m, &c, &s);
        pl.wr[m] = c;
        pl.wi[m] = -s;
    }
    return pl;
}

static void plan_free(orc_plan *pl) { free(pl->wr); free(pl->wi); }

static void fft1d_with_plan(const orc_plan *pl, double *data, double *work,
                            double *tmp, int sign)
{
    /* forward uses W = exp(-i...), the table holds exactly that */
    double sgn = (sign < 0) ? 1.0 : -1.0;
    fft_rec(pl, pl->n, data, 1, work, sgn, tmp);
    memcpy(data, work, sizeof(double) * 2 * pl->n);
}

void orc_fft1d(int n, double *data, int sign)
{
    orc_plan pl = plan_make(n);
    double *work = (double *)malloc(sizeof(double) * 2 * n);
    double *tmp = (double *)malloc(sizeof(double) * 4 * (n + 1));
    fft1d_with_plan(&pl, data, work, tmp, sign);
    free(work); free(tmp);
    plan_free(&pl);
}

/* transform every pencil along one axis (stride `st`, length n) */
static void fft_axis(double *data, int n, long long st, long long npencil,
                     long long (*base)(long long, const int *), const int *dims,
                     int sign)
{
    if (n == 1) return;
    orc_plan pl = plan_make(n);
    #pragma omp parallel
    {
        double *line = (double *)malloc(sizeof(double) * 2 * n);
        double *work = (double *)malloc(sizeof(double) * 2 * n);
        double *tmp = (double *)malloc(sizeof(double) * 4 * (n + 1));
        #pragma omp for schedule(static)
        for (long long p = 0; p < npencil; ++p) {
            long long b = base(p, dims);
            for (int i = 0; i < n; ++i) {
                line[2 * i] = data[2 * (b + i * st)];
                line[2 * i + 1] = data[2 * (b + i * st) + 1];
            }
            fft1d_with_plan(&pl, line, work, tmp, sign);
            for (int i = 0; i < n; ++i) {
                data[2 * (b + i * st)] = line[2 * i];
                data[2 * (b + i * st) + 1] = line[2 * i + 1];
            }
        }
        free(line); free(work); free(tmp);
    }
    plan_free(&pl);
}

static long long base_x(long long p, const int *d) { (void)d; return p * d[0]; }
static long long base_y(long long p, const int *d)
{   /* p = z*nx + x */
    long long x = p % d[0], z = p / d[0];
    return z * (long long)d[0] * d[1] + x;
}
static long long base_z(long long p, const int *d) { (void)d; return p; }

/* 3D C2C transform of the (NZ,NY,NX) row-major array, x fastest: the plan of
   src/parallel.cu:97 (cufftPlan3d(NZ,NY,NX,Z2Z)).  Unnormalised. */
void orc_fft3d(int nx, int ny, int nz, double *data, int sign)
{
    int d[3] = {nx, ny, nz};
    fft_axis(data, nx, 1, (long long)ny * nz, base_x, d, sign);
    fft_axis(data, ny, nx, (long long)nx * nz, base_y, d, sign);
    fft_axis(data, nz, (long long)nx * ny, (long long)nx * ny, base_z, d, sign);
}

/* ---------------------------------------------------------------------- */
/* alg_Laplace (src/baseHamiltonian.cu:86-96) with the Laplace kernel
   (src/parallel.cu:194-205): out = F^-1[ (19.739208802/N) |k|^2 F psi ].
   The truncated literal is the reference's, not 2 pi^2. */
void orc_laplace(const orc_grid *g, const double *psi, double *out)
{
    long long N = (long long)g->nx * g->ny * g->nz;
    memcpy(out, psi, sizeof(double) * 2 * N);
    orc_fft3d(g->nx, g->ny, g->nz, out, -1);
    double scale = 19.739208802 / (double)N;
    #pragma omp parallel for schedule(static)
    for (long long idx = 0; idx < N; ++idx) {
        int ix = (int)(idx % g->nx);
        int iy = (int)((idx / g->nx) % g->ny);
        int iz = (int)(idx / ((long long)g->nx * g->ny));
        double k2 = g->kx[ix] * g->kx[ix] + g->ky[iy] * g->ky[iy] + g->kz[iz] * g->kz[iz];
        out[2 * idx] = (scale * out[2 * idx]) * k2;
        out[2 * idx + 1] = (scale * out[2 * idx + 1]) * k2;
    }
    orc_fft3d(g->nx, g->ny, g->nz, out, +1);
}

/* DipoleDipoleInteraction (src/parallel.cu:326-352):
   Vt = 12 pi a_dd ((k.d/|k|)^2 - 1/3), 0 where |k| < 1e-6. */
void orc_vdd(const orc_grid *g, double a_dd, const double d[3], double *vt)
{
    long long N = (long long)g->nx * g->ny * g->nz;
    for (long long idx = 0; idx < N; ++idx) {
        int ix = (int)(idx % g->nx);
        int iy = (int)((idx / g->nx) % g->ny);
        int iz = (int)(idx / ((long long)g->nx * g->ny));
        double num = g->kx[ix] * d[0] + g->ky[iy] * d[1] + g->kz[iz] * d[2];
        double den = sqrt(g->kx[ix] * g->kx[ix] + g->ky[iy] * g->ky[iy] + g->kz[iz] * g->kz[iz]);
        if (den < 1E-6) vt[idx] = 0.0;
        else vt[idx] = 12.0 * ORC_PI * a_dd * (((num / den) * (num / den)) - 1.0 / 3.0);
    }
}

/* SoftCoreInteraction_1D/2D/3D (src/parallel.cu:371-397, 411-437, 451-477)
   with k = 2 pi |k|, a = R (prms[0]) and U (prms[1]). */
void orc_vsoftcore(const orc_grid *g, int dim, double R, double U, double *vt)
{
    long long N = (long long)g->nx * g->ny * g->nz;
    double a = R;
    for (long long idx = 0; idx < N; ++idx) {
        int ix = (int)(idx % g->nx);
        int iy = (int)((idx / g->nx) % g->ny);
        int iz = (int)(idx / ((long long)g->nx * g->ny));
        double k = 2 * ORC_PI * sqrt(g->kx[ix] * g->kx[ix] + g->ky[iy] * g->ky[iy] + g->kz[iz] * g->kz[iz]);
        double v;
        if (dim == 1) {
            if (k < 1E-6) v = U * 2 * a * (1 - (a * k) * (a * k) / 6.0);
            else v = U * 2 * sin(a * k) / (k);
        } else if (dim == 2) {
            if (k < 1E-6) v = U * ORC_PI * a * a;
            else v = 2 * ORC_PI * a * U * j1(k * a) / k;
        } else {
            if (k < 1E-6) v = U * (4. / 3.) * ORC_PI * pow(a, 3);
            else v = 4 * ORC_PI * U * pow(a, 3) * (sin(k * a) - k * a * cos(k * a)) / (pow(k * a, 3));
        }
        vt[idx] = v;
    }
}

/* k-space convolution Phi = F^-1[ Vt(k) F(dens) ] / N, real part.  Absent
   from the reference (its Vt generators have no call site); written from the
   textbook definition with the unnormalised transform pair and 1/N folded in
   the k-space multiply, as Laplace does (src/parallel.cu:199). */
void orc_convolve(const orc_grid *g, const double *vt, const double *dens, double *phi)
{
    long long N = (long long)g->nx * g->ny * g->nz;
    double *w = (double *)malloc(sizeof(double) * 2 * N);
    for (long long i = 0; i < N; ++i) { w[2 * i] = dens[i]; w[2 * i + 1] = 0.0; }
    orc_fft3d(g->nx, g->ny, g->nz, w, -1);
    double invN = 1.0 / (double)N;
    #pragma omp parallel for schedule(static)
    for (long long i = 0; i < N; ++i) {
        double f = vt[i] * invN;
        w[2 * i] *= f;
        w[2 * i + 1] *= f;
    }
    orc_fft3d(g->nx, g->ny, g->nz, w, +1);
    for (long long i = 0; i < N; ++i) phi[i] = w[2 * i];
    free(w);
}

/* d_quench_as (hamiltonians/BEC_oneComponent.cu:167-178) */
double orc_quench_a(double time, const double prms[6])
{
    double a0 = prms[0], a1 = prms[1], a2 = prms[2];
    double t1 = prms[3], t2 = prms[4], t3 = prms[5];
    if (0 < time && time <= t1) return a0 + (a1 - a0) / (t1) * time;
    if (t1 < time && time <= t2) return a1;
    if (t2 < time && time <= t3) return a1 + (a2 - a1) / (t3 - t2) * (time - t2);
    return a2;
}

/* ---------------------------------------------------------------------- */
/* H application.
   One component: alg_calcHpsi/alg_calcHpsiMU (hamiltonians/BEC_oneComponent.cu
   :92-111) = alg_Laplace + BECHamiltonian/BECHamiltonianMU (:128-165):
     hmu = KE + (V.x + 4 pi a |psi|^2 [+ Phi]) psi + i V.y psi
     hen = KE + (V.x + 2 pi a |psi|^2 [+ Phi/2]) psi + i V.y psi
   Two components: BECHamiltonian_2/BECHamiltonianMU_2
   (hamiltonians/BEC_twoComponent.cu:168-244) with couplings a11, a22, a12.
   Imag(V) of component 2 is V2.y (the reference's BECHamiltonian_2 reads
   V1.y at :208, which only matters for complex traps).
   Phi (the dipolar / soft-core term) acts on the total density. */
/* The moving-frame term of alg_addVPx (src/baseHamiltonian.cu:198-211):
   out += (-vpx / N) F^-1[ k_x F(psi) ]  (BECVPx, src/parallel.cu:283-292;
   AppendArray, hpp/parallel.hpp:231-237). */
static void add_vpx(const orc_grid *g, double vpx, const double *psi, double *out)
{
    long long N = (long long)g->nx * g->ny * g->nz;
    double *aux = (double *)malloc(sizeof(double) * 2 * N);
    memcpy(aux, psi, sizeof(double) * 2 * N);
    orc_fft3d(g->nx, g->ny, g->nz, aux, -1);
    for (long long idx = 0; idx < N; ++idx) {
        double kx = g->kx[idx % g->nx];
        aux[2 * idx] *= kx;
        aux[2 * idx + 1] *= kx;
    }
    orc_fft3d(g->nx, g->ny, g->nz, aux, +1);
    double f = -vpx / (double)N;
    for (long long i = 0; i < 2 * N; ++i) out[i] += f * aux[i];
    free(aux);
}

void orc_hpsi(const orc_grid *g, const orc_ham *h, double t,
              const double *psi, const double *V, double *hmu, double *hen)
{
    orc_hpsi_vpx(g, h, t, psi, V, hmu, hen, hen != NULL);
}

/* The test user pointwise terms of the library's user_term.hpp (not part of
   the reference: they stand for a plugin author's own kernels).
   Term 1: a modulated contact coupling, three-body loss and a ramp along z,
     w = u0 sin(u1 t) n + u3 iz - i u2 n^2 / 2, w_en with half the coupling.
   Term 2: a decaying lattice felt with opposite signs by the components,
     w = w_en = sgn_c u4 exp(-u5 t) cos(u6 ix + u7 iy). */
static void user_term(const orc_ham *h, double t, int ix, int iy, int iz, int c, double n, double *wr, double *wi,
                      double *er, double *ei)
{
    *wr = *wi = *er = *ei = 0.0;
    if (h->user_id == 1) {
        const double dg = h->user[0] * sin(h->user[1] * t);
        const double loss = -0.5 * h->user[2] * n * n;
        const double ramp = h->user[3] * iz;
        *wr = dg * n + ramp;
        *wi = loss;
        *er = 0.5 * dg * n + ramp;
        *ei = loss;
    } else if (h->user_id == 2) {
        const double sgn = c == 0 ? 1.0 : -1.0;
        const double w = sgn * h->user[4] * exp(-h->user[5] * t) * cos(h->user[6] * ix + h->user[7] * iy);
        *wr = w;
        *er = w;
    }
}

static void add_user(const orc_grid *g, const orc_ham *h, double t, int nc, const double *psi, double *hmu,
                     double *hen)
{
    const long long N = (long long)g->nx * g->ny * g->nz;
    for (int c = 0; c < nc; ++c) {
        const double *p = psi + 2 * N * c;
        for (long long i = 0; i < N; ++i) {
            const int ix = (int)(i % g->nx);
            const int iy = (int)((i / g->nx) % g->ny);
            const int iz = (int)(i / ((long long)g->nx * g->ny));
            const double pr = p[2 * i], pi_ = p[2 * i + 1];
            double wr, wi, er, ei;
            user_term(h, t, ix, iy, iz, c, pr * pr + pi_ * pi_, &wr, &wi, &er, &ei);
            double *o = hmu + 2 * N * c;
            o[2 * i] += wr * pr - wi * pi_;
            o[2 * i + 1] += wr * pi_ + wi * pr;
            if (hen) {
                double *e = hen + 2 * N * c;
                e[2 * i] += er * pr - ei * pi_;
                e[2 * i + 1] += er * pi_ + ei * pr;
            }
        }
    }
}

void orc_hpsi_vpx(const orc_grid *g, const orc_ham *h, double t, const double *psi,
                  const double *V, double *hmu, double *hen, int vpx)
{
    long long N = (long long)g->nx * g->ny * g->nz;
    int nc = h->ncomp;
    double *phi = NULL;
    if (h->vtilde) {
        double *dens = (double *)malloc(sizeof(double) * N);
        for (long long i = 0; i < N; ++i) {
            double d = 0.0;
            for (int c = 0; c < nc; ++c) {
                const double *p = psi + 2 * N * c;
                d += p[2 * i] * p[2 * i] + p[2 * i + 1] * p[2 * i + 1];
            }
            dens[i] = d;
        }
        phi = (double *)malloc(sizeof(double) * N);
        orc_convolve(g, h->vtilde, dens, phi);
        free(dens);
    }
    double *ke = (double *)malloc(sizeof(double) * 2 * N * nc);
    for (int c = 0; c < nc; ++c) {
        orc_laplace(g, psi + 2 * N * c, ke + 2 * N * c);
        /* the term lands in both H and the energy operator, like the
           reference's AppendArray on d_hpsi and d_hpsi_en */
        if (vpx && h->vpx[c] != 0.0) add_vpx(g, h->vpx[c], psi + 2 * N * c, ke + 2 * N * c);
    }

    if (nc == 1) {
        double a = h->quench_on ? orc_quench_a(t, h->quench) : h->a[0];
        for (long long i = 0; i < N; ++i) {
            double pr = psi[2 * i], pi_ = psi[2 * i + 1];
            double n = pr * pr + pi_ * pi_;
            double ph = phi ? phi[i] : 0.0;
            double rY = V[2 * i + 1];
            if (hen) {
                double rX = (V[2 * i] + 2 * ORC_PI * a * n);
                if (phi) rX += 0.5 * ph;
                hen[2 * i] = ke[2 * i] + rX * pr - rY * pi_;
                hen[2 * i + 1] = ke[2 * i + 1] + rX * pi_ + rY * pr;
            }
            double rX = (V[2 * i] + 4 * ORC_PI * a * n);
            if (phi) rX += ph;
            hmu[2 * i] = ke[2 * i] + rX * pr - rY * pi_;
            hmu[2 * i + 1] = ke[2 * i + 1] + rX * pi_ + rY * pr;
        }
    } else {
        double a11 = h->a[0], a22 = h->a[1], a12 = h->a[2];
        const double *p1 = psi, *p2 = psi + 2 * N;
        for (long long i = 0; i < N; ++i) {
            double n1 = p1[2 * i] * p1[2 * i] + p1[2 * i + 1] * p1[2 * i + 1];
            double n2 = p2[2 * i] * p2[2 * i] + p2[2 * i + 1] * p2[2 * i + 1];
            double ph = phi ? phi[i] : 0.0;
            for (int c = 0; c < 2; ++c) {
                const double *p = psi + 2 * N * c;
                const double *v = V + 2 * N * c;
                const double *k = ke + 2 * N * c;
                double aself = c == 0 ? a11 : a22;
                double nself = c == 0 ? n1 : n2, nother = c == 0 ? n2 : n1;
                double pr = p[2 * i], pi_ = p[2 * i + 1];
                double rY = v[2 * i + 1];
                if (hen) {
                    double *o = hen + 2 * N * c;
                    double rX = (v[2 * i] + 2 * ORC_PI * aself * nself
                                 + 2 * ORC_PI * a12 * nother);
                    if (phi) rX += 0.5 * ph;
                    o[2 * i] = k[2 * i] + rX * pr - rY * pi_;
                    o[2 * i + 1] = k[2 * i + 1] + rX * pi_ + rY * pr;
                }
                double *o = hmu + 2 * N * c;
                double rX = (v[2 * i] + 4 * ORC_PI * aself * nself
                             + 4 * ORC_PI * a12 * nother);
                if (phi) rX += ph;
                o[2 * i] = k[2 * i] + rX * pr - rY * pi_;
                o[2 * i + 1] = k[2 * i + 1] + rX * pi_ + rY * pr;
            }
        }
    }
    if (h->user_id) add_user(g, h, t, nc, psi, hmu, hen);
    free(ke);
    free(phi);
}

/* ---------------------------------------------------------------------- */
/* alg_calcNorm (src/baseHamiltonian.cu:173-179, NormalizePsi at
   src/parallel.cu:499-527): sum |psi|^2 times DV.  Summed in long double. */
double orc_norm(const orc_grid *g, const double *psi)
{
    long long N = (long long)g->nx * g->ny * g->nz;
    long double s = 0.0L;
    for (long long i = 0; i < N; ++i)
        s += (long double)(psi[2 * i] * psi[2 * i] + psi[2 * i + 1] * psi[2 * i + 1]);
    return (double)s * g->dv;
}

/* alg_calc2Observables (src/baseHamiltonian.cu:230-236 with
   Calculate2Observables at src/parallel.cu:613-645 and _simple_reduction2
   :125-138): o = sum Re(psi* O psi) * DV / norm. */
void orc_obs2(const orc_grid *g, double norm, const double *psi,
              const double *o1psi, const double *o2psi, double *o1, double *o2)
{
    long long N = (long long)g->nx * g->ny * g->nz;
    long double s1 = 0.0L, s2 = 0.0L;
    for (long long i = 0; i < N; ++i) {
        s1 += (long double)(psi[2 * i] * o1psi[2 * i] + psi[2 * i + 1] * o1psi[2 * i + 1]);
        s2 += (long double)(psi[2 * i] * o2psi[2 * i] + psi[2 * i + 1] * o2psi[2 * i + 1]);
    }
    *o1 = (double)s1 * g->dv / norm;
    *o2 = (double)s2 * g->dv / norm;
}

/* UpdateRKStep (src/parallel.cu:238-244): psi = old + c*h, c complex */
static void rk_update(long long n2, double *psi, const double *old,
                      const double *hp, double cr, double ci)
{
    for (long long i = 0; i < n2; ++i) {
        double hx = hp[2 * i], hy = hp[2 * i + 1];
        psi[2 * i] = old[2 * i] + (cr * hx - ci * hy);
        psi[2 * i + 1] = old[2 * i + 1] + (cr * hy + ci * hx);
    }
}

/* Real-time RK4 of src/oneComponentGPSolver.cu:122-217 (two components:
   src/twoComponentGPSolver.cu:120-209, all stages with the MU kernel):
     old = psi; H1 = H(psi, t); k = H1; psi = old + c/2 H1
     H2 = H(psi, t+dt/2); k += 2 H2; psi = old + c/2 H2
     H3 = H(psi, t+dt/2); k += 2 H3; psi = old + c H3
     H4 = H(psi, t+dt);   psi = old + c (k + H4)/6         (FinalRKStep :261-271)
   with c = -i dt.  imaginary != 0 runs the same scheme with c = -dt and
   renormalises to npart after every step (the rk4_renorm imaginary-time
   scheme; the reference's imaginary time is heavy-ball, see below). */
void orc_rk4(const orc_grid *g, const orc_ham *h, double *psi, const double *V,
             double dt, double t0, int nsteps, int imaginary, double npart)
{
    long long N = (long long)g->nx * g->ny * g->nz;
    long long n2 = N * h->ncomp;   /* complex points across components */
    size_t bytes = sizeof(double) * 2 * n2;
    double *old = (double *)malloc(bytes);
    double *k = (double *)malloc(bytes);
    double *hp = (double *)malloc(bytes);
    double hr, hi, fr, fi;
    if (imaginary) { hr = -0.5 * dt; hi = 0.0; fr = -1.0 * dt; fi = 0.0; }
    else           { hr = 0.0; hi = -0.5 * dt; fr = 0.0; fi = -1.0 * dt; }
    for (int it = 0; it < nsteps; ++it) {
        double t = t0 + it * dt;
        memcpy(old, psi, bytes);
        orc_hpsi(g, h, t, psi, V, hp, NULL);
        memcpy(k, hp, bytes);
        rk_update(n2, psi, old, hp, hr, hi);

        t = t0 + (it + 0.5) * dt;
        orc_hpsi(g, h, t, psi, V, hp, NULL);
        for (long long i = 0; i < 2 * n2; ++i) k[i] = k[i] + 2.0 * hp[i];
        rk_update(n2, psi, old, hp, hr, hi);

        orc_hpsi(g, h, t, psi, V, hp, NULL);
        for (long long i = 0; i < 2 * n2; ++i) k[i] = k[i] + 2.0 * hp[i];
        rk_update(n2, psi, old, hp, fr, fi);

        t = t0 + (it + 1) * dt;
        /* stage 4 of the one-component step is alg_calcHpsi (with the
           moving-frame term), stages 1-3 alg_calcHpsiMU
           (src/oneComponentGPSolver.cu:170-187); two components use the MU
           form throughout (src/twoComponentGPSolver.cu:157-175) */
        orc_hpsi_vpx(g, h, t, psi, V, hp, NULL, h->ncomp == 1);
        for (long long i = 0; i < n2; ++i) {
            double aX = k[2 * i] + hp[2 * i];
            double aY = k[2 * i + 1] + hp[2 * i + 1];
            psi[2 * i] = old[2 * i] + (fr * aX - fi * aY) / 6.0;
            psi[2 * i + 1] = old[2 * i + 1] + (fr * aY + fi * aX) / 6.0;
        }
        if (imaginary) {
            for (int c = 0; c < h->ncomp; ++c) {
                double *p = psi + 2 * N * c;
                double nrm = orc_norm(g, p);
                double s = pow(npart / nrm, 0.5);
                for (long long i = 0; i < 2 * N; ++i) p[i] = s * p[i];
            }
        }
    }
    free(old); free(k); free(hp);
}

/* Imaginary time (heavy ball).  One component, src/oneComponentGPSolver.cu:26-104:
     H, Hen = H(psi, it*dt)                         alg_calcHpsi
     new = (1 + 0*mu dt) psi - dt H + beta (psi - old)   BECUpdatePsi
                                                    (src/parallel.cu:294-302)
     norm = |new|^2 DV; old = psi; psi = sqrt(npart/norm) new
     mu, E = <psi|H>, <psi|Hen> * DV / npart        alg_calc2Observables
     stop when |E - E_prev| / E_prev < eps          (:84-88, no abs on E_prev)
   Two components, src/twoComponentGPSolver.cu:15-112 (PROPER_2BEC_NORMALIZATION
   undefined): the update uses mu_c of the previous iteration, each component
   is renormalised to npart on its own, mu_c and E_c divide by the unscaled
   norm_c (PARS[NORM_c] is not reset, :67-69), and the loop stops when
   |E_c - E_c,prev| < eps for both (:94).
   Returns iterations run (the converging one included); hist gets
   (norm, mu, E) per component per iteration. */
int orc_imag_heavy_ball(const orc_grid *g, const orc_ham *h, double *psi,
                        const double *V, double dt, double beta, double npart,
                        double eps, int niter, double *hist)
{
    long long N = (long long)g->nx * g->ny * g->nz;
    int nc = h->ncomp;
    size_t bytes = sizeof(double) * 2 * N * nc;
    double *old = (double *)malloc(bytes);
    double *nw = (double *)malloc(bytes);
    double *hp = (double *)malloc(bytes);
    double *he = (double *)malloc(bytes);
    memcpy(old, psi, bytes);
    double en[2] = {0.0, 0.0}, mu[2] = {0.0, 0.0};
    int it;
    for (it = 0; it < niter; ++it) {
        double en_prev[2] = {en[0], en[1]};
        double t = it * dt;
        orc_hpsi(g, h, t, psi, V, hp, he);
        for (int c = 0; c < nc; ++c) {
            double m = nc == 1 ? 0 * mu[0] : mu[c];
            long long o = 2 * N * c;
            for (long long i = 0; i < 2 * N; ++i)
                nw[o + i] = (1.0 + m * dt) * psi[o + i] - dt * hp[o + i] + beta * (psi[o + i] - old[o + i]);
        }
        memcpy(old, psi, bytes);
        for (int c = 0; c < nc; ++c) {
            long long o = 2 * N * c;
            double nrm = orc_norm(g, nw + o);
            double s = pow(npart / nrm, 0.5);
            for (long long i = 0; i < 2 * N; ++i) psi[o + i] = s * nw[o + i];
            double narg = nc == 1 ? npart : nrm;
            orc_obs2(g, narg, psi + o, hp + o, he + o, &mu[c], &en[c]);
            if (hist) {
                hist[3 * (nc * it + c)] = nrm;
                hist[3 * (nc * it + c) + 1] = mu[c];
                hist[3 * (nc * it + c) + 2] = en[c];
            }
        }
        int conv;
        if (nc == 1) conv = fabs(en[0] - en_prev[0]) / en_prev[0] < eps;
        else conv = (fabs(en[0] - en_prev[0]) < eps) && (fabs(en[1] - en_prev[1]) < eps);
        if (conv) { ++it; break; }
    }
    free(old); free(nw); free(hp); free(he);
    return it;
}
