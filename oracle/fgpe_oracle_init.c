/*
 * fgpe_oracle_init.c -- CPU FP64 restatement of the reference's initial
 * state and trap builders (Psi, src/psi.cu:80-372; Vext, src/vext.cu:60-199)
 * with the device path's noise generator.
 *
 * TEST INFRASTRUCTURE ONLY: the checker for libfgpe's fgpe_psi_* and
 * fgpe_vext_* kernels.  Nothing in the product links it.
 *
 * The loops run in the reference's operation order, built with
 * -ffp-contract=off like the reference's host code.  The noise is
 * Philox4x32-10 (Salmon, Moraes, Dror, Shaw, SC'11), which the device path
 * uses in place of the reference's sequential std::mt19937 draws.
 *
 * Pinning (tests/test_oracle_init.py):
 *   orc_philox4x32_10  the published Random123 known-answer vectors
 *   orc_psi_init       the reference-shaped C++ host classes (line-by-line
 *                      mirror of src/psi.cu, zero noise), mirror symmetry of
 *                      the cell centres, Gaussian widths per axis
 *   orc_psi_normalize  sum |psi|^2 dv = npart afterwards; shape unchanged
 *   orc_psi_vortex     |psi| unchanged; phase winds ncirc times along x
 *   orc_psi_force      the C++ host classes; amp at the lattice origin;
 *                      periodicity 2 pi / k
 *   orc_vext_*         the C++ host classes; closed-form curvature
 */
#include <math.h>
#include <stdint.h>

#include "fgpe_oracle.h"

#define ORC_PI 3.141592653589793   /* hpp/header.hpp:28 */

void orc_philox4x32_10(const uint32_t ctr[4], const uint32_t key[2], uint32_t out[4])
{
    uint32_t c0 = ctr[0], c1 = ctr[1], c2 = ctr[2], c3 = ctr[3];
    uint32_t k0 = key[0], k1 = key[1];
    for (int r = 0; r < 10; ++r) {
        if (r > 0) {
            k0 += 0x9E3779B9u;
            k1 += 0xBB67AE85u;
        }
        const uint64_t p0 = (uint64_t)0xD2511F53u * c0;
        const uint64_t p1 = (uint64_t)0xCD9E8D57u * c2;
        const uint32_t n0 = (uint32_t)(p1 >> 32) ^ c1 ^ k0;
        const uint32_t n2 = (uint32_t)(p0 >> 32) ^ c3 ^ k1;
        c0 = n0;
        c1 = (uint32_t)p1;
        c2 = n2;
        c3 = (uint32_t)p0;
    }
    out[0] = c0;
    out[1] = c1;
    out[2] = c2;
    out[3] = c3;
}

/* uniform on [-1, 1) from 53 bits of (hi:lo) */
static double pm1(uint32_t lo, uint32_t hi)
{
    const uint64_t r = (((uint64_t)hi << 32) | lo) >> 11;
    return (double)r * 0x1p-52 - 1.0;
}

void orc_noise(long long i, uint32_t stream, uint64_t seed, double *u1, double *u2)
{
    const uint32_t ctr[4] = {(uint32_t)i, (uint32_t)((uint64_t)i >> 32), stream, 0u};
    const uint32_t key[2] = {(uint32_t)seed, (uint32_t)(seed >> 32)};
    uint32_t o[4];
    orc_philox4x32_10(ctr, key, o);
    *u1 = pm1(o[0], o[1]);
    *u2 = pm1(o[2], o[3]);
}

/* cell centre of index i on axis a (src/psi.cu:134-150) */
static double centre(const orc_geom *g, int a, int i) { return -g->xmax[a] + i * g->dx[a] + 0.5 * g->dx[a]; }

static long long npts(const orc_geom *g) { return (long long)g->n[0] * g->n[1] * g->n[2]; }

static void index3(const orc_geom *g, long long i, int *ix, int *iy, int *iz)
{
    *ix = (int)(i % g->n[0]);
    *iy = (int)((i / g->n[0]) % g->n[1]);
    *iz = (int)(i / ((long long)g->n[0] * g->n[1]));
}

/* Psi::initHarmonic / initTube / initVoid (src/psi.cu:80-131) */
void orc_psi_init(const orc_geom *g, int kind, const double ax[3], double rr, double ci, uint64_t seed, double *psi)
{
    const long long n = npts(g);
#pragma omp parallel for schedule(static)
    for (long long i = 0; i < n; ++i) {
        int ix, iy, iz;
        index3(g, i, &ix, &iy, &iz);
        double u1, u2;
        orc_noise(i, 0u, seed, &u1, &u2);
        /* (complex(1,0) + noise) */
        const double a = 1.0 + rr * u1, b = 0.0 + ci * u2;
        double re = a, im = b;
        if (kind != 2) {
            const double y = centre(g, 1, iy) / ax[1];
            const double z = centre(g, 2, iz) / ax[2];
            double q;
            if (kind == 0) {
                const double x = centre(g, 0, ix) / ax[0];
                q = (y * y) + (z * z) + (x * x);
            } else {
                q = (y * y) + (z * z);
            }
            const double e = exp(-0.5 * 0.1 * q);
            /* (a + ib) * (e + i0), std::complex product order */
            re = a * e - b * 0.0;
            im = a * 0.0 + b * e;
        }
        psi[2 * i] = re;
        psi[2 * i + 1] = im;
    }
}

/* Psi::addNoise (src/psi.cu:167-176) with Philox stream `stream` */
void orc_psi_add_noise(const orc_geom *g, double rr, double ci, uint64_t seed, uint32_t stream, double *psi)
{
    const long long n = npts(g);
#pragma omp parallel for schedule(static)
    for (long long i = 0; i < n; ++i) {
        double u1, u2;
        orc_noise(i, stream, seed, &u1, &u2);
        const double a = 1.0 + rr * u1, b = 0.0 + ci * u2;
        const double c = psi[2 * i], d = psi[2 * i + 1];
        psi[2 * i] = a * c - b * d;
        psi[2 * i + 1] = a * d + b * c;
    }
}

/* Psi::imprintVortexTube (src/psi.cu:155-164) */
void orc_psi_vortex(const orc_geom *g, double ncirc, double *psi)
{
    const double xlen = 2 * g->xmax[0];
    const long long n = npts(g);
#pragma omp parallel for schedule(static)
    for (long long i = 0; i < n; ++i) {
        int ix, iy, iz;
        index3(g, i, &ix, &iy, &iz);
        const double x = centre(g, 0, ix);
        const double c = cos(ncirc * 2 * ORC_PI / xlen * x), s = sin(ncirc * 2 * ORC_PI / xlen * x);
        const double re = psi[2 * i], im = psi[2 * i + 1];
        psi[2 * i] = c * re - s * im;
        psi[2 * i + 1] = c * im + s * re;
    }
}

/* force profile of each lattice (src/psi.cu:179-347) */
static double force_of(int lat, double x, double y, double z, double amp, double k)
{
    const double r3 = sqrt(3.0), r2 = sqrt(2.0);
    switch (lat) {
    case 0:
        return amp * (1 - cos(k * x)) / 2.;
    case 1:
        return amp * (cos(k * x) + cos(k * (-0.5 * x + (r3 / 2) * y)) + cos(k * (-0.5 * x - (r3 / 2) * y))) / 3;
    case 2:
        return amp * (cos(k * y) + cos(k * x)) / 2;
    case 3: {
        const double x1 = (-x + r3 * y) / 2., x4 = (-x - r3 * y) / 2., x2 = x, x3 = z;
        return amp * (cos(k * x1) + cos(k * x2) + cos(k * x3) + cos(k * x4)) / 4.0;
    }
    case 4: {
        const double x1 = (x + y - z) / r3, x2 = (y + z - x) / r3, x3 = (z + x - y) / r3;
        const double x4 = (x - y - z) / r3, x5 = (y - z - x) / r3, x6 = (z - x - y) / r3;
        const double x7 = (x + y + z) / r3, x8 = (y + z + x) / r3, x9 = (z + x + y) / r3;
        return amp *
               (cos(k * x1) + cos(k * x2) + cos(k * x3) + cos(k * x4) + cos(k * x5) + cos(k * x6) + cos(k * x7) +
                cos(k * x8) + cos(k * x9)) /
               9.0;
    }
    case 5: {
        const double x1 = (x + y) / r2, x2 = (z + y) / r2, x3 = (z + x) / r2;
        const double x4 = (x - y) / r2, x5 = (z - y) / r2, x6 = (z - x) / r2;
        return amp * (cos(k * x1) + cos(k * x2) + cos(k * x3) + cos(k * x4) + cos(k * x5) + cos(k * x6)) / 6;
    }
    default:
        return amp * (cos(k * z) + cos(k * y) + cos(k * x)) / 3;
    }
}

double orc_force(int lattice, double x, double y, double z, double amp, double k)
{
    return force_of(lattice, x, y, z, amp, k);
}

/* psi = force * avg + psi with avg = sqrt(npart / dv) (src/psi.cu:179-347,
   without the trailing normalize) */
void orc_psi_force(const orc_geom *g, int lattice, double amp, double k, const double off[3], double npart,
                   double *psi)
{
    const double avg = pow(npart / (g->dx[0] * g->dx[1] * g->dx[2]), 0.5);
    const long long n = npts(g);
#pragma omp parallel for schedule(static)
    for (long long i = 0; i < n; ++i) {
        int ix, iy, iz;
        index3(g, i, &ix, &iy, &iz);
        const double x = centre(g, 0, ix) - off[0], y = centre(g, 1, iy) - off[1], z = centre(g, 2, iz) - off[2];
        psi[2 * i] = force_of(lattice, x, y, z, amp, k) * avg + psi[2 * i];
    }
}

/* Psi::normalize (src/psi.cu:351-372); returns sum |psi|^2 dx dy dz after */
double orc_psi_normalize(const orc_geom *g, double npart, double *psi)
{
    const long long n = npts(g);
    double norm = 0;
    for (long long i = 0; i < n; ++i) norm += psi[2 * i] * psi[2 * i] + psi[2 * i + 1] * psi[2 * i + 1];
    const double f = pow(npart / (norm * g->dx[0] * g->dx[1] * g->dx[2]), 0.5);
    for (long long i = 0; i < 2 * n; ++i) psi[i] = psi[i] * f;
    norm = 0;
    for (long long i = 0; i < n; ++i)
        norm += (psi[2 * i] * psi[2 * i] + psi[2 * i + 1] * psi[2 * i + 1]) * g->dx[0] * g->dx[1] * g->dx[2];
    return norm;
}

/* Vext::initHarmonic / initTube / initVoid (src/vext.cu:60-118), real part */
void orc_vext_init(const orc_geom *g, int kind, const double ax[3], double *v)
{
    const long long n = npts(g);
#pragma omp parallel for schedule(static)
    for (long long i = 0; i < n; ++i) {
        int ix, iy, iz;
        index3(g, i, &ix, &iy, &iz);
        double r = 0;
        if (kind != 2) {
            const double y = centre(g, 1, iy) / ax[1], z = centre(g, 2, iz) / ax[2];
            if (kind == 0) {
                const double x = centre(g, 0, ix) / ax[0];
                r = 0.5 * ((x * x) / (ax[0] * ax[0]) + (y * y) / (ax[1] * ax[1]) + (z * z) / (ax[2] * ax[2]));
            } else {
                r = 0.5 * ((y * y) / (ax[1] * ax[1]) + (z * z) / (ax[2] * ax[2]));
            }
        }
        v[i] = r;
    }
}

/* Vext::initBox (src/vext.cu:88-109) */
void orc_vext_box(const orc_geom *g, double V, double B, const double w[3], double *v)
{
    const long long n = npts(g);
#pragma omp parallel for schedule(static)
    for (long long i = 0; i < n; ++i) {
        int ix, iy, iz;
        index3(g, i, &ix, &iy, &iz);
        const double x = centre(g, 0, ix), y = centre(g, 1, iy), z = centre(g, 2, iz);
        v[i] = 0.5 * V *
               ((tanh(B * (x - w[0])) - tanh(B * (x + w[0]))) * (tanh(B * (y - w[1])) - tanh(B * (y + w[1]))) *
                (tanh(B * (z - w[2])) - tanh(B * (z + w[2]))));
    }
}

/* V += A cos(n 2 pi / L x), L = 2 xmax_x (src/vext.cu:122-179) */
void orc_vext_add_cos(const orc_geom *g, double A, double nn, double *v)
{
    const double L = 2 * g->xmax[0];
    const long long n = npts(g);
#pragma omp parallel for schedule(static)
    for (long long i = 0; i < n; ++i) {
        int ix, iy, iz;
        index3(g, i, &ix, &iy, &iz);
        v[i] += A * cos(nn * 2 * ORC_PI / L * centre(g, 0, ix));
    }
}
