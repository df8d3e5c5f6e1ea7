// psi.cpp -- the host-side initial wave function (Psi).  It follows the
// reference's initialisers, src/psi.cu:7-395, in what it computes: the
// envelopes, the lattice profiles and the normalisation keep the
// reference's expressions and their evaluation order, and the noise keeps its
// generator and draw order, so that a seeded run gives the reference's
// numbers.  Every initialiser is one walk over the grid (CellGrid).
#include "../hpp/psi.hpp"

#include <cmath>
#include <cstdio>

#include "cell_grid.hpp"

using fgpe_host::CellGrid;

// One complex noise value.  The reference writes it as
// complex(re_amp * u(g), im_amp * u(g)) (src/psi.cu:93); g++, the host
// compiler behind nvcc, evaluates those two arguments right to left, so the
// imaginary part takes the first number of the stream.
complex Psi::Noise::draw(double re_amp, double im_amp)
{
    const double ui = u_(gen_);
    const double ur = u_(gen_);
    return complex(re_amp * ur, im_amp * ui);
}

namespace {

// a tube along x when some trap frequency vanishes and omega_x does
// (src/psi.cu:19-20, :62-64)
bool tube_trap(const double* omega)
{
    bool some_zero = false;
    for (int a = 0; a < 3; ++a) some_zero = some_zero || std::abs(omega[a]) < 1E-6;
    return some_zero && std::abs(omega[0]) < 1E-6;
}

double npart_of(const ParamsBase& p, int type) { return p._npart(type == 1 ? 1 : 2); }

}  // namespace

Psi::Psi(const ParamsBase& p, int type, int seed, const std::string& read_prefix) : noise(seed)
{
    psi = (complex*)malloc(sizeof(complex) * (size_t)p.Npoints);
    if (p.getBool("read_psi")) {
        std::cout << "# Reading initial wavefunction from: " << p.inpath << std::endl;
        readPsi(read_prefix, p.Npoints);
        return;
    }
    const bool tube = tube_trap(p._omega(type == 1 ? 1 : 2));
    std::cout << (tube ? "# Initializing tubular psi_init..." : "# Initializing harmonic psi_init...") << std::endl;
    fill(p, type, tube ? Envelope::Tube : Envelope::Gaussian);
    normalize(p, type);
}

// One component: seed rndseed, normalised with the default type (-1).
Psi::Psi(const Params& p, const std::string&) : Psi(p, -1, p.getInt("rndseed"), p.inpath) {}

// Component type of two: seed rndseed + type, file <inprefix><type>psi_final.wdat.
Psi::Psi(const ParamsBase& p, const int& type)
    : Psi(p, type, p.getInt("rndseed") + type, p.inpath + std::to_string(type))
{
}

void Psi::fill(const ParamsBase& p, int type, Envelope env)
{
    const double rn = p.getDouble("rand_pct");
    const double cn = p.getDouble("crand_pct");
    const double* ax = p._ax(type == 1 ? 1 : 2);
    CellGrid(p).walk([&](long long i, double X, double Y, double Z) {
        const complex one_plus = complex(1, 0) + noise.draw(rn, cn);
        if (env == Envelope::Flat) {
            psi[i] = one_plus;
            return;
        }
        const double y = Y / ax[1], z = Z / ax[2];
        double r2 = (y * y) + (z * z);
        if (env == Envelope::Gaussian) {
            const double x = X / ax[0];
            r2 = r2 + (x * x);
        }
        psi[i] = one_plus * complex(std::exp(-0.5 * 0.1 * r2), 0);
    });
}

void Psi::initTube(const ParamsBase& p, const int& type) { fill(p, type, Envelope::Tube); }
void Psi::initHarmonic(const ParamsBase& p, const int& type) { fill(p, type, Envelope::Gaussian); }
void Psi::initVoid(const ParamsBase& p, const int& type) { fill(p, type, Envelope::Flat); }

double Psi::getX(const ParamsBase& p, int i) { return fgpe_host::cell_centre(p, 0, iX(i, p.NX[0], p.NX[1])); }
double Psi::getY(const ParamsBase& p, int i) { return fgpe_host::cell_centre(p, 1, iY(i, p.NX[0], p.NX[1])); }
double Psi::getZ(const ParamsBase& p, int i) { return fgpe_host::cell_centre(p, 2, iZ(i, p.NX[0], p.NX[1])); }

// psi *= exp(i Ncirc 2 pi x / (2 XMAX_x))
void Psi::imprintVortexTube(const ParamsBase& p, const double& Ncirc)
{
    const double len = 2 * p.XMAX[0];
    CellGrid(p).walk([&](long long i, double x, double, double) {
        const double phase = Ncirc * 2 * pi / len * x;
        psi[i] = complex(std::cos(phase), std::sin(phase)) * psi[i];
    });
}

// psi *= 1 + noise on the first `length` points
void Psi::addNoise(complex& noise_amp, const int& length)
{
    for (int i = 0; i < length; ++i) psi[i] = (complex(1, 0) + noise.draw(noise_amp.real(), noise_amp.imag())) * psi[i];
}

void Psi::addProfile(const ParamsBase& p, int type, const double off[3],
                     const std::function<double(double, double, double)>& f)
{
    const double avg = std::pow(npart_of(p, type) / fgpe_host::cell_volume(p), 0.5);
    CellGrid(p).walk([&](long long i, double x, double y, double z) {
        psi[i] = f(x - off[0], y - off[1], z - off[2]) * avg + psi[i];
    });
    // the reference normalises with component 1's particle number either way
    normalize(p, 1);
}

// ---- the lattice profiles (src/psi.cu:179-347): amp times the mean of
// cosines along the lattice's directions
void Psi::force1DLattice(const ParamsBase& p, const double& amp, const double& k, const double& off,
                         const int& type)
{
    const double o[3] = {off, 0, 0};
    addProfile(p, type, o, [&](double x, double, double) { return amp * (1 - std::cos(k * x)) / 2.; });
}

void Psi::force2DTriangularLattice(const ParamsBase& p, const double& amp, const double& k, const double& offx,
                                   const double& offy, const int& type)
{
    const double o[3] = {offx, offy, 0};
    addProfile(p, type, o, [&](double x, double y, double) {
        const double s = std::cos(k * x) + std::cos(k * (-0.5 * x + (std::sqrt(3) / 2) * y)) +
                         std::cos(k * (-0.5 * x - (std::sqrt(3) / 2) * y));
        return amp * s / 3;
    });
}

void Psi::force2DSquareLattice(const ParamsBase& p, const double& amp, const double& k, const double& offx,
                               const double& offy, const int& type)
{
    const double o[3] = {offx, offy, 0};
    addProfile(p, type, o, [&](double x, double y, double) { return amp * (std::cos(k * y) + std::cos(k * x)) / 2; });
}

void Psi::force3DHEXLattice(const ParamsBase& p, const double& amp, const double& k, const double& offx,
                            const double& offy, const double& offz, const int& type)
{
    const double o[3] = {offx, offy, offz};
    addProfile(p, type, o, [&](double x, double y, double z) {
        // directions -60 deg, 0 and +60 deg in the xy plane, and z
        const double d1 = (-x + std::sqrt(3.) * y) / 2., d4 = (-x - std::sqrt(3.) * y) / 2.;
        return amp * (std::cos(k * d1) + std::cos(k * x) + std::cos(k * z) + std::cos(k * d4)) / 4.0;
    });
}

void Psi::force3DFCCLattice(const ParamsBase& p, const double& amp, const double& k, const double& offx,
                            const double& offy, const double& offz, const int& type)
{
    const double o[3] = {offx, offy, offz};
    addProfile(p, type, o, [&](double x, double y, double z) {
        const double r = std::sqrt(3.);
        // the reference's nine projections, the (1,1,1) one written three ways
        const double d[9] = {(x + y - z) / r, (y + z - x) / r, (z + x - y) / r, (x - y - z) / r, (y - z - x) / r,
                             (z - x - y) / r, (x + y + z) / r, (y + z + x) / r, (z + x + y) / r};
        double s = 0.0;
        for (int j = 0; j < 9; ++j) s = j ? s + std::cos(k * d[j]) : std::cos(k * d[j]);
        return amp * s / 9.0;
    });
}

void Psi::force3DBCCLattice(const ParamsBase& p, const double& amp, const double& k, const double& offx,
                            const double& offy, const double& offz, const int& type)
{
    const double o[3] = {offx, offy, offz};
    addProfile(p, type, o, [&](double x, double y, double z) {
        const double r = std::sqrt(2);
        const double d[6] = {(x + y) / r, (z + y) / r, (z + x) / r, (x - y) / r, (z - y) / r, (z - x) / r};
        double s = 0.0;
        for (int j = 0; j < 6; ++j) s = j ? s + std::cos(k * d[j]) : std::cos(k * d[j]);
        return amp * s / 6;
    });
}

void Psi::force3DSCLattice(const ParamsBase& p, const double& amp, const double& k, const double& offx,
                           const double& offy, const double& offz, const int& type)
{
    const double o[3] = {offx, offy, offz};
    addProfile(p, type, o, [&](double x, double y, double z) {
        return amp * (std::cos(k * z) + std::cos(k * y) + std::cos(k * x)) / 3;
    });
}

// psi *= sqrt(npart / (sum |psi|^2 dV)), npart of component 1 for type 1 and
// of component 2 otherwise (one component: its only npart)
void Psi::normalize(const ParamsBase& p, const int& type)
{
    const long long n = p.Npoints;
    double s = 0;
    for (long long i = 0; i < n; ++i) {
        const double m = std::abs(psi[i]);
        s += m * m;
    }
    const double scale = std::pow(npart_of(p, type) / (s * p.DX[0] * p.DX[1] * p.DX[2]), 0.5);
    for (long long i = 0; i < n; ++i) psi[i] = psi[i] * scale;
    double check = 0;
    for (long long i = 0; i < n; ++i) {
        const double m = std::abs(psi[i]);
        check += m * m * p.DX[0] * p.DX[1] * p.DX[2];
    }
    std::cout << "# Init Psi norm: " << check << std::endl;
}

// <prefix>psi_final.wdat, raw complex doubles.  A missing file is an error
// (exit 2); the reference would read through a null FILE* (src/psi.cu:383-394).
void Psi::readPsi(const std::string& prefix, const int& length)
{
    const std::string name = prefix + "psi_final.wdat";
    FILE* f = std::fopen(name.c_str(), "rb");
    if (!f) {
        std::fprintf(stderr, "Error: Could not open file %s!\n", name.c_str());
        std::exit(2);
    }
    const size_t got = std::fread(psi, sizeof(complex), (size_t)length, f);
    if (got != (size_t)length)
        std::fprintf(stderr, "Error: Only %zu elements were read instead of %d!\n", got, length);
    std::fclose(f);
}
