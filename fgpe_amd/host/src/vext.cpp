// vext.cpp -- the host-side trap (Vext).  It follows the reference's
// initialisers, src/vext.cu:6-218, in what it computes: each potential keeps
// the reference's expression and evaluation order (a trap's bits decide the
// run's).  Every initialiser is one walk over the grid (CellGrid).
#include "../hpp/vext.hpp"

#include <cmath>
#include <cstdio>

#include "cell_grid.hpp"

using fgpe_host::CellGrid;

namespace {

// a tube along x when some trap frequency vanishes and omega_x does
// (src/vext.cu:17-18)
bool tube_trap(const double* omega)
{
    bool some_zero = false;
    for (int a = 0; a < 3; ++a) some_zero = some_zero || std::abs(omega[a]) < 1E-6;
    return some_zero && std::abs(omega[0]) < 1E-6;
}

}  // namespace

Vext::Vext(const ParamsBase& p, int type, const std::string& read_prefix, bool say)
{
    vext = (complex*)malloc(sizeof(complex) * (size_t)p.Npoints);
    if (p.getBool("read_vext")) {
        readVext(read_prefix, p.Npoints);
        return;
    }
    const bool tube = tube_trap(p._omega(type == 1 ? 1 : 2));
    if (say) std::cout << (tube ? "# Initializing tubular Vext..." : "# Initializing harmonic Vext...") << std::endl;
    harmonic(p, type, !tube);
}

Vext::Vext(const Params& p, const std::string&) : Vext(p, -1, p.vpath, true) {}

Vext::Vext(const Params2& p, const int& type) : Vext(p, type, p.vpath + std::to_string(type), false) {}

// 1/2 sum_a (x_a / ax_a)^2 / ax_a^2 over y, z and, unless a tube, x
// (src/vext.cu:60-85: the coordinate is divided by ax twice)
void Vext::harmonic(const ParamsBase& p, int type, bool along_x)
{
    const double* ax = p._ax(type == 1 ? 1 : 2);
    const double a0 = std::pow(ax[0], 2), a1 = std::pow(ax[1], 2), a2 = std::pow(ax[2], 2);
    CellGrid(p).walk([&](long long i, double X, double Y, double Z) {
        const double y = Y / ax[1], z = Z / ax[2];
        double v;
        if (along_x) {
            const double x = X / ax[0];
            v = 0.5 * ((x * x) / a0 + (y * y) / a1 + (z * z) / a2);
        } else {
            v = 0.5 * ((y * y) / a1 + (z * z) / a2);
        }
        vext[i] = v;
    });
}

void Vext::initTube(const ParamsBase& p, const int& type) { harmonic(p, type, false); }
void Vext::initHarmonic(const ParamsBase& p, const int& type) { harmonic(p, type, true); }

// 1/2 V prod_a (tanh(B (x_a - w_a)) - tanh(B (x_a + w_a))), w_a = {x,y,z}box
// over the oscillator length _aho(0), which for two components is component
// 2's (src/vext.cu:88-109)
void Vext::initBox(const ParamsBase& p, const int&)
{
    const double V = p.getDouble("Vbox"), B = p.getDouble("Bbox");
    const double w[3] = {p.getDouble("xbox") / p._aho(0), p.getDouble("ybox") / p._aho(0),
                         p.getDouble("zbox") / p._aho(0)};
    auto wall = [&](double x, int a) { return std::tanh(B * (x - w[a])) - std::tanh(B * (x + w[a])); };
    CellGrid(p).walk([&](long long i, double x, double y, double z) {
        vext[i] = 0.5 * V * (wall(x, 0) * wall(y, 1) * wall(z, 2));
    });
}

void Vext::initVoid(const ParamsBase& p, const int&)
{
    for (long long i = 0; i < p.Npoints; ++i) vext[i] = 0;
}

void Vext::addCosine(const ParamsBase& p, double A, double n)
{
    const double L = 2 * p.XMAX[0];
    CellGrid(p).walk([&](long long i, double x, double, double) { vext[i] += A * std::cos(n * 2 * pi / L * x); });
}

void Vext::addProtocolPotential(const ParamsBase& p, const int&, const int sign)
{
    addCosine(p, p.getDouble("aseed") * sign, p.getDouble("nseed"));
}

void Vext::addOpticalLattice(const ParamsBase& p, const int&)
{
    addCosine(p, p.getDouble("aopt"), p.getDouble("nopt"));
}

void Vext::addWeightedProtocolPotential(const Params2& p, const int& type)
{
    addCosine(p, weightedProtocolAmplitude(p, type), p.getDouble("nseed"));
}

// The seed amplitude of each component for a given counterflow c0 (input in
// mm/s): aseed when c0 = 0, else split by xi = -(4 pi (a11 - a12) - c0^2 /
// (n1 omho^2)) / (4 pi (a22 - a12) - c0^2 / (n2 omho^2)) (src/vext.cu:122-153).
double Vext::weightedProtocolAmplitude(const Params2& p, const int& type)
{
    const double c0 = p.getDouble("c0") / (p.aho1 * 1E-3);
    const double aseed = p.getDouble("aseed");
    double A = aseed;
    if (std::abs(c0) >= 1E-6) {
        const double n1 = p.npart1 / p.V, n2 = p.npart2 / p.V;
        const double num = 4 * pi * (p.a11 - p.a12) - c0 * c0 / (n1 * p.omho1 * p.omho1);
        const double den = 4 * pi * (p.a22 - p.a12) - c0 * c0 / (n2 * p.omho1 * p.omho1);
        const double xi = -1 * num / den;
        const double d = xi > 0 ? 1 + xi : 1 - xi;
        A = (type == 1) ? 2 * aseed * xi / d : 2 * aseed / d;
    }
    std::cout << "# The lambda imbalance: " << A << std::endl;
    return A;
}

double Vext::getX(const ParamsBase& p, int i) { return fgpe_host::cell_centre(p, 0, iX(i, p.NX[0], p.NX[1])); }
double Vext::getY(const ParamsBase& p, int i) { return fgpe_host::cell_centre(p, 1, iY(i, p.NX[0], p.NX[1])); }
double Vext::getZ(const ParamsBase& p, int i) { return fgpe_host::cell_centre(p, 2, iZ(i, p.NX[0], p.NX[1])); }

// <prefix>vext.wdat, raw complex doubles.  A missing file is an error
// (exit 2); the reference would read through a null FILE* (src/vext.cu:201-218).
void Vext::readVext(const std::string& prefix, const int& length)
{
    const std::string name = prefix + "vext.wdat";
    FILE* f = std::fopen(name.c_str(), "rb");
    if (!f) {
        std::fprintf(stderr, "Error: Could not open the Vext file %s!\n", name.c_str());
        std::exit(2);
    }
    const size_t got = std::fread(vext, sizeof(complex), (size_t)length, f);
    if (got != (size_t)length)
        std::fprintf(stderr, "Error: Only %zu elements were read instead of %d!\n", got, length);
    std::fclose(f);
}
