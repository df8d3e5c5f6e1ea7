// vext.cpp -- Vext initialisers (src/vext.cu:6-218).
#include "../hpp/vext.hpp"

#include <cstdio>

Vext::Vext(const Params& p, const std::string&)
{
    vext = (complex*)malloc(p.Npoints * sizeof(complex));
    if (p.getBool("read_vext")) {
        readVext(p.vpath, p.Npoints);
    } else {
        bool flat = std::any_of(p.omega, p.omega + 3, [](double x) { return std::abs(x) < 1E-6; });
        if (flat && std::abs(p.omega[0]) < 1E-6) {
            std::cout << "# Initializing tubular Vext..." << std::endl;
            initTube(p, 0);
        } else {
            std::cout << "# Initializing harmonic Vext..." << std::endl;
            initHarmonic(p, 0);
        }
    }
}

Vext::Vext(const Params2& p, const int& type)
{
    vext = (complex*)malloc(p.Npoints * sizeof(complex));
    if (p.getBool("read_vext")) {
        readVext(p.vpath + std::to_string(type), p.Npoints);
    } else {
        const double* omega = p._omega(type);
        bool flat = std::any_of(omega, omega + 3, [](double x) { return std::abs(x) < 1E-6; });
        if (flat && std::abs(omega[0]) < 1E-6) initTube(p, type);
        else initHarmonic(p, type);
    }
}

void Vext::initTube(const ParamsBase& p, const int& type)
{
    const double* ax = p._ax(type);
    for (int i = 0; i < p.Npoints; ++i) {
        double y = getY(p, i) / ax[1];
        double z = getZ(p, i) / ax[2];
        vext[i] = 0.5 * (std::pow(y, 2) / std::pow(ax[1], 2) + std::pow(z, 2) / std::pow(ax[2], 2));
    }
}

void Vext::initHarmonic(const ParamsBase& p, const int& type)
{
    const double* ax = p._ax(type);
    for (int i = 0; i < p.Npoints; ++i) {
        double x = getX(p, i) / ax[0];
        double y = getY(p, i) / ax[1];
        double z = getZ(p, i) / ax[2];
        vext[i] = 0.5 * (std::pow(x, 2) / std::pow(ax[0], 2) + std::pow(y, 2) / std::pow(ax[1], 2) +
                         std::pow(z, 2) / std::pow(ax[2], 2));
    }
}

// tanh box of half-widths xbox/aho, ybox/aho, zbox/aho (the reference reads
// _aho(0): for two components that is component 2's length)
void Vext::initBox(const ParamsBase& p, const int&)
{
    const double V = p.getDouble("Vbox");
    const double xx = p.getDouble("xbox") / p._aho(0);
    const double yy = p.getDouble("ybox") / p._aho(0);
    const double zz = p.getDouble("zbox") / p._aho(0);
    const double B = p.getDouble("Bbox");
    for (int i = 0; i < p.Npoints; ++i) {
        const double x = getX(p, i), y = getY(p, i), z = getZ(p, i);
        vext[i] = 0.5 * V *
                  ((std::tanh(B * (x - xx)) - std::tanh(B * (x + xx))) *
                   (std::tanh(B * (y - yy)) - std::tanh(B * (y + yy))) *
                   (std::tanh(B * (z - zz)) - std::tanh(B * (z + zz))));
    }
}

void Vext::initVoid(const ParamsBase& p, const int&)
{
    for (int i = 0; i < p.Npoints; ++i) vext[i] = 0;
}

void Vext::addWeightedProtocolPotential(const Params2& p, const int& type)
{
    const double A = weightedProtocolAmplitude(p, type);
    double L = 2 * p.XMAX[0];
    double n = p.getDouble("nseed");
    for (int i = 0; i < p.Npoints; ++i) vext[i] += A * std::cos(n * 2 * pi / L * getX(p, i));
}

double Vext::weightedProtocolAmplitude(const Params2& p, const int& type)
{
    const double n1 = p.npart1 / p.V;
    const double n2 = p.npart2 / p.V;
    const double c0 = p.getDouble("c0") / (p.aho1 * 1E-3);
    double A = 0;
    if (std::abs(c0) < 1E-6) {
        A = p.getDouble("aseed");
    } else {
        const double num = 4 * pi * (p.a11 - p.a12) - c0 * c0 / (n1 * p.omho1 * p.omho1);
        const double den = 4 * pi * (p.a22 - p.a12) - c0 * c0 / (n2 * p.omho1 * p.omho1);
        const double xi = -1 * num / den;
        if (xi > 0) A = (type == 1) ? 2 * p.getDouble("aseed") * xi / (1 + xi) : 2 * p.getDouble("aseed") / (1 + xi);
        else A = (type == 1) ? 2 * p.getDouble("aseed") * xi / (1 - xi) : 2 * p.getDouble("aseed") / (1 - xi);
    }
    std::cout << "# The lambda imbalance: " << A << std::endl;
    return A;
}

void Vext::addProtocolPotential(const ParamsBase& p, const int&, const int sign)
{
    double L = 2 * p.XMAX[0];
    double n = p.getDouble("nseed");
    double A = p.getDouble("aseed") * sign;
    for (int i = 0; i < p.Npoints; ++i) vext[i] += A * std::cos(n * 2 * pi / L * getX(p, i));
}

void Vext::addOpticalLattice(const ParamsBase& p, const int&)
{
    double L = 2 * p.XMAX[0];
    double n = p.getDouble("nopt");
    double A = p.getDouble("aopt");
    for (int i = 0; i < p.Npoints; ++i) vext[i] += A * std::cos(n * 2 * pi / L * getX(p, i));
}

double Vext::getX(const ParamsBase& p, int i) { return -p.XMAX[0] + iX(i, p.NX[0], p.NX[1]) * p.DX[0] + 0.5 * p.DX[0]; }
double Vext::getY(const ParamsBase& p, int i) { return -p.XMAX[1] + iY(i, p.NX[0], p.NX[1]) * p.DX[1] + 0.5 * p.DX[1]; }
double Vext::getZ(const ParamsBase& p, int i) { return -p.XMAX[2] + iZ(i, p.NX[0], p.NX[1]) * p.DX[2] + 0.5 * p.DX[2]; }

void Vext::readVext(const std::string& prefix, const int& length)
{
    std::string name = prefix + "vext.wdat";
    FILE* file = fopen(name.c_str(), "rb");
    if (file == NULL) {
        fprintf(stderr, "Error: Could not open the Vext file %s!\n", name.c_str());
        exit(2);
    }
    size_t n = fread(vext, sizeof(complex), length, file);
    if (n != (size_t)length)
        fprintf(stderr, "Error: Only %zu elements were read instead of %d!\n", n, length);
    fclose(file);
}
