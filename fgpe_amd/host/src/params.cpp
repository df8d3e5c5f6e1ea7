// params.cpp -- Params::InitializeParams / calculateLengthScales /
// RedefineLatticeSoft (src/params.cu:35-279) and the two-component Params2 (src/params2.cu:
// 28-186), including their unit conversion: lengths in a_ho of the geometric
// mean trap frequency, a in a_ho, dt in ms converted at the call sites.
#include "../hpp/params.hpp"

Params::Params(const InputVariables& iv_) { iv = iv_; InitializeParams(this); }
Params::Params(const Params& other) { iv = other.getIV(); InitializeParams(&other); }
Params& Params::operator=(const Params& other)
{
    if (this == &other) return *this;
    iv = other.getIV();
    InitializeParams(&other);
    return *this;
}

// src/params.cu:35-131.  The inter-droplet distance a_soft / (akrot / 2 pi)
// is stretched per axis for the expected lattice, and each box length is
// rounded to a whole number of those distances along the active axes (BCC's
// half-cell offset is overwritten at once in the reference, and so here).
void Params::RedefineLatticeSoft(const Params* par)
{
    std::cout << "Old numerical domain lengths: \nLX: " << 2 * XMAX[0] << "\nLY: " << 2 * XMAX[1]
              << "\nLZ: " << 2 * XMAX[2] << std::endl;
    const InputVariables& v = par->iv;
    const double akrot = v.getDouble("akrot") / (2 * pi);
    const double as = v.getDouble("a_soft");
    const double idd = as / akrot;
    const double triangular = 2 * std::sqrt(3.) / 2.;
    double d[3] = {idd, idd, idd};
    const std::string lt = v.getString("Latticetype");
    if (lt == "triangular" || lt == "HEX") d[1] = idd * triangular;
    if (lt == "BCC") d[0] = d[1] = d[2] = idd * std::sqrt(2);
    if (lt == "SC") d[0] = d[1] = d[2] = idd / std::sqrt(2);
    if (lt == "FCC") d[0] = d[1] = d[2] = idd * std::sqrt(3);
    const int active = DIM < 1 ? 0 : (DIM > 3 ? 3 : DIM);
    for (int i = 0; i < active; ++i) {
        const int n = (int)std::round((2 * XMAX[i]) / d[i]);
        XMAX[i] = 0.5 * n * d[i];
    }
    std::cout << "New numerical domain lengths: \nLX: " << 2 * XMAX[0] << "\nLY: " << 2 * XMAX[1]
              << "\nLZ: " << 2 * XMAX[2] << std::endl;
}

void Params::InitializeParams(const Params* par)
{
    const InputVariables& v = par->iv;
    outpath = v.getString("outprefix");
    inpath = v.getString("inprefix");
    vpath = v.getString("vprefix");
    NX[0] = v.getInt("nx");
    NX[1] = v.getInt("ny");
    NX[2] = v.getInt("nz");
    // dimension test of src/params.cu:145-154
    DIM = 3;
    if (((NX[0] - 1) * (NX[1] - 1) == 1) || ((NX[1] - 1) * (NX[2] - 1) == 0)) {
        if ((NX[0] - 1) * (NX[1] - 1) == (NX[1] - 1) * (NX[2] - 1)) DIM = 1;
        else DIM = 2;
    }
    Npoints = NX[0] * NX[1] * NX[2];
    XMAX[0] = v.getDouble("xmax");
    XMAX[1] = v.getDouble("ymax");
    XMAX[2] = v.getDouble("zmax");
    if (v.getString("hamiltonian") == "soft") RedefineLatticeSoft(par);
    for (int i = 0; i < 3; ++i) DX[i] = 2 * XMAX[i] / NX[i];
    omega[0] = 2 * pi * v.getDouble("omx");
    omega[1] = 2 * pi * v.getDouble("omy");
    omega[2] = 2 * pi * v.getDouble("omz");
    mass = v.getDouble("mass");
    if (v.getVariable("n1D") != "None") npart = v.getDouble("n1D") * (2 * XMAX[0]);
    else npart = v.getDouble("npart");
    a = v.getDouble("a");
    if (v.getVariable("add") != "None") {
        add = v.getDouble("add");
        edd = add / a;
    } else {
        add = 0.0;
        edd = 0.0;
    }
    calculateLengthScales();
    dt = v.getDouble("dt");
    time0 = v.getDouble("time0");
    epsilon_e = v.getDouble("epsilon_e");
    niter = v.getInt("niter");
    itmod = v.getInt("itmod");
    for (int i = 0; i < 3; ++i) {
        KXMAX[i] = 1.0 / (2.0 * DX[i]);
        DKX[i] = 1.0 / (NX[i] * DX[i]);
    }
    DV = DX[0] * DX[1] * DX[2];
    V = DV * Npoints;
}

void Params::calculateLengthScales()
{
    if ((std::abs(omega[0]) < 1E-6) && (std::abs(omega[1]) < 1E-6) && (std::abs(omega[2]) < 1E-6)) {
        aho = std::sqrt(hbar / (mass));
        omho = 1.0;
    } else if ((std::abs(omega[0]) < 1E-6) && (std::abs(omega[1]) < 1E-6)) {
        ax[2] = std::sqrt(hbar / (mass * omega[2]));
        aho = ax[2];
        ax[2] /= aho;
        omho = omega[2];
    } else if (std::abs(omega[0]) < 1E-6) {
        ax[1] = std::sqrt(hbar / (mass * omega[1]));
        ax[2] = std::sqrt(hbar / (mass * omega[2]));
        aho = std::sqrt(ax[1] * ax[2]);
        ax[1] /= aho;
        ax[2] /= aho;
        omho = std::sqrt(omega[2] * omega[1]);
    } else {
        for (int i = 0; i < 3; ++i) ax[i] = std::sqrt(hbar / (mass * omega[i]));
        aho = std::pow(ax[0] * ax[1] * ax[2], 1. / 3.);
        for (int i = 0; i < 3; ++i) ax[i] /= aho;
        omho = std::pow(omega[0] * omega[1] * omega[2], 1. / 3.);
    }
    a = a * abohr / aho;
    if (add != 0.0) {
        add = add * abohr / aho;
        if ((std::abs(add) < 1E-6) && (std::abs(a) < 1E-6)) edd = 0.0;
        else edd = add / a;
    }
    for (int i = 0; i < 3; ++i) {
        XMAX[i] = XMAX[i] / aho;
        DX[i] = DX[i] / aho;
    }
}

Params2::Params2(const InputVariables& iv_) { iv = iv_; InitializeParams(this); }
Params2::Params2(const Params2& other) { iv = other.getIV(); InitializeParams(&other); }
Params2& Params2::operator=(const Params2& other)
{
    if (this == &other) return *this;
    iv = other.getIV();
    InitializeParams(&other);
    return *this;
}

void Params2::InitializeParams(const Params2* par)
{
    const InputVariables& v = par->iv;
    isTwoComponent = true;
    outpath = v.getString("outprefix");
    inpath = v.getString("inprefix");
    vpath = v.getString("vprefix");
    NX[0] = v.getInt("nx");
    NX[1] = v.getInt("ny");
    NX[2] = v.getInt("nz");
    DIM = 3;
    if ((NX[0] * NX[1] == 0) || (NX[1] * NX[2] == 0)) {
        if (NX[0] * NX[1] == NX[1] * NX[2]) DIM = 1;
        else DIM = 2;
    }
    Npoints = NX[0] * NX[1] * NX[2];
    XMAX[0] = v.getDouble("xmax");
    XMAX[1] = v.getDouble("ymax");
    XMAX[2] = v.getDouble("zmax");
    for (int i = 0; i < 3; ++i) DX[i] = 2 * XMAX[i] / NX[i];
    omega1[0] = 2 * pi * v.getDouble("omx1");
    omega1[1] = 2 * pi * v.getDouble("omy1");
    omega1[2] = 2 * pi * v.getDouble("omz1");
    omega2[0] = 2 * pi * v.getDouble("omx2");
    omega2[1] = 2 * pi * v.getDouble("omy2");
    omega2[2] = 2 * pi * v.getDouble("omz2");
    mass1 = v.getDouble("mass1");
    npart1 = v.getDouble("npart1");
    a11 = v.getDouble("a11");
    mass2 = v.getDouble("mass2");
    npart2 = v.getDouble("npart2");
    a22 = v.getDouble("a22");
    a12 = v.getDouble("a12");
    calculateLengthScales();
    dt = v.getDouble("dt");
    time0 = v.getDouble("time0");
    epsilon_e = v.getDouble("epsilon_e");
    niter = v.getInt("niter");
    itmod = v.getInt("itmod");
    for (int i = 0; i < 3; ++i) {
        KXMAX[i] = 1.0 / (2.0 * DX[i]);
        DKX[i] = 1.0 / (NX[i] * DX[i]);
    }
    DV = DX[0] * DX[1] * DX[2];
    V = DV * Npoints;
}

static void lengths(const double* om, double mass, double* ax, double& aho, double& omho, bool second)
{
    if ((std::abs(om[0]) < 1E-6) && (std::abs(om[1]) < 1E-6)) {
        ax[2] = std::sqrt(hbar / (mass * om[2]));
        aho = ax[2];
        ax[2] /= aho;
        omho = om[2];
    } else if (std::abs(om[0]) < 1E-6) {
        ax[1] = std::sqrt(hbar / (mass * om[1]));
        ax[2] = std::sqrt(hbar / (mass * om[2]));
        aho = std::sqrt(ax[1] * ax[2]);
        ax[1] /= aho;
        ax[2] /= aho;
        omho = std::sqrt(om[2] * om[1]);
    } else {
        for (int i = 0; i < 3; ++i) ax[i] = std::sqrt(hbar / (mass * om[i]));
        // the second component's geometric mean omits ax[0] in the reference
        // (src/params2.cu:150); kept so that its ax2 and traps agree
        aho = second ? std::pow(ax[1] * ax[2], 1. / 3.) : std::pow(ax[0] * ax[1] * ax[2], 1. / 3.);
        for (int i = 0; i < 3; ++i) ax[i] /= aho;
        omho = std::pow(om[0] * om[1] * om[2], 1. / 3.);
    }
}

void Params2::calculateLengthScales()
{
    lengths(omega1, mass1, ax1, aho1, omho1, false);
    lengths(omega2, mass2, ax2, aho2, omho2, true);
    if (std::abs(omho2 - omho1) > 1E-6) {
        std::cerr << "ERROR: Different traps for each spin component currently not supported. Exiting..."
                  << std::endl;
        exit(1);
    }
    a11 = a11 * abohr / aho1;
    a22 = a22 * abohr / aho1;
    a12 = a12 * abohr / aho1;
    for (int i = 0; i < 3; ++i) {
        XMAX[i] = XMAX[i] / aho1;
        DX[i] = DX[i] / aho1;
    }
}
