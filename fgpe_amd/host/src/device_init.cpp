// device_init.cpp -- Psi / Vext construction on the device for the plugins
// (the host path is src/psi.cpp and src/vext.cpp).
#include "../hpp/device_init.hpp"

namespace {

bool key_true(const ParamsBase& p, const char* key)
{
    return p.getIV().getVariable(key) != "None" && p.getBool(key);
}

bool flat(const double* omega)
{
    return std::any_of(omega, omega + 3, [](double x) { return std::abs(x) < 1E-6; }) && std::abs(omega[0]) < 1E-6;
}

}  // namespace

bool deviceInitPsi(const ParamsBase& p) { return key_true(p, "device_init") && !p.getBool("read_psi"); }
bool deviceInitVext(const ParamsBase& p) { return key_true(p, "device_init") && !p.getBool("read_vext"); }

fgpe_geom geomOf(const ParamsBase& p)
{
    fgpe_geom g{};
    for (int a = 0; a < 3; ++a) {
        g.xmax[a] = p.XMAX[a];
        g.dx[a] = p.DX[a];
    }
    return g;
}

void devicePsi(fgpe_ctx* ctx, const ParamsBase& p, int comp, int type, bool vortex, complex* host_psi)
{
    const fgpe_geom g = geomOf(p);
    const int kind = flat(p._omega(type)) ? FGPE_INIT_TUBE : FGPE_INIT_HARMONIC;
    std::cout << (kind == FGPE_INIT_TUBE ? "# Initializing tubular psi_init (device)..."
                                         : "# Initializing harmonic psi_init (device)...")
              << std::endl;
    const unsigned long long seed = (unsigned long long)(p.getInt("rndseed") + type);
    FCE(fgpe_psi_init(ctx, comp, kind, &g, p._ax(type), p.getDouble("rand_pct"), p.getDouble("crand_pct"), seed),
        "fgpe_psi_init");
    // Psi::normalize: component 2 uses npart2, everything else npart1
    const double npart = (p.isTwoComponent && type == 2) ? p._npart(2) : p._npart(1);
    double norm = 0;
    FCE(fgpe_psi_normalize(ctx, comp, npart, &g, &norm), "fgpe_psi_normalize");
    std::cout << "# Init Psi norm: " << norm << std::endl;
    if (vortex) FCE(fgpe_psi_imprint_vortex_tube(ctx, comp, &g, p.getDouble("api_prc")), "fgpe_psi_imprint_vortex_tube");
    FCE(fgpe_download_psi(ctx, comp, (double*)host_psi), "fgpe_download_psi");
}

void deviceVext(fgpe_ctx* ctx, const ParamsBase& p, int comp, int type, bool seed, double A, complex* host_vext)
{
    const fgpe_geom g = geomOf(p);
    const int kind = flat(p._omega(type)) ? FGPE_INIT_TUBE : FGPE_INIT_HARMONIC;
    FCE(fgpe_vext_init(ctx, comp, kind, &g, p._ax(type)), "fgpe_vext_init");
    if (seed) FCE(fgpe_vext_add_cos(ctx, comp, &g, A, p.getDouble("nseed")), "fgpe_vext_add_cos");
    FCE(fgpe_vext_commit(ctx, comp, (double*)host_vext), "fgpe_vext_commit");
}
