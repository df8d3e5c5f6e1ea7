// params_dump.cpp -- prints the code-unit parameters an input file yields
// (Params / Params2) as one JSON object, for tests and for checking an input
// file before a run.  `gpe_params <inputfile>`.
//
// `gpe_params <inputfile> --init <prefix> [box] [<lattice> amp k offx offy offz]`
// also builds the initial state and trap on the host with the Psi and Vext
// classes (the reference's CPU path: std::mt19937 noise) and writes them as
// raw complex doubles to <prefix>psi<c>.bin and <prefix>vext<c>.bin (c = 1,
// or 1 and 2).  `box` replaces the trap by Vext::initBox; a lattice name
// (1d, 2d_triangular, 2d_square, 3d_hex, 3d_fcc, 3d_bcc, 3d_sc) adds that
// profile to component 1.
#include <cstdio>
#include <cstdlib>
#include <cstring>
#include <string>

#include "hpp/params.hpp"
#include "hpp/psi.hpp"
#include "hpp/vext.hpp"

namespace {

void write_raw(const std::string& name, const complex* a, int n)
{
    FILE* f = std::fopen(name.c_str(), "wb");
    if (!f || std::fwrite(a, sizeof(complex), n, f) != (size_t)n) {
        std::fprintf(stderr, "Error: cannot write %s\n", name.c_str());
        std::exit(2);
    }
    std::fclose(f);
}

void force(Psi& psi, const ParamsBase& p, const char* name, const double* v)
{
    const std::string l(name);
    if (l == "1d") psi.force1DLattice(p, v[0], v[1], v[2]);
    else if (l == "2d_triangular") psi.force2DTriangularLattice(p, v[0], v[1], v[2], v[3]);
    else if (l == "2d_square") psi.force2DSquareLattice(p, v[0], v[1], v[2], v[3]);
    else if (l == "3d_hex") psi.force3DHEXLattice(p, v[0], v[1], v[2], v[3], v[4]);
    else if (l == "3d_fcc") psi.force3DFCCLattice(p, v[0], v[1], v[2], v[3], v[4]);
    else if (l == "3d_bcc") psi.force3DBCCLattice(p, v[0], v[1], v[2], v[3], v[4]);
    else if (l == "3d_sc") psi.force3DSCLattice(p, v[0], v[1], v[2], v[3], v[4]);
    else {
        std::fprintf(stderr, "Error: unknown lattice %s\n", name);
        std::exit(1);
    }
}

// argv[3] = prefix, then optional "box", then optional lattice + 5 numbers
int dump_init(const InputVariables& iv, int argc, const char* argv[])
{
    const std::string prefix = argv[3];
    int a = 4;
    const bool box = a < argc && std::strcmp(argv[a], "box") == 0;
    if (box) ++a;
    const char* lattice = nullptr;
    double lv[5] = {0, 0, 0, 0, 0};
    if (a < argc) {
        if (argc - a != 6) {
            std::fprintf(stderr, "Error: a lattice takes amp k offx offy offz\n");
            return 1;
        }
        lattice = argv[a];
        for (int j = 0; j < 5; ++j) lv[j] = std::atof(argv[a + 1 + j]);
    }
    if (!iv.getBool("t2component")) {
        Params p(iv);
        Psi psi(p, "");
        if (lattice) force(psi, p, lattice, lv);
        Vext vext(p, "");
        if (box) vext.initBox(p, 0);
        write_raw(prefix + "psi1.bin", psi.getPsi(), p.Npoints);
        write_raw(prefix + "vext1.bin", vext.getVext(), p.Npoints);
    } else {
        Params2 p(iv);
        for (int c = 1; c <= 2; ++c) {
            Psi psi(p, c);
            if (lattice && c == 1) force(psi, p, lattice, lv);
            Vext vext(p, c);
            if (box) vext.initBox(p, c);
            write_raw(prefix + "psi" + std::to_string(c) + ".bin", psi.getPsi(), p.Npoints);
            write_raw(prefix + "vext" + std::to_string(c) + ".bin", vext.getVext(), p.Npoints);
        }
    }
    return 0;
}

}  // namespace

int main(int argc, const char* argv[])
{
    if (argc != 2 && !(argc >= 4 && std::strcmp(argv[2], "--init") == 0)) {
        std::fprintf(stderr, "Usage: %s <filename> [--init <prefix> [box] [<lattice> amp k offx offy offz]]\n",
                     argv[0]);
        return 1;
    }
    InputVariables iv(argv[1]);
    if (argc >= 4) return dump_init(iv, argc, argv);
    if (!iv.getBool("t2component")) {
        Params p(iv);   // before the JSON: its set-up may print (the soft-core box lines)
        std::cout.flush();
        std::printf("{\"ncomp\": 1, \"aho\": %.17g, \"omho\": %.17g, \"a\": %.17g, \"add\": %.17g, \"edd\": %.17g, "
                    "\"npart\": %.17g, \"ax\": [%.17g, %.17g, %.17g], ",
                    p.aho, p.omho, p.a, p.add, p.edd, p.npart, p.ax[0], p.ax[1], p.ax[2]);
        std::printf("\"DIM\": %d, \"NX\": [%d, %d, %d], \"XMAX\": [%.17g, %.17g, %.17g], \"DX\": [%.17g, %.17g, "
                    "%.17g], \"DKX\": [%.17g, %.17g, %.17g], \"KXMAX\": [%.17g, %.17g, %.17g], \"DV\": %.17g, "
                    "\"dt\": %.17g, \"niter\": %d, \"itmod\": %d, \"epsilon_e\": %.17g",
                    p.DIM, p.NX[0], p.NX[1], p.NX[2], p.XMAX[0], p.XMAX[1], p.XMAX[2], p.DX[0], p.DX[1], p.DX[2],
                    p.DKX[0], p.DKX[1], p.DKX[2], p.KXMAX[0], p.KXMAX[1], p.KXMAX[2], p.DV, p.dt, p.niter, p.itmod,
                    p.epsilon_e);
    } else {
        Params2 p(iv);
        std::cout.flush();
        std::printf("{\"ncomp\": 2, \"aho\": %.17g, \"omho\": %.17g, \"aho2\": %.17g, \"a11\": %.17g, \"a22\": %.17g, "
                    "\"a12\": %.17g, \"npart1\": %.17g, \"npart2\": %.17g, \"ax1\": [%.17g, %.17g, %.17g], "
                    "\"ax2\": [%.17g, %.17g, %.17g], ",
                    p.aho1, p.omho1, p.aho2, p.a11, p.a22, p.a12, p.npart1, p.npart2, p.ax1[0], p.ax1[1], p.ax1[2],
                    p.ax2[0], p.ax2[1], p.ax2[2]);
        std::printf("\"DIM\": %d, \"NX\": [%d, %d, %d], \"XMAX\": [%.17g, %.17g, %.17g], \"DX\": [%.17g, %.17g, "
                    "%.17g], \"DKX\": [%.17g, %.17g, %.17g], \"KXMAX\": [%.17g, %.17g, %.17g], \"DV\": %.17g, "
                    "\"dt\": %.17g, \"niter\": %d, \"itmod\": %d, \"epsilon_e\": %.17g",
                    p.DIM, p.NX[0], p.NX[1], p.NX[2], p.XMAX[0], p.XMAX[1], p.XMAX[2], p.DX[0], p.DX[1], p.DX[2],
                    p.DKX[0], p.DKX[1], p.DKX[2], p.KXMAX[0], p.KXMAX[1], p.KXMAX[2], p.DV, p.dt, p.niter, p.itmod,
                    p.epsilon_e);
    }
    std::printf("}\n");
    return 0;
}
