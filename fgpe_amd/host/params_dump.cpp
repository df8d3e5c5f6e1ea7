// params_dump.cpp -- prints the code-unit parameters an input file yields
// (Params / Params2) as one JSON object, for tests and for checking an input
// file before a run.  `gpe_params <inputfile>`.
#include <cstdio>

#include "hpp/params.hpp"

int main(int argc, const char* argv[])
{
    if (argc != 2) {
        std::fprintf(stderr, "Usage: %s <filename>\n", argv[0]);
        return 1;
    }
    InputVariables iv(argv[1]);
    std::printf("{");
    if (!iv.getBool("t2component")) {
        Params p(iv);
        std::printf("\"ncomp\": 1, \"aho\": %.17g, \"omho\": %.17g, \"a\": %.17g, \"add\": %.17g, \"edd\": %.17g, "
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
        std::printf("\"ncomp\": 2, \"aho\": %.17g, \"omho\": %.17g, \"aho2\": %.17g, \"a11\": %.17g, \"a22\": %.17g, "
                    "\"a12\": %.17g, \"npart1\": %.17g, \"npart2\": %.17g, ",
                    p.aho1, p.omho1, p.aho2, p.a11, p.a22, p.a12, p.npart1, p.npart2);
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
