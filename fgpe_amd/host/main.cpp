// main.cpp -- the `gpe <inputfile>` front end (main.cu:20-82): single- or
// two-component mode from t2component, the plugin from `hamiltonian`.  Exit
// codes as in the reference: 1 wrong usage or unknown single-component
// Hamiltonian, 2 unknown two-component Hamiltonian (or unreadable input).
// Extra in this build: "dbec" (constant a, dipolar / soft-core terms),
// "template" (hamiltonians/single_template), and `precision mixed` in the
// input file selects the mixed-precision bulk.
#include "hamiltonians/BEC_oneComponent.hpp"
#include "hamiltonians/BEC_twoComponent.hpp"
#include "hamiltonians/single_template.hpp"

int main(int argc, const char* argv[])
{
    if (argc != 2) {
        std::cerr << "Usage: " << argv[0] << " <filename>" << std::endl;
        return 1;
    }
    std::string inputFile = argv[1];
    InputVariables reader(inputFile);
    int precision = FGPE_PREC_F64;
    if (reader.getVariable("precision") != "None" && reader.getString("precision") == "mixed")
        precision = FGPE_PREC_MIXED;

    if (!reader.getBool("t2component")) {
        std::cout << "# Single component mode." << std::endl;
        Params params(reader);
        std::string hamiltonian = params.getString("hamiltonian");
        if (hamiltonian == "bec") {
            std::cout << "# BEC hamiltonian" << std::endl;
            BEConeComponent bec(params, precision);
        } else if (hamiltonian == "dbec") {
            std::cout << "# dipolar BEC hamiltonian" << std::endl;
            BECDipolar bec(params, precision);
        } else if (hamiltonian == "template") {
            std::cout << "# template hamiltonian" << std::endl;
            SingleTemplate st(params, precision);
        } else {
            std::cerr << "No hamiltonian available for: " << hamiltonian << std::endl;
            exit(1);
        }
    } else {
        std::cout << "# Two component mode." << std::endl;
        Params2 params2(reader);
        std::string hamiltonian = params2.getString("hamiltonian");
        if (hamiltonian == "2bec") {
            std::cout << "# 2BEC hamiltonian" << std::endl;
            BECtwoComponent bec(params2, precision);
        } else {
            std::cerr << "No hamiltonian available for: " << hamiltonian << std::endl;
            exit(2);
        }
    }
    return 0;
}
