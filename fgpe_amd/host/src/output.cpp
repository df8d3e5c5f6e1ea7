// output.cpp -- byte formats of src/output.cu: .wdat = raw complex<double>
// (or double) arrays in ID(x,y,z) order, psi.wdat appended once per cycle
// (:193-197, 262-277); info.wtxt field layout (:26-173); status.out reopened
// every 10 rows so it can be followed during a run (:220-256).
#include "../hpp/output.hpp"

Output::Output(const std::string& prefix_) : prefix(prefix_) {}

static void wtxt_common(FILE* f, int ncycle, int dim, const int* NX, const double* DX, const double* L,
                        const std::string& outpath, double time0, double dt)
{
    fprintf(f, "cycles                   %d\n\n", ncycle);
    fprintf(f, "dim                      %d\n", dim);
    fprintf(f, "nx                       %d\n", NX[0]);
    fprintf(f, "ny                       %d\n", NX[1]);
    fprintf(f, "nz                       %d\n", NX[2]);
    fprintf(f, "dx                       %lf\n", DX[0]);
    fprintf(f, "dy                       %lf\n", DX[1]);
    fprintf(f, "dz                       %lf\n", DX[2]);
    fprintf(f, "lx                       %lf\n", L[0]);
    fprintf(f, "ly                       %lf\n", L[1]);
    fprintf(f, "lz                       %lf\n", L[2]);
    fprintf(f, "datadim                  %d\n", 3);
    fprintf(f, "prefix                   %s\n", outpath.c_str());
    fprintf(f, "t0                       %lf\n", time0);
    fprintf(f, "dt                       %lf\n\n", dt);
}

void Output::WriteWtxt(const Params& p, const int& ncycle)
{
    FILE* f = fopen((prefix + "info.wtxt").c_str(), "w");
    if (f == NULL) { perror("Wtxt file opening failed"); return; }
    // lx/ly/lz use NX[0] for all three axes in the reference (src/output.cu:42-44)
    const double L[3] = {p.NX[0] * p.DX[0] * p.aho, p.NX[0] * p.DX[1] * p.aho, p.NX[0] * p.DX[2] * p.aho};
    wtxt_common(f, ncycle, p.DIM, p.NX, p.DX, L, p.outpath, p.time0, p.dt);
    fprintf(f, "%-15s %-15s %-15s %-15s %-15s\n", "# tag", "name", "value", "unit", "format");
    fprintf(f, "%-15s %-15s %-15s %-15s %-15s\n", "var", "psi", "complex", "none", "wdat");
    fprintf(f, "%-15s %-15s %-15s %-15s %-15s\n", "var", "vext", "complex", "none", "wdat");
    fprintf(f, "%-15s %-15s %-15s %-15s %-15s\n\n", "", "", "", "", "");
    fprintf(f, "%-15s %-15s %-15s %-15s\n", "# tag", "name", "value", "unit");
    fprintf(f, "%-15s %-15s %-15.1f %-15s\n", "const", "hbar       ", hbar, "u*(um)^2/s");
    fprintf(f, "%-15s %-15s %-15.1f %-15s\n", "const", "kB         ", kB, "[hbar]/nK");
    fprintf(f, "%-15s %-15s %-15.2f %-15s\n", "const", "npart      ", p.npart, "none");
    fprintf(f, "%-15s %-15s %-15.6f %-15s\n", "const", "mass       ", p.mass, "u");
    fprintf(f, "%-15s %-15s %-15.6E %-15s\n", "const", "abohr      ", abohr, "mu_m");
    fprintf(f, "%-15s %-15s %-15.6f %-15s\n", "const", "a          ", p.a / abohr * p.aho, "a0");
    if (p.getIV().getVariable("add") != "None") {
        fprintf(f, "%-15s %-15s %-15.6f %-15s\n", "const", "add        ", p.add / abohr * p.aho, "a0");
        fprintf(f, "%-15s %-15s %-15.6f %-15s\n", "const", "dipole_x   ", p.getDouble("dipole_x"), "1");
        fprintf(f, "%-15s %-15s %-15.6f %-15s\n", "const", "dipole_y   ", p.getDouble("dipole_y"), "1");
        fprintf(f, "%-15s %-15s %-15.6f %-15s\n", "const", "dipole_z   ", p.getDouble("dipole_z"), "1");
    }
    fprintf(f, "%-15s %-15s %-15.6f %-15s\n", "const", "omx        ", p.omega[0] / (2 * pi), "2pi");
    fprintf(f, "%-15s %-15s %-15.6f %-15s\n", "const", "omy        ", p.omega[1] / (2 * pi), "2pi");
    fprintf(f, "%-15s %-15s %-15.6f %-15s\n", "const", "omz        ", p.omega[2] / (2 * pi), "2pi");
    fprintf(f, "%-15s %-15s %-15.6f %-15s\n", "const", "aho        ", p.aho, "?");
    fprintf(f, "%-15s %-15s %-15.6f %-15s\n", "const", "ax         ", p.ax[0], "code");
    fprintf(f, "%-15s %-15s %-15.6f %-15s\n", "const", "ay         ", p.ax[1], "code");
    fprintf(f, "%-15s %-15s %-15.6f %-15s\n", "const", "az         ", p.ax[2], "code");
    fprintf(f, "%-15s %-15s %-15.6f %-15s\n", "const", "omho       ", p.omho, "?");
    fprintf(f, "%-15s %-15s %-15.1f %-15s\n", "const", "Eho        ", p.omho * hbar, "Energy_scale");
    fprintf(f, "%-15s %-15s %-15d %-15s\n", "const", "itmod      ", p.itmod, ".");
    fclose(f);
}

void Output::WriteWtxt_2c(const Params2& p, const int& ncycle)
{
    FILE* f = fopen((prefix + "info.wtxt").c_str(), "w");
    if (f == NULL) { perror("Wtxt file opening failed"); return; }
    const double L[3] = {p.NX[0] * p.DX[0] * p.aho1, p.NX[1] * p.DX[1] * p.aho1, p.NX[2] * p.DX[2] * p.aho1};
    wtxt_common(f, ncycle, p.DIM, p.NX, p.DX, L, p.outpath, p.time0, p.dt);
    fprintf(f, "%-15s %-15s %-15s %-15s %-15s\n", "# tag", "name", "value", "unit", "format");
    fprintf(f, "%-15s %-15s %-15s %-15s %-15s\n", "var", "1psi", "complex", "none", "wdat");
    fprintf(f, "%-15s %-15s %-15s %-15s %-15s\n", "var", "2psi", "complex", "none", "wdat");
    fprintf(f, "%-15s %-15s %-15s %-15s %-15s\n", "var", "1vext", "complex", "none", "wdat");
    fprintf(f, "%-15s %-15s %-15s %-15s %-15s\n", "var", "2vext", "complex", "none", "wdat");
    fprintf(f, "%-15s %-15s %-15s %-15s %-15s\n\n", "", "", "", "", "");
    fprintf(f, "%-15s %-15s %-15s %-15s\n", "# tag", "name", "value", "unit");
    fprintf(f, "%-15s %-15s %-15.1f %-15s\n", "const", "hbar       ", hbar, "u*(um)^2/s");
    fprintf(f, "%-15s %-15s %-15.1f %-15s\n", "const", "kB         ", kB, "[hbar]/nK");
    fprintf(f, "%-15s %-15s %-15.6E %-15s\n", "const", "abohr      ", abohr, "mu_m");
    fprintf(f, "%-15s %-15s %-15.1f %-15s\n", "const", "npart1      ", p.npart1, "none");
    fprintf(f, "%-15s %-15s %-15.6f %-15s\n", "const", "mass1       ", p.mass1, "u");
    fprintf(f, "%-15s %-15s %-15.6f %-15s\n", "const", "a11         ", p.a11 / abohr * p.aho1, "a0");
    fprintf(f, "%-15s %-15s %-15.1f %-15s\n", "const", "npart2      ", p.npart2, "none");
    fprintf(f, "%-15s %-15s %-15.6f %-15s\n", "const", "mass2       ", p.mass2, "u");
    fprintf(f, "%-15s %-15s %-15.6f %-15s\n", "const", "a22         ", p.a22 / abohr * p.aho1, "a0");
    fprintf(f, "%-15s %-15s %-15.6f %-15s\n", "const", "a12         ", p.a12 / abohr * p.aho1, "a0");
    const char* om1[3] = {"omx1        ", "omy1        ", "omz1        "};
    const char* om2[3] = {"omx2        ", "omy2        ", "omz2        "};
    for (int i = 0; i < 3; ++i) fprintf(f, "%-15s %-15s %-15.2f %-15s\n", "const", om1[i], p.omega1[i] / (2 * pi), "2pi");
    for (int i = 0; i < 3; ++i) fprintf(f, "%-15s %-15s %-15.2f %-15s\n", "const", om2[i], p.omega2[i] / (2 * pi), "2pi");
    fprintf(f, "%-15s %-15s %-15.6f %-15s\n", "const", "aho1        ", p.aho1, "?");
    fprintf(f, "%-15s %-15s %-15.6f %-15s\n", "const", "ax1         ", p.ax1[0], "code");
    fprintf(f, "%-15s %-15s %-15.6f %-15s\n", "const", "ay1         ", p.ax1[1], "code");
    fprintf(f, "%-15s %-15s %-15.6f %-15s\n", "const", "az1         ", p.ax1[2], "code");
    fprintf(f, "%-15s %-15s %-15.6f %-15s\n", "const", "omho1       ", p.omho1, "?");
    fprintf(f, "%-15s %-15s %-15.1f %-15s\n", "const", "Eho1        ", p.omho1 * hbar, "Energy_scale");
    fprintf(f, "%-15s %-15s %-15.6f %-15s\n", "const", "aho2        ", p.aho2, "?");
    fprintf(f, "%-15s %-15s %-15.6f %-15s\n", "const", "ax2         ", p.ax2[0], "code");
    fprintf(f, "%-15s %-15s %-15.6f %-15s\n", "const", "ay2         ", p.ax2[1], "code");
    fprintf(f, "%-15s %-15s %-15.6f %-15s\n", "const", "az2         ", p.ax2[2], "code");
    fprintf(f, "%-15s %-15s %-15.6f %-15s\n", "const", "omho2       ", p.omho2, "?");
    fprintf(f, "%-15s %-15s %-15.1f %-15s\n", "const", "Eho2        ", p.omho2 * hbar, "Energy_scale");
    fprintf(f, "%-15s %-15s %-15d %-15s\n", "const", "itmod      ", p.itmod, ".");
    fprintf(f, "%-15s %-15s %-15.6f %-15s\n", "const", "aho        ", p.aho1, "?");
    fprintf(f, "%-15s %-15s %-15.1f %-15s\n", "const", "Eho        ", p.omho1 * hbar, "Energy_scale");
    fprintf(f, "%-15s %-15s %-15.6f %-15s\n", "const", "omho       ", p.omho1, "?");
    fclose(f);
}

void Output::WriteVariable2WTXT(const char* name, const double& value, const char* unit)
{
    FILE* f = fopen((prefix + "info.wtxt").c_str(), "a");
    if (f == NULL) { perror("Wtxt file opening failed"); return; }
    fprintf(f, "%-15s %-15s %-15.6f %-15s\n", "const", name, value, unit);
    fclose(f);
}

void Output::WriteVariable2WTXT(const char* name, const int& value, const char* unit)
{
    FILE* f = fopen((prefix + "info.wtxt").c_str(), "a");
    if (f == NULL) { perror("Wtxt file opening failed"); return; }
    fprintf(f, "%-15s %-15s %-15d %-15s\n", "const", name, value, unit);
    fclose(f);
}

void Output::WritePsi(const complex* psi, const int length, const std::string number)
{
    Write3DMatrix(psi, length, number + "psi.wdat", "ab");
}
void Output::WritePsiInit(const complex* psi, const int length, const std::string number)
{
    Write3DMatrix(psi, length, number + "psi_init.wdat");
}
void Output::WritePsiFinal(const complex* psi, const int length, const std::string number)
{
    Write3DMatrix(psi, length, number + "psi_final.wdat");
}
void Output::WriteVext(const complex* vext, const int length, const std::string number)
{
    Write3DMatrix(vext, length, number + "vext.wdat");
}

void Output::CreateStatusFile(const char* header)
{
    status = fopen((prefix + "status.out").c_str(), "w");
    if (status) fprintf(status, "%s", header);
}
void Output::CreateRealStatusFile(const char* header)
{
    status_real = fopen((prefix + "status.out").c_str(), "w");
    if (status_real) fprintf(status_real, "%s", header);
}
void Output::CloseStatusFile()
{
    if (status) fclose(status);
    status = nullptr;
}
void Output::CloseRealStatusFile()
{
    if (status_real) fclose(status_real);
    status_real = nullptr;
}

void Output::WriteStatus(const char* line)
{
    status_count++;
    if (status_count % 10 == 1) {
        CloseStatusFile();
        status = fopen((prefix + "status.out").c_str(), "a");
    }
    if (status) fprintf(status, "%s", line);
}

void Output::WriteStatusReal(const char* line)
{
    status_real_count++;
    if (status_real_count % 10 == 1) {
        CloseRealStatusFile();
        status_real = fopen((prefix + "status.out").c_str(), "a");
    }
    if (status_real) fprintf(status_real, "%s", line);
}

void Output::Write3DMatrix(const complex* matrix, const int length, const std::string& filename, const char* mode)
{
    FILE* f = fopen((prefix + filename).c_str(), mode);
    if (f == NULL) {
        fprintf(stderr, "%s: Error: Could not open file!\n", filename.c_str());
        return;
    }
    size_t n = fwrite(matrix, sizeof(complex), length, f);
    if (n != (size_t)length) fprintf(stderr, "Error: Only %zu elements were written instead of %d!\n", n, length);
    fclose(f);
}

void Output::Write3DMatrix(const double* matrix, const int length, const std::string& filename, const char* mode)
{
    FILE* f = fopen((prefix + filename).c_str(), mode);
    if (f == NULL) {
        fprintf(stderr, "%s: Error: Could not open file!\n", filename.c_str());
        return;
    }
    size_t n = fwrite(matrix, sizeof(double), length, f);
    if (n != (size_t)length) fprintf(stderr, "Error: Only %zu elements were written instead of %d!\n", n, length);
    fclose(f);
}
