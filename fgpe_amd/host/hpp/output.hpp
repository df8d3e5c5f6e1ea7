// output.hpp -- file output (hpp/output.hpp): raw complex<double> .wdat
// arrays, the info.wtxt key/value file and the status.out table.
#ifndef FGPE_HOST_OUTPUT_HPP
#define FGPE_HOST_OUTPUT_HPP

#include "params.hpp"

class Output {
private:
    std::string prefix;
    FILE* status = nullptr;
    FILE* status_real = nullptr;
    int status_count = 0;
    int status_real_count = 0;

public:
    std::string getPrefix() const { return prefix; }
    Output(const std::string& prefix);
    Output(const Output& other) : prefix(other.getPrefix()) {}
    Output& operator=(const Output& other)
    {
        if (this != &other) prefix = other.getPrefix();
        return *this;
    }

    void WriteVariable2WTXT(const char* name, const double& value, const char* unit);
    void WriteVariable2WTXT(const char* name, const int& value, const char* unit);
    void WriteWtxt(const Params& p, const int& ncycle);
    void WriteWtxt_2c(const Params2& p, const int& ncycle);
    void WritePsi(const complex* psi, const int length, const std::string number = "");
    void WritePsiInit(const complex* psi, const int length, const std::string number = "");
    void WritePsiFinal(const complex* psi, const int length, const std::string number = "");
    void WriteVext(const complex* vext, const int length, const std::string number = "");
    void CreateStatusFile(const char* header);
    void CreateRealStatusFile(const char* header);
    void CloseStatusFile();
    void CloseRealStatusFile();
    void WriteStatus(const char* line);
    void WriteStatusReal(const char* line);
    void Write3DMatrix(const double* matrix, const int length, const std::string& filename, const char* mode = "wb");
    void Write3DMatrix(const complex* matrix, const int length, const std::string& filename, const char* mode = "wb");
};

#endif
