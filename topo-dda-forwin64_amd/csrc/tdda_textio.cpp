// tdda_textio.cpp -- the reference's text result files, written byte for
// byte as its C++ streams write them, so its plotting scripts read ours.
//
// The reference writes with std::ofstream defaults: numbers in the "%g"
// form with 6 significant digits, and Eigen vectors one coefficient per
// line, each right-aligned with spaces to the widest coefficient of that
// vector (Eigen's default IOFormat; the vendored Eigen's IO.h).  Complex
// entries are written as real part, then "+" unless the imaginary part is
// negative, then the imaginary part and "j" (DDAModel.cpp:589-697).
#include <cstdio>
#include <string>
#include <vector>

#include "topodda.h"

namespace tdda {
int set_error(int code, const char* msg);  // tdda_capi.hip
}

namespace {

using tdda::set_error;

std::string num(double v)
{
    char b[64];
    std::snprintf(b, sizeof b, "%g", v);
    return b;
}

std::string num(int v) { return std::to_string(v); }

// Eigen's default print of a column vector, without the trailing newline.
template <typename T>
void column(std::FILE* f, const T* v, long n)
{
    std::vector<std::string> s((size_t)n);
    size_t w = 0;
    for (long i = 0; i < n; i++) {
        s[(size_t)i] = num(v[i]);
        if (s[(size_t)i].size() > w) w = s[(size_t)i].size();
    }
    for (long i = 0; i < n; i++) {
        if (i) std::fputc('\n', f);
        for (size_t k = s[(size_t)i].size(); k < w; k++) std::fputc(' ', f);
        std::fputs(s[(size_t)i].c_str(), f);
    }
}

void complex_lines(std::FILE* f, const double* z, long n)
{
    for (long i = 0; i < n; i++) {
        const double re = z[2 * i], im = z[2 * i + 1];
        if (im < 0) std::fprintf(f, "%s%sj\n", num(re).c_str(), num(im).c_str());
        else std::fprintf(f, "%s+%sj\n", num(re).c_str(), num(im).c_str());
    }
}

int finish(std::FILE* f)
{
    const bool bad = std::ferror(f) != 0;
    if (std::fclose(f) != 0 || bad) return set_error(TDDA_ERR_IO, "write failed");
    return TDDA_OK;
}

}  // namespace

extern "C" {

int tdda_write_structure(const char* path, int Nx, int Ny, int Nz, int N, const int* R, const double* diel_old,
                         double d, int full)
{
    if (!path || N < 0 || (N > 0 && (!diel_old || (full && !R)))) return set_error(TDDA_ERR_ARG, "bad argument");
    std::FILE* f = std::fopen(path, "w");
    if (!f) return set_error(TDDA_ERR_IO, "cannot open the output file");
    if (full) {
        // CoreStructure.cpp:155-163 and the "normal" mode of 165-178
        std::fprintf(f, "%d\n%d\n%d\n%d\n", Nx, Ny, Nz, N);
        column(f, R, 3L * N);
        std::fputc('\n', f);
    }
    column(f, diel_old, 3L * N);
    std::fputc('\n', f);
    if (full) std::fprintf(f, "%s\n", num(d).c_str());
    return finish(f);
}

int tdda_write_model_results(const char* path, int layout, const double* P, int nP, const double* EResult, int nE,
                             const int* RResult, int nR, const double* nK, const double* nE0)
{
    if (!path || nP < 0 || nE < 0 || nR < 0 || (nP && !P) || (nE && !EResult))
        return set_error(TDDA_ERR_ARG, "bad argument");
    if (layout != TDDA_RESULTS_FULL && layout != TDDA_RESULTS_E_P && layout != TDDA_RESULTS_E)
        return set_error(TDDA_ERR_ARG, "unknown layout");
    if (layout == TDDA_RESULTS_FULL && (!nK || !nE0 || (nR && !RResult)))
        return set_error(TDDA_ERR_ARG, "the full layout needs nK, nE0 and RResult");
    std::FILE* f = std::fopen(path, "w");
    if (!f) return set_error(TDDA_ERR_IO, "cannot open the output file");
    if (layout == TDDA_RESULTS_FULL) {
        // DDAModel.cpp:589-616
        complex_lines(f, P, nP);
        column(f, nK, 3);
        std::fputc('\n', f);
        column(f, nE0, 3);
        std::fprintf(f, "\n%d\n", nE);
        column(f, RResult, nR);
        std::fputc('\n', f);
        complex_lines(f, EResult, nE);
    } else {
        // DDAModel.cpp:618-656 (E then P) and 658-697 (E only)
        complex_lines(f, EResult, nE);
        if (layout == TDDA_RESULTS_E_P) complex_lines(f, P, nP);
    }
    return finish(f);
}

}  // extern "C"
