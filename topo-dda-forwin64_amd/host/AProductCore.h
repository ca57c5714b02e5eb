// AProductCore.h -- the reference's operator class (AProductCore.h:51-80),
// backed by the device operator of libtopodda_hip.so.
//
// Construction builds the Green tensor on the device and keeps one octant of
// its spectrum there (AProductCore.cpp:12-249); Aproduct applies the
// off-diagonal interaction on the device (AProductCore.cpp:697-758).  The
// Eigen/cuFFT members of the reference are replaced by an opaque tdda_op.
#pragma once
#include <string>

#include "CoreStructure.h"
#include "dda_types.h"
#include "topodda.h"

namespace topodda {

// Process-wide default device context (device 0 unless set before first use).
tdda_ctx default_context();
void set_default_device(int device);
// Throws std::runtime_error with tdda_last_error() when rc != TDDA_OK.
void check(int rc);

class AProductCore {
public:
    // AProductCore.cpp:12-249.  AMatrixMethod "LDR" or "FCD" (the filtered
    // tensor, which needs K d < pi - 0.1); any other string throws
    // std::invalid_argument, as the reference's A_dic_generator does
    // (AProductCore.cpp:560-571).  grid = {0,0,0} picks the FFT grid.
    AProductCore(CoreStructure* CStr, double lam, VectorXcd material, double nback, std::string AMatrixMethod,
                 int precisions = TDDA_C128 | TDDA_C64, const int* grid = nullptr);
    // Periodic overload (AProductCore.cpp:251-536): images m = -MAXm..MAXm,
    // n = -MAXn..MAXn shifted by (m Lm, n Ln, 0), damped per image
    // (AProductCore.cpp:573-586).
    AProductCore(CoreStructure* CStr, double lam, VectorXcd material, double nback, int MAXm, int MAXn, double Lm,
                 double Ln, std::string AMatrixMethod, int precisions = TDDA_C128 | TDDA_C64,
                 const int* grid = nullptr);
    ~AProductCore();
    AProductCore(const AProductCore&) = delete;
    AProductCore& operator=(const AProductCore&) = delete;

    // A_offdiag b, no diagonal term (AProductCore.h:56).
    VectorXcd Aproduct(VectorXcd& b);
    // Interaction tensor between sites separated by (x, y, z)
    // (AProductCore.cpp:560-623).
    Matrix3cd A_dic_generator(double x, double y, double z);
    Matrix3cd LDR_inter(double x, double y, double z);
    Matrix3cd FCD_inter(double x, double y, double z);  // AProductCore.cpp:625-695

    CoreStructure* get_CStr() { return CStr; }
    int get_N() { return CStr->get_N(); }
    int get_Nx() { return CStr->get_Nx(); }
    int get_Ny() { return CStr->get_Ny(); }
    int get_Nz() { return CStr->get_Nz(); }
    VectorXi* get_R() { return CStr->get_R(); }
    double get_d() { return CStr->get_d(); }
    double get_lam() { return lam; }
    double get_K() { return K; }
    VectorXd* get_diel_old() { return CStr->get_diel_old(); }
    VectorXd* get_diel_old_max() { return CStr->get_diel_old_max(); }
    VectorXcd* get_material() { return &material; }
    double get_nback() { return nback; }
    tdda_op handle() { return op; }

private:
    void build(const int* periodic_mn, const double* periodic_L, int precisions, const int* grid);
    CoreStructure* CStr;
    double K, lam, nback;
    VectorXcd material;
    std::string AMatrixMethod;
    tdda_op op = nullptr;
};

}  // namespace topodda

// The reference's code names these classes without a namespace; define
// TOPODDA_NO_GLOBAL_NAMES to keep them in namespace topodda only.
#ifndef TOPODDA_NO_GLOBAL_NAMES
using topodda::AProductCore;
#endif
