// ref_shim.cpp -- C bindings over the reference's own tools.cpp, compiled from
// the reference sources where they lie (see oracle/Makefile, target _ref).
//
// TEST INFRASTRUCTURE ONLY: loaded by tests/ and tests/golden/make_golden.py to
// pin the oracle's restatement of Get_Alpha (tools.cpp:617-630),
// nEPerpinXZ (tools.cpp:686-709) and build_a_bulk (tools.cpp:497-510).
#include <complex>
#include "Tools.h"

extern "C" {

void ref_get_alpha(double lam, double K, double d, double diel_re, double diel_im,
                   const double* nE, const double* nK, double* out2)
{
    Vector3d e(nE[0], nE[1], nE[2]);
    Vector3d k(nK[0], nK[1], nK[2]);
    std::complex<double> a = Get_Alpha(lam, K, d, std::complex<double>(diel_re, diel_im), e, k);
    out2[0] = a.real();
    out2[1] = a.imag();
}

void ref_nE_perp_in_xz(double theta, double phi, double* out3)
{
    Vector3d v = nEPerpinXZ(theta, phi);
    out3[0] = v(0); out3[1] = v(1); out3[2] = v(2);
}

void ref_build_a_bulk(int Nx, int Ny, int Nz, int* out)
{
    VectorXi r = build_a_bulk(Nx, Ny, Nz);
    for (int i = 0; i < r.size(); i++) out[i] = r(i);
}

}
