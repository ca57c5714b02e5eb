// CoreStructure.h -- the geometry/design data contract the solve path reads
// (CoreStructure.h:12-50 of the reference): dipole count N, box extents
// Nx/Ny/Nz, lattice spacing d, integer lattice coordinates R (3N) and the
// per-component design values diel_old (3N, t in [0, 1] or material labels).
//
// Built from a SpacePara as the reference does (the optimisation tasks), or
// directly from R and diel_old.
#pragma once
#include <stdexcept>
#include <string>

#include "SpacePara.h"
#include "dda_types.h"

namespace topodda {

class CoreStructure {
public:
    // CoreStructure.cpp:6-29: sites, box and initial design from a
    // SpacePara (diel_old of site i = Para(geometryPara(i)))
    CoreStructure(SpacePara* spacepara_, double d_);
    CoreStructure(int Nx, int Ny, int Nz, double d, const VectorXi& R, const VectorXd& diel_old);
    // Full Nx x Ny x Nz box in build_a_bulk order (tools.cpp:497-510).
    static CoreStructure Bulk(int Nx, int Ny, int Nz, double d, const VectorXd& diel_old);
    static VectorXi build_a_bulk(int Nx, int Ny, int Nz);

    // CoreStructure.cpp:31-104: one optimisation step of the SpacePara's
    // free parameters (step has one value per free parameter): with a
    // filter, beta scheduled at current_it of Max_it, Para_origin += step
    // clamped to [0, 1], cone filter over FreeWeight, SmoothDensity; without,
    // Para += step clamped; then diel_old from Para.  Needs the SpacePara
    // constructor; a size mismatch throws.
    void UpdateStr(VectorXd step, int current_it, int Max_it);
    // Replace the design (the effect of UpdateStr, CoreStructure.cpp:99-109).
    void UpdateStr(const VectorXd& diel_old_new);
    void UpdateStrSingle(int idx, double value);  // sets the 3 components of dipole idx

    // The reference's text files (CoreStructure.cpp:155-190):
    // "CoreStructure.txt" with Nx, Ny, Nz, N, R, diel_old, d; and save_position
    // + "CoreStructure" + iteration + ".txt", the same for mode "normal",
    // diel_old only for any other mode.
    void output_to_file();
    void output_to_file(std::string save_position, int iteration, std::string mode);

    int get_N() const { return N; }
    int get_Nx() const { return Nx; }
    int get_Ny() const { return Ny; }
    int get_Nz() const { return Nz; }
    double get_d() const { return d; }
    VectorXi* get_R() { return &R; }
    VectorXd* get_diel_old() { return &diel_old; }
    VectorXd* get_diel_old_max() { return &diel_old_max; }
    SpacePara* get_spacepara() { return spacepara; }

private:
    void diel_from_para();
    SpacePara* spacepara = nullptr;
    int N, Nx, Ny, Nz;
    double d;
    VectorXi R;
    VectorXd diel_old;
    VectorXd diel_old_max;
};

}  // namespace topodda

// The reference's code names these classes without a namespace; define
// TOPODDA_NO_GLOBAL_NAMES to keep them in namespace topodda only.
#ifndef TOPODDA_NO_GLOBAL_NAMES
using topodda::CoreStructure;
#endif
