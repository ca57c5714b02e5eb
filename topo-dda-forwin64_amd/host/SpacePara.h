// SpacePara.h -- the reference's design parameterization (SpacePara.h):
// which parameter drives each dipole (geometryPara), the parameter values
// (Para, and for filtered designs Para_origin / Para_filtered), the free
// parameters (Free) and, when filtered, each free parameter's cone-filter
// weights (FreeWeight), including the periodic wrap and the 4-fold
// symmetric parameter sharing.  The constructor is the reference's
// input-geometry one (SpacePara.cpp:1087-1384), which the 2D optimisation
// tasks use (main.cpp:349, 457).
#pragma once
#include <string>
#include <vector>

#include "Structure.h"
#include "filterOption.h"

namespace topodda {

struct WeightPara {
    double weight;
    int position;
};

class SpacePara {
public:
    // Sites: the Space's structures in list order; structures marked as
    // parameters become free parameters, one per (x, y) column of bind(2)
    // sites (one per 4 mirrored columns with symmetry "4fold" about
    // symaxis), the others one fixed parameter per site.  Initial values
    // from InputGeo / Inputdiel (first component of each site).  Filter
    // needs bind(0) = bind(1) = 1 and Filterstats; Periodic wraps the filter
    // over periods Lx, Ly (lattice units).  Errors throw.
    SpacePara(Space* space_, Vector3i bind_, VectorXi* InputGeo, VectorXd* Inputdiel, bool Filter_ = false,
              FilterOption* Filterstats_ = nullptr, std::string symmetry = "None",
              std::vector<double> symaxis = std::vector<double>{0.0, 0.0}, bool Periodic_ = false, int Lx_ = 0,
              int Ly_ = 0);
    // SpacePara.cpp:1388-1679: FreeWeight again with Filterstats' current radius
    void ChangeFilter();

    Space* get_space() { return space; }
    VectorXi get_geometry() { return geometry; }
    VectorXi* get_geometryPara() { return &geometryPara; }
    VectorXd* get_Para() { return &Para; }
    VectorXd* get_Para_origin() { return &Para_origin; }
    VectorXd* get_Para_filtered() { return &Para_filtered; }
    Vector3i* get_bind() { return &bind; }
    VectorXi* get_Free() { return &FreeparatoPara; }
    bool get_Filter() const { return Filter; }
    FilterOption* get_Filterstats() { return Filterstats; }
    std::vector<std::vector<WeightPara>>* get_FreeWeight() { return &FreeWeight; }
    std::vector<int>* get_ParaDividePos() { return &ParaDividePos; }
    std::vector<std::vector<int>>* get_Paratogeometry() { return &Paratogeometry; }
    bool get_Periodic() const { return Periodic; }

private:
    void build_weights();
    Space* space;
    VectorXi geometry, geometryPara;
    std::vector<std::vector<int>> Paratogeometry;
    VectorXd Para, Para_origin, Para_filtered;
    Vector3i bind;
    VectorXi FreeparatoPara;
    bool Filter;
    FilterOption* Filterstats = nullptr;
    std::vector<std::vector<WeightPara>> FreeWeight;
    std::vector<int> ParaDividePos;
    bool Periodic;
    int Lx, Ly;
};

}  // namespace topodda

// The reference's code names these classes without a namespace; define
// TOPODDA_NO_GLOBAL_NAMES to keep them in namespace topodda only.
#ifndef TOPODDA_NO_GLOBAL_NAMES
using topodda::WeightPara;
using topodda::SpacePara;
#endif
