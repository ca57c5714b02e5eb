// ObjDDAModel.h -- the reference's optimisation objectives (ObjDDAModel.h,
// ObjDDAModel.cpp:5-177) over a DDAModel: PointE, the field magnitude at a
// point, and IntegratedE, log of the weighted field energy in a box.
// GetVal evaluates the objective on the device from the model's solution;
// AdjointSource writes its closed-form adjoint right-hand side 2 dF/dP into
// the model's E on the device (what EvoDDAModel::devp estimates by finite
// differences, EvoDDAModel.cpp:353-391).  SingleResponse / GroupResponse
// are the reference's incremental host forms, used by the finite-difference
// design derivative (EvoDDAModel::devx_and_Adevxp).
#pragma once
#include <string>
#include <vector>

#include "DDAModel.h"

namespace topodda {

class ObjDDAModel {
public:
    virtual ~ObjDDAModel() = default;
    virtual double GetVal() = 0;
    virtual void SingleResponse(int idx, bool deduction) = 0;
    virtual double GroupResponse() = 0;
    virtual void Reset() = 0;
    // 2 dF/dP at the model's current P into the model's E (device and host
    // mirror); returns the value of F at that P
    virtual double AdjointSource() = 0;
    bool Have_Devx = false;
};

// ObjDDAModel.cpp:20-97: parameters (x, y, z), a physical position in the
// unit of d; F = |E_ext(r0) - sum_i A(r0 - d R_i) P_i|.
class ObjPointEDDAModel : public ObjDDAModel {
public:
    ObjPointEDDAModel(std::vector<double> parameters, DDAModel* model_);
    double GetVal() override;
    void SingleResponse(int idx, bool deduction) override;
    double GroupResponse() override;
    void Reset() override;
    double AdjointSource() override;

private:
    void host_sum();
    DDAModel* model;
    double r0[3];
    cd E_ext[3], E_sum[3];
    double value = 0.0;
    bool host_valid = false;
};

// ObjDDAModel.cpp:99-177: parameters (powNum, xMin, xMax, yMin, yMax, zMin,
// zMax, ita, beta); F = log sum over the box's sites of
// SmoothDensity(diel_old) sum_c |al_c P_c|^powNum, the box test written as
// the reference writes it ((lo <= v) <= hi, C semantics).
class ObjIntegratedEDDAModel : public ObjDDAModel {
public:
    ObjIntegratedEDDAModel(std::vector<double> parameters, DDAModel* model_);
    double GetVal() override;
    void SingleResponse(int idx, bool deduction) override;
    double GroupResponse() override;
    void Reset() override;
    double AdjointSource() override;

private:
    DDAModel* model;
    int powNum, box[6];
    double ita, beta;
    double E_int = 0.0;
};

}  // namespace topodda

// The reference's code names these classes without a namespace; define
// TOPODDA_NO_GLOBAL_NAMES to keep them in namespace topodda only.
#ifndef TOPODDA_NO_GLOBAL_NAMES
using topodda::ObjDDAModel;
using topodda::ObjPointEDDAModel;
using topodda::ObjIntegratedEDDAModel;
#endif
