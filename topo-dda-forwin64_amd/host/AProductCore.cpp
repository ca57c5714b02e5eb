#include "AProductCore.h"

#include <mutex>
#include <stdexcept>

namespace topodda {

namespace {
std::mutex g_ctx_mu;
tdda_ctx g_ctx = nullptr;
int g_device = 0;
}  // namespace

void check(int rc)
{
    if (rc != TDDA_OK) throw std::runtime_error(std::string("topodda: ") + tdda_last_error());
}

void set_default_device(int device)
{
    std::lock_guard<std::mutex> lk(g_ctx_mu);
    if (g_ctx) throw std::logic_error("default context already created");
    g_device = device;
}

tdda_ctx default_context()
{
    std::lock_guard<std::mutex> lk(g_ctx_mu);
    if (!g_ctx) check(tdda_ctx_create(g_device, &g_ctx));
    return g_ctx;
}

AProductCore::AProductCore(CoreStructure* CStr_, double lam_, VectorXcd material_, double nback_,
                           std::string AMatrixMethod_, int precisions, const int* grid)
    : CStr(CStr_), lam(lam_), nback(nback_), material(material_), AMatrixMethod(AMatrixMethod_)
{
    build(nullptr, nullptr, precisions, grid);
}

AProductCore::AProductCore(CoreStructure* CStr_, double lam_, VectorXcd material_, double nback_, int MAXm, int MAXn,
                           double Lm, double Ln, std::string AMatrixMethod_, int precisions, const int* grid)
    : CStr(CStr_), lam(lam_), nback(nback_), material(material_), AMatrixMethod(AMatrixMethod_)
{
    const int mn[2] = {MAXm, MAXn};
    const double L[2] = {Lm, Ln};
    build(mn, L, precisions, grid);
}

void AProductCore::build(const int* periodic_mn, const double* periodic_L, int precisions, const int* grid)
{
    // AProductCore.cpp:560-571: "FCD" or "LDR", anything else throws
    if (AMatrixMethod != "LDR" && AMatrixMethod != "FCD")
        throw std::invalid_argument("\"" + AMatrixMethod + "\" is not a valid polarizabiliy calculation method.");
    K = 2 * M_PI / (lam / nback);  // AProductCore.cpp:17
    // AProductCore.cpp:19-24 (periodic constructor: AProductCore.cpp:261-267,
    // which tests |eps| < 1.01 instead of <= 1.01)
    double* mat = reinterpret_cast<double*>(material.data());
    check(periodic_mn ? tdda_material_rule_periodic((int)material.size(), mat, nback)
                      : tdda_material_rule((int)material.size(), mat, nback));
    tdda_op_desc D{};
    D.N = CStr->get_N();
    D.Nx = CStr->get_Nx();
    D.Ny = CStr->get_Ny();
    D.Nz = CStr->get_Nz();
    D.R = CStr->get_R()->data();
    D.d = CStr->get_d();
    D.lam = lam;
    D.nback = nback;
    if (grid) {
        D.Mx = grid[0];
        D.My = grid[1];
        D.Mz = grid[2];
    }
    D.precisions = precisions;
    tdda_op_ext E{};
    E.method = AMatrixMethod == "FCD" ? TDDA_METHOD_FCD : TDDA_METHOD_LDR;
    E.max_m = periodic_mn ? periodic_mn[0] : -1;
    E.max_n = periodic_mn ? periodic_mn[1] : -1;
    E.Lm = periodic_L ? periodic_L[0] : 0.0;
    E.Ln = periodic_L ? periodic_L[1] : 0.0;
    check(tdda_op_create_ex(default_context(), &D, &E, &op));
}

AProductCore::~AProductCore()
{
    if (op) tdda_op_destroy(op);
}

VectorXcd AProductCore::Aproduct(VectorXcd& b)
{
    VectorXcd y(b.size());
    check(tdda_aproduct(op, reinterpret_cast<const double*>(b.data()), reinterpret_cast<double*>(y.data()), 1,
                        nullptr, 0, TDDA_C128));
    return y;
}

namespace {
Matrix3cd tensor_matrix(const double* a)
{
    Matrix3cd m;
    const int first[6] = {0, 0, 0, 1, 1, 2};
    const int second[6] = {0, 1, 2, 1, 2, 2};
    for (int l = 0; l < 6; l++) {
        m(first[l], second[l]) = cd(a[2 * l], a[2 * l + 1]);
        m(second[l], first[l]) = m(first[l], second[l]);
    }
    return m;
}
}  // namespace

Matrix3cd AProductCore::LDR_inter(double x, double y, double z)
{
    double a[12];
    check(tdda_green_tensor(K, x, y, z, a));
    return tensor_matrix(a);
}

Matrix3cd AProductCore::FCD_inter(double x, double y, double z)
{
    double a[12];
    check(tdda_green_tensor_fcd(K, CStr->get_d(), x, y, z, a));
    return tensor_matrix(a);
}

Matrix3cd AProductCore::A_dic_generator(double x, double y, double z)
{
    return AMatrixMethod == "FCD" ? FCD_inter(x, y, z) : LDR_inter(x, y, z);
}

}  // namespace topodda
