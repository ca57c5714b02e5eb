#include "Structure.h"

#include <stdexcept>

namespace topodda {

MatrixXi find_scope_3_dim(VectorXi* x)
{
    const int N = (int)(x->size() / 3);
    MatrixXi s(3, 2);
    if (N == 0) return s;
    for (int a = 0; a < 3; a++) s(a, 0) = s(a, 1) = (*x)(a);
    for (int i = 0; i < N; i++)
        for (int a = 0; a < 3; a++) {
            const int v = (*x)(3 * (std::size_t)i + a);
            if (s(a, 0) >= v) s(a, 0) = v;
            if (s(a, 1) <= v) s(a, 1) = v;
        }
    return s;
}

Structure::Structure(VectorXi* total_space, VectorXi* geometry_, bool para_cond_) : para_cond(para_cond_)
{
    if (!total_space || !geometry_) throw std::invalid_argument("Structure: null geometry");
    cut(total_space, geometry_);
}

void Structure::cut(VectorXi* big, VectorXi* smalll)
{
    const MatrixXi sc = find_scope_3_dim(big);
    const int n = (int)(smalll->size() / 3);
    std::vector<int> in;
    for (int i = 0; i < n; i++) {
        bool inside = true;
        for (int a = 0; a < 3; a++) {
            const int v = (*smalll)(3 * (std::size_t)i + a);
            if (v < sc(a, 0) || v > sc(a, 1)) inside = false;
        }
        if (inside) in.push_back(i);
    }
    geometry = VectorXi::Zero(3 * in.size());
    for (std::size_t k = 0; k < in.size(); k++)
        for (int a = 0; a < 3; a++) geometry(3 * k + a) = (*smalll)(3 * (std::size_t)in[k] + a);
}

Space::Space(VectorXi* total_space_, int Nx_, int Ny_, int Nz_, int N_, std::vector<Structure>* ln_)
    : total_space(total_space_), Nx(Nx_), Ny(Ny_), Nz(Nz_), N(N_), ln(ln_)
{
}

Space operator+(const Space& s1, Structure& s2)
{
    s1.ln->push_back(s2);
    return Space(s1.total_space, s1.Nx, s1.Ny, s1.Nz, s1.N + s2.get_geometry_size(), s1.ln);
}

}  // namespace topodda
