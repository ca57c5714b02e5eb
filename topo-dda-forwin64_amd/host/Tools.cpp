#include "Tools.h"

#include <fstream>
#include <stdexcept>

#include "CoreStructure.h"

namespace topodda {

VectorXi build_a_bulk(int Nx, int Ny, int Nz) { return CoreStructure::build_a_bulk(Nx, Ny, Nz); }

std::pair<VectorXi, VectorXd> getInputStr(std::string pathCommonData, std::string pathPara)
{
    std::ifstream f1(pathCommonData), f2(pathPara);
    if (!f1 || !f2) throw std::runtime_error("getInputStr: cannot open " + pathCommonData + " or " + pathPara);
    int Nx, Ny, Nz, N;
    f1 >> Nx >> Ny >> Nz >> N;
    VectorXi g = VectorXi::Zero(3 * (std::size_t)N);
    for (std::size_t i = 0; i < 3 * (std::size_t)N; i++) f1 >> g(i);
    VectorXd v = VectorXd::Zero(3 * (std::size_t)N);
    for (std::size_t i = 0; i < 3 * (std::size_t)N; i++) f2 >> v(i);
    return {g, v};
}

std::tuple<int, int, int> getInputNs(std::string pathCommonData)
{
    std::ifstream f(pathCommonData);
    if (!f) throw std::runtime_error("getInputNs: cannot open " + pathCommonData);
    int Nx, Ny, Nz, N;
    f >> Nx >> Ny >> Nz >> N;
    return std::make_tuple(Nx, Ny, Nz);
}

}  // namespace topodda
