#include "CoreStructure.h"

#include "AProductCore.h"

namespace topodda {

CoreStructure::CoreStructure(int Nx_, int Ny_, int Nz_, double d_, const VectorXi& R_, const VectorXd& diel_old_)
    : N((int)(R_.size() / 3)), Nx(Nx_), Ny(Ny_), Nz(Nz_), d(d_), R(R_), diel_old(diel_old_), diel_old_max(diel_old_)
{
    if (R_.size() % 3 != 0) throw std::invalid_argument("R must hold 3N coordinates");
    if (diel_old_.size() != R_.size()) throw std::invalid_argument("diel_old must hold 3N values");
}

VectorXi CoreStructure::build_a_bulk(int Nx, int Ny, int Nz)
{
    VectorXi r = VectorXi::Zero(3 * (std::size_t)Nx * Ny * Nz);
    for (int x = 0; x < Nx; x++)
        for (int y = 0; y < Ny; y++)
            for (int z = 0; z < Nz; z++) {
                const std::size_t pos = (std::size_t)z + Nz * ((std::size_t)y + Ny * x);
                r(3 * pos) = x;
                r(3 * pos + 1) = y;
                r(3 * pos + 2) = z;
            }
    return r;
}

CoreStructure CoreStructure::Bulk(int Nx, int Ny, int Nz, double d, const VectorXd& diel_old)
{
    return CoreStructure(Nx, Ny, Nz, d, build_a_bulk(Nx, Ny, Nz), diel_old);
}

void CoreStructure::UpdateStr(const VectorXd& diel_old_new)
{
    if (diel_old_new.size() != diel_old.size()) throw std::invalid_argument("diel_old size mismatch");
    diel_old = diel_old_new;
}

void CoreStructure::UpdateStrSingle(int idx, double value)
{
    if (idx < 0 || idx >= N) throw std::out_of_range("dipole index");
    diel_old(3 * idx) = value;
    diel_old(3 * idx + 1) = value;
    diel_old(3 * idx + 2) = value;
}

void CoreStructure::output_to_file()
{
    check(tdda_write_structure("CoreStructure.txt", Nx, Ny, Nz, N, R.data(), diel_old.data(), d, 1));
}

void CoreStructure::output_to_file(std::string save_position, int iteration, std::string mode)
{
    const std::string name = save_position + "CoreStructure" + std::to_string(iteration) + ".txt";
    check(tdda_write_structure(name.c_str(), Nx, Ny, Nz, N, R.data(), diel_old.data(), d, mode == "normal" ? 1 : 0));
}

}  // namespace topodda
