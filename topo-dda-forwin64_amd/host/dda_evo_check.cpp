// dda_evo_check -- the reference's "EvoOpt 2D input" task body
// (main.cpp:291-371) over the host classes: an extruded Nx x Ny x Nz
// structure whose columns are the free parameters, the cone filter,
// CoreStructure, AProductCore, one DDAModel and EvoDDAModel::EvoOptimization
// with the PointE objective and Adam, solves on the device.
//
// usage: dda_evo_check OUTDIR Nx Ny Nz d lam seed iterations MAX_ERROR
// Writes OUTDIR/evo.txt: the initial column values, then the objective of
// every iteration, then Para_origin, Para_filtered and Para after the last
// step (17 significant digits); the reference's own result files go under
// OUTDIR/ as EvoOptimization writes them.
#include <cstdio>
#include <random>
#include <string>

#include "EvoDDAModel.h"

int main(int argc, char** argv)
{
    if (argc != 10) {
        std::fprintf(stderr, "usage: %s OUTDIR Nx Ny Nz d lam seed iterations MAX_ERROR\n", argv[0]);
        return 2;
    }
    const std::string out = argv[1];
    const int Nx = atoi(argv[2]), Ny = atoi(argv[3]), Nz = atoi(argv[4]);
    const double d = atof(argv[5]), lam = atof(argv[6]);
    const unsigned seed = (unsigned)atoi(argv[7]);
    const int iters = atoi(argv[8]);
    const double max_error = atof(argv[9]);
    try {
        VectorXi total_space = CoreStructure::build_a_bulk(Nx, Ny, Nz);
        std::vector<Structure> ln;
        Space S(&total_space, Nx, Ny, Nz, 0, &ln);
        VectorXi inputGeo = total_space;
        const int N = Nx * Ny * Nz;
        VectorXd inputDiel = VectorXd::Zero(3 * (std::size_t)N);
        std::mt19937_64 rng(seed);
        std::uniform_real_distribution<double> U(0.0, 1.0);
        std::vector<double> col((std::size_t)Nx * Ny);
        for (double& v : col) v = U(rng);
        for (int i = 0; i < N; i++) {
            const int x = inputGeo(3 * (std::size_t)i), y = inputGeo(3 * (std::size_t)i + 1);
            for (int c = 0; c < 3; c++) inputDiel(3 * (std::size_t)i + c) = col[(std::size_t)x * Ny + y];
        }
        Structure s1(S.get_total_space(), &inputGeo, true);
        S = S + s1;
        Vector3i bind(1, 1, Nz);
        std::vector<filterinfo> fl{filterinfo{0, 2.0}};
        FilterOption filterOpt(1.0, 50.0, 0.5, "exp", fl);
        SpacePara spacepara(&S, bind, &inputGeo, &inputDiel, true, &filterOpt, "None", std::vector<double>{0.0, 0.0});
        CoreStructure CStr(&spacepara, d);
        VectorXcd material = VectorXcd::Zero(2);
        material(0) = cd(1.0, 0.0);
        material(1) = cd(2.25, 0.01);
        AProductCore Core(&CStr, lam, material, 1.0, "LDR");
        DDAModel TestModel(&Core, Vector3d(0, 0, 1), 1.0, Vector3d(1, 0, 0), false);
        std::vector<DDAModel*> modelPtrs{&TestModel};
        const std::vector<double> objPara{0.5 * d * (Nx - 1), 0.5 * d * (Ny - 1), d * (Nz + 4)};
        EvoDDAModel evoModel("PointE", objPara, 1.0, false, false, false, out + "/", &CStr, modelPtrs);
        evoModel.EvoOptimization(100000, max_error, iters, "Adam");
        FILE* f = std::fopen((out + "/evo.txt").c_str(), "w");
        if (!f) return 3;
        for (double v : col) std::fprintf(f, "%.17g\n", v);
        for (double v : evoModel.get_objective_history()) std::fprintf(f, "%.17g\n", v);
        for (VectorXd* v : {spacepara.get_Para_origin(), spacepara.get_Para_filtered(), spacepara.get_Para()})
            for (std::size_t i = 0; i < v->size(); i++) std::fprintf(f, "%.17g\n", (*v)(i));
        std::fclose(f);
        return 0;
    } catch (const std::exception& e) {
        std::fprintf(stderr, "dda_evo_check: %s\n", e.what());
        return 1;
    }
}
