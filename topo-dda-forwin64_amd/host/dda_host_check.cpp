// dda_host_check -- drives the host classes (CoreStructure, AProductCore,
// DDAModel, EvoDDAModel) through one forward and one adjoint solve and one
// Aproductwithalb, writing the results for comparison with the oracle.
//
// usage: dda_host_check OUT.bin Nx Ny Nz d lam seed [MAXm MAXn Lm Ln] [LDR|FCD]
// (the optional numbers select the periodic AProductCore overload,
// AProductCore.cpp:251-536; the last word the interaction tensor, LDR by
// default)
// OUT.bin layout (little endian): int32 N, int32 it_fwd, int32 it_adj, int32 grid[3],
// then complex128 vectors P_fwd, lambda_adj, Aproductwithalb(E), al, E, diel (3N each),
// then int32 adjoint target count and the targets.
// Beside OUT.bin it writes the reference's text files of the forward solve
// (EvoDDAModel.cpp:428-447 order): CoreStructure0.txt (mode "normal"),
// Model_results0it0.txt (EResult = Einternal, then P), and
// Model_resultsit1.txt: EResult of a model with RResult = 6 points around
// the box, from the forward P.
#include <cstdio>
#include <cstdlib>
#include <random>
#include <string>

#include "EvoDDAModel.h"

using namespace topodda;

int main(int argc, char** argv)
{
    if (argc < 8) {
        std::fprintf(stderr, "usage: %s OUT.bin Nx Ny Nz d lam seed [MAXm MAXn Lm Ln] [LDR|FCD]\n", argv[0]);
        return 2;
    }
    const int Nx = atoi(argv[2]), Ny = atoi(argv[3]), Nz = atoi(argv[4]);
    const double d = atof(argv[5]), lam = atof(argv[6]);
    const unsigned seed = (unsigned)atoi(argv[7]);
    const int N = Nx * Ny * Nz;
    std::mt19937_64 rng(seed);
    std::uniform_real_distribution<double> U(0.0, 1.0);
    VectorXd diel_old = VectorXd::Zero(3 * (std::size_t)N);
    for (int i = 0; i < N; i++) {
        const double t = U(rng);
        for (int c = 0; c < 3; c++) diel_old(3 * (std::size_t)i + c) = t;
    }
    CoreStructure CStr = CoreStructure::Bulk(Nx, Ny, Nz, d, diel_old);
    VectorXcd material = VectorXcd::Zero(2);
    material(0) = cd(1.0, 0.0);
    material(1) = cd(2.25, 0.01);
    std::string method = "LDR";
    int nargs = argc;
    if (argc > 8 && (std::string(argv[argc - 1]) == "LDR" || std::string(argv[argc - 1]) == "FCD"))
        method = argv[--nargs];
    const bool periodic = nargs >= 12;
    AProductCore Core = periodic ? AProductCore(&CStr, lam, material, 1.0, atoi(argv[8]), atoi(argv[9]), atof(argv[10]),
                                                atof(argv[11]), method)
                                 : AProductCore(&CStr, lam, material, 1.0, method);
    DDAModel model(&Core, Vector3d(0, 0, 1), 1.0, Vector3d(1, 0, 0), false);
    EvoDDAModel evo(&CStr, {&model});
    evo.ForwardSolve(1000, 1e-10);
    const int it_fwd = model.get_ITERATION();
    VectorXcd Pf = *model.get_P();
    std::string dir = argv[1];
    dir = dir.find('/') == std::string::npos ? std::string() : dir.substr(0, dir.rfind('/') + 1);
    CStr.output_to_file(dir, 0, "normal");
    model.update_E_in_structure();
    model.solve_E();
    model.output_to_file(dir, 0, 0);
    {
        VectorXi pts(18);
        const int p[18] = {-2, 0, 0, Nx + 1, 1, 0, 0, -3, 1, 1, Ny + 2, Nz - 1, 0, 0, -2, Nx / 2, Ny / 2, Nz + 3};
        for (int k = 0; k < 18; k++) pts(k) = p[k];
        DDAModel probe(&Core, Vector3d(0, 0, 1), 1.0, Vector3d(1, 0, 0), &pts, false);
        probe.InitializeP(Pf);
        probe.solve_E();
        probe.output_to_file(dir, 1);
    }
    // adjoint source: conj(E_int) at 8 targets
    VectorXcd devp = VectorXcd::Zero(3 * (std::size_t)N);
    std::vector<int> targets;
    for (int k = 0; k < 8; k++) targets.push_back((int)(rng() % (unsigned)N));
    for (int t : targets)
        for (int c = 0; c < 3; c++) devp(3 * (std::size_t)t + c) = std::conj((*model.get_Einternal())(3 * (std::size_t)t + c));
    VectorXcd lambdaT = evo.AdjointSolve(0, devp, 1000, 1e-10);
    const int it_adj = model.get_ITERATION();
    VectorXcd E = *model.get_E();
    VectorXcd y = model.Aproductwithalb(E);
    FILE* f = std::fopen(argv[1], "wb");
    if (!f) return 3;
    int hdr[6] = {N, it_fwd, it_adj, 0, 0, 0};
    check(tdda_op_info(Core.handle(), hdr + 3, nullptr));
    std::fwrite(hdr, sizeof(int), 6, f);
    for (const VectorXcd* v : {&Pf, &lambdaT, &y, model.get_al(), &E, model.get_diel()})
        std::fwrite(v->data(), sizeof(cd), v->size(), f);
    const int nt = (int)targets.size();
    std::fwrite(&nt, sizeof(int), 1, f);
    std::fwrite(targets.data(), sizeof(int), nt, f);
    std::fclose(f);
    std::printf("{\"N\": %d, \"it_fwd\": %d, \"it_adj\": %d}\n", N, it_fwd, it_adj);
    return 0;
}
