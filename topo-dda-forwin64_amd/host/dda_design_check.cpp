// dda_design_check -- drives the host design classes (Space, Structure,
// FilterOption, SpacePara, CoreStructure) through a design case with no
// device: the parameterization, the cone-filter weights and a sequence of
// CoreStructure::UpdateStr steps, writing every result for comparison with
// the reference's own classes.
//
// usage: dda_design_check CASE.txt OUT.txt
// CASE.txt (whitespace separated):
//   Nx Ny Nz bindz
//   nstruct, then per structure: para n x0 y0 z0 ... (n sites)
//   nin, then nin sites x y z, then nin design values
//   filter beta_min beta_max ita beta_type nfilter (iteration rfilter)*
//   symmetry ax ay          (None or 4fold)
//   periodic Lx Ly
//   nsteps max_it, then per step: iteration change_filter values (NFpara)
// OUT.txt: N Npara NFpara, geometryPara, the weights (count, then weight
// position pairs) of every free parameter, then after the steps Para,
// Para_origin and Para_filtered (filtered designs), diel_old and the final
// weights again; doubles printed with 17 significant digits.
#include <cstdio>
#include <fstream>
#include <iostream>
#include <string>

#include "CoreStructure.h"

using namespace topodda;

namespace {

void write_weights(FILE* f, SpacePara& sp)
{
    for (const auto& w : *sp.get_FreeWeight()) {
        std::fprintf(f, "%zu", w.size());
        for (const WeightPara& q : w) std::fprintf(f, " %.17g %d", q.weight, q.position);
        std::fprintf(f, "\n");
    }
}

void write_vec(FILE* f, const VectorXd& v)
{
    for (std::size_t i = 0; i < v.size(); i++) std::fprintf(f, "%.17g\n", v(i));
}

}  // namespace

int main(int argc, char** argv)
{
    if (argc != 3) {
        std::fprintf(stderr, "usage: %s CASE.txt OUT.txt\n", argv[0]);
        return 2;
    }
    try {
        std::ifstream in(argv[1]);
        int Nx, Ny, Nz, bindz, ns;
        in >> Nx >> Ny >> Nz >> bindz >> ns;
        VectorXi total_space = CoreStructure::build_a_bulk(Nx, Ny, Nz);
        std::vector<Structure> ln;
        std::vector<VectorXi> geoms(ns);
        std::vector<int> paras(ns);
        for (int s = 0; s < ns; s++) {
            int n;
            in >> paras[s] >> n;
            geoms[s] = VectorXi::Zero(3 * (std::size_t)n);
            for (int k = 0; k < 3 * n; k++) in >> geoms[s](k);
        }
        Space S(&total_space, Nx, Ny, Nz, 0, &ln);
        for (int s = 0; s < ns; s++) {
            Structure st(S.get_total_space(), &geoms[s], paras[s] != 0);
            S = S + st;
        }
        int nin;
        in >> nin;
        VectorXi inGeo = VectorXi::Zero(3 * (std::size_t)nin);
        VectorXd inDiel = VectorXd::Zero(3 * (std::size_t)nin);
        for (int k = 0; k < 3 * nin; k++) in >> inGeo(k);
        for (int k = 0; k < nin; k++) {
            double v;
            in >> v;
            for (int c = 0; c < 3; c++) inDiel(3 * (std::size_t)k + c) = v;
        }
        int filter, nfl;
        double bmin, bmax, ita;
        std::string btype;
        in >> filter >> bmin >> bmax >> ita >> btype >> nfl;
        std::vector<filterinfo> fl(nfl);
        for (auto& f : fl) in >> f.iteration >> f.rfilter;
        std::string sym;
        double ax, ay;
        in >> sym >> ax >> ay;
        int periodic, Lx, Ly;
        in >> periodic >> Lx >> Ly;
        FilterOption fo(bmin, bmax, ita, btype, fl);
        SpacePara sp(&S, Vector3i(1, 1, bindz), &inGeo, &inDiel, filter != 0, &fo, sym, std::vector<double>{ax, ay},
                     periodic != 0, Lx, Ly);
        CoreStructure cs(&sp, 1.0);
        FILE* f = std::fopen(argv[2], "w");
        if (!f) return 3;
        const int N = cs.get_N();
        const int NF = (int)sp.get_Free()->size();
        std::fprintf(f, "%d %zu %d\n", N, sp.get_Para()->size(), NF);
        for (int i = 0; i < N; i++) std::fprintf(f, "%d\n", (*sp.get_geometryPara())(i));
        if (filter) write_weights(f, sp);
        int nsteps, max_it;
        in >> nsteps >> max_it;
        for (int k = 0; k < nsteps; k++) {
            int it, change;
            in >> it >> change;
            VectorXd step = VectorXd::Zero(NF);
            for (int i = 0; i < NF; i++) in >> step(i);
            if (change && filter && fo.filterchange(it)) sp.ChangeFilter();
            cs.UpdateStr(step, it, max_it);
        }
        write_vec(f, *sp.get_Para());
        if (filter) {
            write_vec(f, *sp.get_Para_origin());
            write_vec(f, *sp.get_Para_filtered());
        }
        write_vec(f, *cs.get_diel_old());
        if (filter) write_weights(f, sp);
        std::fclose(f);
        return 0;
    } catch (const std::exception& e) {
        std::fprintf(stderr, "dda_design_check: %s\n", e.what());
        return 1;
    }
}
