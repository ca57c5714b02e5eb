// ref_shim.cpp -- C bindings over the reference's own tools.cpp, compiled from
// the reference sources where they lie (see oracle/Makefile, target _ref).
//
// TEST INFRASTRUCTURE ONLY: loaded by tests/ and tests/golden/make_golden.py to
// pin the oracle's restatement of Get_Alpha (tools.cpp:617-630),
// nEPerpinXZ (tools.cpp:686-709), build_a_bulk (tools.cpp:497-510), the beta
// schedules (tools.cpp:776-798) and the filtered design update
// (CoreStructure::UpdateStr, CoreStructure.cpp:31-104, over SpacePara's
// FreeWeight cone filter, SpacePara.cpp:37-48,1087-1270, and
// FilterOption::SmoothDensity, FilterOption.cpp:49-60), and the text result
// files (CoreStructure::output_to_file, CoreStructure.cpp:165-190, through
// the reference's class; DDAModel's complex lines and Eigen prints,
// DDAModel.cpp:589-616, restated with std::ofstream since DDAModel.cpp
// includes CUDA headers).
#include <complex>
#include <cstdio>
#include <fstream>
#include <vector>
#include "Tools.h"
#include "CoreStructure.h"

extern "C" {

void ref_get_alpha(double lam, double K, double d, double diel_re, double diel_im,
                   const double* nE, const double* nK, double* out2)
{
    Vector3d e(nE[0], nE[1], nE[2]);
    Vector3d k(nK[0], nK[1], nK[2]);
    std::complex<double> a = Get_Alpha(lam, K, d, std::complex<double>(diel_re, diel_im), e, k);
    out2[0] = a.real();
    out2[1] = a.imag();
}

void ref_nE_perp_in_xz(double theta, double phi, double* out3)
{
    Vector3d v = nEPerpinXZ(theta, phi);
    out3[0] = v(0); out3[1] = v(1); out3[2] = v(2);
}

void ref_build_a_bulk(int Nx, int Ny, int Nz, int* out)
{
    VectorXi r = build_a_bulk(Nx, Ny, Nz);
    for (int i = 0; i < r.size(); i++) out[i] = r(i);
}

double ref_beta_update(int kind, double x, double x_max, double y_min, double y_max)
{
    if (kind == 0) return exp_update(x, x_max, y_min, y_max);
    if (kind == 1) return piecewise_update(x, x_max, y_min, y_max);
    return linear_update(x, x_max, y_min, y_max);
}

// One filtered design step of an extruded Nx x Ny x Nz box through the
// reference's own classes: one free parameter per (x, y) column (bind
// (1, 1, Nz), no symmetry, not periodic), Para_origin = para0, then
// UpdateStr(step) with the filter radius, beta and eta given.  Outputs the
// per-site diel_old (3N), Para_filtered and Para (Nx Ny each).
int ref_design_update(int Nx, int Ny, int Nz, const double* para0, const double* step, double rfilter, double beta,
                      double eta, double* diel_out, double* filtered_out, double* para_out)
{
    try {
        const int N = Nx * Ny * Nz;
        VectorXi total_space = build_a_bulk(Nx, Ny, Nz);
        VectorXi geometry = build_a_bulk(Nx, Ny, Nz);
        vector<Structure> ln;
        ln.push_back(Structure(&total_space, &geometry, true));
        Space space(&total_space, Nx, Ny, Nz, N, &ln);
        VectorXd diel0 = VectorXd::Zero(3 * N);
        for (int i = 0; i < N; i++) {
            const int x = geometry(3 * i), y = geometry(3 * i + 1);
            diel0(3 * i) = diel0(3 * i + 1) = diel0(3 * i + 2) = para0[x * Ny + y];
        }
        vector<filterinfo> fl{filterinfo{0, rfilter}};
        FilterOption fo(beta, beta, eta, "linear", fl, false, 0);
        SpacePara sp(&space, Vector3i(1, 1, Nz), &geometry, &diel0, true, &fo, "None", vector<double>{}, false, 0, 0);
        CoreStructure cs(&sp, 1.0);
        const int np = (int)sp.get_Free()->size();
        if (np != Nx * Ny) return -1;
        VectorXd st(np);
        for (int i = 0; i < np; i++) st(i) = step[i];
        cs.UpdateStr(st, 0, 1);
        const VectorXd& dl = *cs.get_diel_old();
        for (int i = 0; i < 3 * N; i++) diel_out[i] = dl(i);
        const VectorXd& pf = *sp.get_Para_filtered();
        const VectorXd& pp = *sp.get_Para();
        for (int i = 0; i < np; i++) {
            filtered_out[i] = pf(i);
            para_out[i] = pp(i);
        }
        return 0;
    } catch (...) {
        return -2;
    }
}

// CoreStructure::output_to_file(save_position, iteration, mode)
// (CoreStructure.cpp:165-190) of the extruded Nx x Ny x Nz box built as in
// ref_design_update, with the column design para0 and lattice spacing d;
// normal != 0 selects mode "normal".  Writes save_position +
// "CoreStructure" + iteration + ".txt".
int ref_structure_output(int Nx, int Ny, int Nz, const double* para0, double d, const char* save_position,
                         int iteration, int normal)
{
    try {
        const int N = Nx * Ny * Nz;
        VectorXi total_space = build_a_bulk(Nx, Ny, Nz);
        VectorXi geometry = build_a_bulk(Nx, Ny, Nz);
        vector<Structure> ln;
        ln.push_back(Structure(&total_space, &geometry, true));
        Space space(&total_space, Nx, Ny, Nz, N, &ln);
        VectorXd diel0 = VectorXd::Zero(3 * N);
        for (int i = 0; i < N; i++) {
            const int x = geometry(3 * i), y = geometry(3 * i + 1);
            diel0(3 * i) = diel0(3 * i + 1) = diel0(3 * i + 2) = para0[x * Ny + y];
        }
        vector<filterinfo> fl{filterinfo{0, 1.0}};
        FilterOption fo(0.0, 0.0, 0.5, "linear", fl, false, 0);
        SpacePara sp(&space, Vector3i(1, 1, Nz), &geometry, &diel0, false, &fo, "None", vector<double>{}, false, 0, 0);
        CoreStructure cs(&sp, d);
        cs.output_to_file(string(save_position), iteration, normal ? "normal" : "simple");
        return 0;
    } catch (...) {
        return -2;
    }
}

// DDAModel::output_to_file() (DDAModel.cpp:589-616) restated statement for
// statement on std::ofstream and Eigen types: P lines, n_K, n_E0, the count
// of EResult, RResult, EResult lines.
int ref_model_results_full(const char* path, const double* P, int nP, const double* E, int nE, const int* RR, int nR,
                           const double* nK, const double* nE0)
{
    ofstream fout(path);
    auto lines = [&](const double* z, int n) {
        for (int i = 0; i < n; i++) {
            const complex<double> c(z[2 * i], z[2 * i + 1]);
            if (c.imag() < 0) fout << c.real() << c.imag() << "j" << endl;
            else fout << c.real() << "+" << c.imag() << "j" << endl;
        }
    };
    lines(P, nP);
    fout << Vector3d(nK[0], nK[1], nK[2]) << endl;
    fout << Vector3d(nE0[0], nE0[1], nE0[2]) << endl;
    VectorXcd ER = VectorXcd::Zero(nE);
    fout << ER.size() << endl;
    VectorXi R(nR);
    for (int i = 0; i < nR; i++) R(i) = RR[i];
    fout << R << endl;
    lines(E, nE);
    fout.close();
    return fout.fail() ? -1 : 0;
}


// A design case through the reference's own Space / Structure /
// FilterOption / SpacePara (input-geometry constructor, SpacePara.cpp:
// 1087-1384; ChangeFilter) / CoreStructure (CoreStructure.cpp:6-104):
// parameterization, cone-filter weights and UpdateStr steps.  Reads and
// writes the text formats of topo-dda-forwin64_amd/host/dda_design_check.cpp,
// which drives the build's host classes through the same case.
static void ref_write_weights(FILE* f, SpacePara& sp)
{
    for (auto& w : *sp.get_FreeWeight()) {
        fprintf(f, "%zu", w.size());
        for (auto& q : w) fprintf(f, " %.17g %d", q.weight, q.position);
        fprintf(f, "\n");
    }
}

static void ref_write_vec(FILE* f, const VectorXd& v)
{
    for (int i = 0; i < v.size(); i++) fprintf(f, "%.17g\n", v(i));
}

int ref_spacepara_case(const char* in_path, const char* out_path)
{
    try {
        ifstream in(in_path);
        int Nx, Ny, Nz, bindz, ns;
        in >> Nx >> Ny >> Nz >> bindz >> ns;
        VectorXi total_space = build_a_bulk(Nx, Ny, Nz);
        vector<Structure> ln;
        vector<VectorXi> geoms(ns);
        vector<int> paras(ns);
        for (int s = 0; s < ns; s++) {
            int n;
            in >> paras[s] >> n;
            geoms[s] = VectorXi::Zero(3 * n);
            for (int k = 0; k < 3 * n; k++) in >> geoms[s](k);
        }
        Space S(&total_space, Nx, Ny, Nz, 0, &ln);
        for (int s = 0; s < ns; s++) {
            Structure st(S.get_total_space(), &geoms[s], paras[s] != 0);
            S = S + st;
        }
        int nin;
        in >> nin;
        VectorXi inGeo = VectorXi::Zero(3 * nin);
        VectorXd inDiel = VectorXd::Zero(3 * nin);
        for (int k = 0; k < 3 * nin; k++) in >> inGeo(k);
        for (int k = 0; k < nin; k++) {
            double v;
            in >> v;
            inDiel(3 * k) = inDiel(3 * k + 1) = inDiel(3 * k + 2) = v;
        }
        int filter, nfl;
        double bmin, bmax, ita;
        string btype;
        in >> filter >> bmin >> bmax >> ita >> btype >> nfl;
        vector<filterinfo> fl(nfl);
        for (auto& f : fl) in >> f.iteration >> f.rfilter;
        string sym;
        double ax, ay;
        in >> sym >> ax >> ay;
        int periodic, Lx, Ly;
        in >> periodic >> Lx >> Ly;
        FilterOption fo(bmin, bmax, ita, btype, fl);
        SpacePara sp(&S, Vector3i(1, 1, bindz), &inGeo, &inDiel, filter != 0, &fo, sym, vector<double>{ax, ay},
                     periodic != 0, Lx, Ly);
        CoreStructure cs(&sp, 1.0);
        FILE* f = fopen(out_path, "w");
        if (!f) return -3;
        const int N = cs.get_N();
        const int NF = (int)sp.get_Free()->size();
        fprintf(f, "%d %d %d\n", N, (int)sp.get_Para()->size(), NF);
        for (int i = 0; i < N; i++) fprintf(f, "%d\n", (*sp.get_geometryPara())(i));
        if (filter) ref_write_weights(f, sp);
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
        ref_write_vec(f, *sp.get_Para());
        if (filter) {
            ref_write_vec(f, *sp.get_Para_origin());
            ref_write_vec(f, *sp.get_Para_filtered());
        }
        ref_write_vec(f, *cs.get_diel_old());
        if (filter) ref_write_weights(f, sp);
        fclose(f);
        return 0;
    } catch (...) {
        return -2;
    }
}
}
