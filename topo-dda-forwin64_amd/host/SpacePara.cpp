#include "SpacePara.h"

#include <algorithm>
#include <cmath>
#include <map>
#include <set>
#include <stdexcept>

namespace topodda {

namespace {

using Key3 = std::vector<int>;

std::set<Key3> site_set(VectorXi* g)
{
    std::set<Key3> s;
    const int n = (int)(g->size() / 3);
    for (int i = 0; i < n; i++) {
        Key3 k{(*g)(3 * (std::size_t)i), (*g)(3 * (std::size_t)i + 1), (*g)(3 * (std::size_t)i + 2)};
        if (!s.insert(k).second) throw std::invalid_argument("SpacePara: a lattice site appears twice");
    }
    return s;
}

// Parameter of a new (x, y) column: with 4-fold symmetry a column whose
// mirror images about symaxis already have one shares it
// (SpacePara.cpp:8-35).
void column_param(std::map<std::vector<int>, int>& cur, const std::vector<int>& xy, int& next, const std::string& mode,
                  const std::vector<double>& ax)
{
    if (mode == "None") {
        cur[xy] = next++;
        return;
    }
    const int mx = (int)std::round(2 * ax[0] - xy[0]), my = (int)std::round(2 * ax[1] - xy[1]);
    const std::vector<int> s1{mx, xy[1]}, s2{mx, my}, s3{xy[0], my};
    for (const auto* s : {&s1, &s2, &s3}) {
        auto it = cur.find(*s);
        if (it != cur.end()) {
            cur[xy] = it->second;
            return;
        }
    }
    cur[xy] = next++;
}

}  // namespace

SpacePara::SpacePara(Space* space_, Vector3i bind_, VectorXi* InputGeo, VectorXd* Inputdiel, bool Filter_,
                     FilterOption* Filterstats_, std::string symmetry, std::vector<double> symaxis, bool Periodic_,
                     int Lx_, int Ly_)
    : space(space_), bind(bind_), Filter(Filter_), Periodic(Periodic_), Lx(Lx_), Ly(Ly_)
{
    // the reference's test, with its operator precedence (SpacePara.cpp:1098)
    if ((Filter && bind(0) != 1) || bind(1) != 1)
        throw std::invalid_argument("SpacePara: if Filter = true, bind(0) and bind(1) must be 1");
    int Nx, Ny, Nz, N;
    std::tie(Nx, Ny, Nz, N) = space->get_Ns();
    std::vector<Structure>* ln = space->get_ln();
    geometry = VectorXi::Zero(3 * (std::size_t)N);
    std::size_t n1 = 0;
    std::vector<VectorXi*> F, B;
    for (Structure& s : *ln) {
        VectorXi* g = s.get_geometry();
        for (std::size_t j = 0; j < g->size(); j++) geometry(n1 + j) = (*g)(j);
        n1 += g->size();
        (s.para_or_not() ? F : B).push_back(g);
    }
    if (n1 != 3 * (std::size_t)N) throw std::invalid_argument("SpacePara: the Space's site count does not match its structures");
    geometryPara = VectorXi::Zero(N);
    // free sites (structures must not overlap, SpacePara.cpp:121-126)
    std::set<Key3> Fset;
    for (VectorXi* g : F)
        for (const Key3& k : site_set(g))
            if (!Fset.insert(k).second) throw std::invalid_argument("SpacePara: parameter structures overlap");
    int dividesym;
    if (symmetry == "None") dividesym = 1;
    else if (symmetry == "4fold") dividesym = 4;
    else throw std::invalid_argument("SpacePara: symmetry must be None or 4fold");
    // counts in the reference's integer arithmetic (SpacePara.cpp:1145-1160)
    int nF = 0;
    for (VectorXi* g : F) nF += (int)g->size();
    const int NFpara = nF / 3 / bind(2) / dividesym;
    int cnt = 0;
    for (VectorXi* g : F) {
        ParaDividePos.push_back(cnt);
        cnt += (int)g->size() / 3 / bind(2) / dividesym;
    }
    FreeparatoPara = VectorXi::Zero(NFpara);
    for (int i = 0; i < NFpara; i++) FreeparatoPara(i) = i;
    int Npara = NFpara;
    for (VectorXi* g : B) {
        ParaDividePos.push_back(Npara);
        Npara += (int)(g->size() / 3);
    }
    Para = VectorXd::Zero(Npara);
    std::vector<std::set<Key3>> Bsets;
    std::vector<int> Bnext;
    for (std::size_t i = 0; i < B.size(); i++) {
        Bsets.push_back(site_set(B[i]));
        Bnext.push_back(ParaDividePos[F.size() + i]);
    }
    std::map<std::vector<int>, int> cur;
    int next = 0;
    for (int i = 0; i < N; i++) {
        const Key3 k{geometry(3 * (std::size_t)i), geometry(3 * (std::size_t)i + 1), geometry(3 * (std::size_t)i + 2)};
        if (Fset.count(k)) {
            const std::vector<int> xy{k[0], k[1]};
            if (!cur.count(xy)) column_param(cur, xy, next, symmetry, symaxis);
            geometryPara(i) = cur[xy];
        } else {
            for (std::size_t j = 0; j < Bsets.size(); j++)
                if (Bsets[j].count(k)) geometryPara(i) = Bnext[j]++;
        }
    }
    if (next != NFpara) throw std::invalid_argument("SpacePara: free columns do not match bind(2) and the symmetry");
    Paratogeometry.assign(Npara, {});
    for (int i = 0; i < N; i++) Paratogeometry[geometryPara(i)].push_back(i);
    // initial values (SpacePara.cpp:1214-1230)
    if (InputGeo->size() != Inputdiel->size())
        throw std::invalid_argument("SpacePara: InputGeo and Inputdiel sizes differ");
    std::map<Key3, double> in;
    for (std::size_t i = 0; i < InputGeo->size() / 3; i++)
        in.emplace(Key3{(*InputGeo)(3 * i), (*InputGeo)(3 * i + 1), (*InputGeo)(3 * i + 2)}, (*Inputdiel)(3 * i));
    for (int i = 0; i < Npara; i++) {
        if (Paratogeometry[i].empty()) throw std::invalid_argument("SpacePara: a parameter drives no site");
        const int p = Paratogeometry[i][0];
        auto it = in.find(Key3{geometry(3 * (std::size_t)p), geometry(3 * (std::size_t)p + 1), geometry(3 * (std::size_t)p + 2)});
        if (it != in.end()) Para(i) = it->second;
    }
    if (Filter) {
        Para_origin = Para;
        Para_filtered = Para;
        Filterstats = Filterstats_;
        if (!Filterstats) throw std::invalid_argument("SpacePara: Filter = true needs Filterstats");
        build_weights();
    }
}

// Cone-filter weights of every free parameter (SpacePara.cpp:1232-1380):
// for the origin site of free parameter i (its first site), every site of
// every parameter j in the same z plane within dist^2 <= r^2 - 0.01 adds
// (r - dist, j), in the order j, then the site's place in parameter j, then
// the periodic image (own cell, then (0,-Ly), (-Lx,-Ly), (-Lx,0), (-Lx,Ly),
// (0,Ly), (Lx,Ly), (Lx,0), (Lx,-Ly)).  The sites are found through a
// lattice index instead of the reference's scan over all sites; the order of
// the weights, which the filtered sums follow, is the same.
void SpacePara::build_weights()
{
    const double r = Filterstats->get_rfilter();
    const int N = (int)geometryPara.size();
    const int NFpara = (int)FreeparatoPara.size();
    MatrixXi sc = find_scope_3_dim(&geometry);
    const int X0 = sc(0, 0), Y0 = sc(1, 0), Z0 = sc(2, 0);
    const int SX = sc(0, 1) - X0 + 1, SY = sc(1, 1) - Y0 + 1, SZ = sc(2, 1) - Z0 + 1;
    std::vector<int> site((std::size_t)SX * SY * SZ, -1), rank(N, 0);
    for (int i = 0; i < N; i++)
        site[((std::size_t)(geometry(3 * (std::size_t)i) - X0) * SY + (geometry(3 * (std::size_t)i + 1) - Y0)) * SZ +
             (geometry(3 * (std::size_t)i + 2) - Z0)] = i;
    for (const auto& sites : Paratogeometry)
        for (std::size_t k = 0; k < sites.size(); k++) rank[sites[k]] = (int)k;
    const int img[9][2] = {{0, 0}, {0, -1}, {-1, -1}, {-1, 0}, {-1, 1}, {0, 1}, {1, 1}, {1, 0}, {1, -1}};
    const int nimg = Periodic ? 9 : 1;
    const int R = (int)std::ceil(r) + 1;
    FreeWeight.assign(NFpara, {});
    struct Cand {
        int j, k, m;
        double w;
    };
    for (int i = 0; i < NFpara; i++) {
        const int po = Paratogeometry[FreeparatoPara(i)][0];
        const int xo = geometry(3 * (std::size_t)po), yo = geometry(3 * (std::size_t)po + 1),
                  zo = geometry(3 * (std::size_t)po + 2);
        std::vector<Cand> c;
        for (int m = 0; m < nimg; m++) {
            const int sx = img[m][0] * Lx, sy = img[m][1] * Ly;
            // sites (xr, yr, zo) with (xr + sx, yr + sy) within r of (xo, yo)
            for (int dx = -R; dx <= R; dx++)
                for (int dy = -R; dy <= R; dy++) {
                    const int xr = xo + dx - sx, yr = yo + dy - sy;
                    if (xr < X0 || xr >= X0 + SX || yr < Y0 || yr >= Y0 + SY) continue;
                    const int p = site[((std::size_t)(xr - X0) * SY + (yr - Y0)) * SZ + (zo - Z0)];
                    if (p < 0) continue;
                    const int x = xr + sx, y = yr + sy;
                    if (double(x - xo) * double(x - xo) + double(y - yo) * double(y - yo) > r * r - 0.01) continue;
                    const double w = r - std::sqrt(double(x - xo) * double(x - xo) + double(y - yo) * double(y - yo));
                    c.push_back(Cand{geometryPara(p), rank[p], m, w});
                }
        }
        std::sort(c.begin(), c.end(), [](const Cand& a, const Cand& b) {
            return a.j != b.j ? a.j < b.j : a.k != b.k ? a.k < b.k : a.m < b.m;
        });
        for (const Cand& q : c) FreeWeight[i].push_back(WeightPara{q.w, q.j});
    }
}

void SpacePara::ChangeFilter()
{
    if (!Filter) throw std::logic_error("SpacePara::ChangeFilter without a filter");
    build_weights();
}

}  // namespace topodda
