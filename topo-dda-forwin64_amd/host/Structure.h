// Structure.h, Space.h -- the reference's geometry containers (Structure.h,
// Space.h): a Structure is a list of lattice sites (3 ints each) marked as
// free design parameters or fixed; a Space is the Nx x Ny x Nz box with the
// list of Structures placed in it.  Only the constructors the optimisation
// tasks use are provided (Structure.cpp:7-10 with Structure::cut,
// Structure.cpp:122-160; Space.cpp:16-23, 55-59); the geometric builders
// (spheres, cylinders, bulk blocks) are set-up code outside the solve path.
#pragma once
#include <tuple>
#include <vector>

#include "dda_types.h"

namespace topodda {

// Scope of 3N coordinates: rows x, y, z; columns min, max (tools.cpp:402-433).
struct MatrixXi {
    int r = 0, c = 0;
    std::vector<int> a;
    MatrixXi() = default;
    MatrixXi(int rows, int cols) : r(rows), c(cols), a((std::size_t)rows * cols, 0) {}
    int& operator()(int i, int j) { return a[(std::size_t)i * c + j]; }
    int operator()(int i, int j) const { return a[(std::size_t)i * c + j]; }
};
MatrixXi find_scope_3_dim(VectorXi* x);

struct Vector3i {
    int v[3] = {0, 0, 0};
    Vector3i() = default;
    Vector3i(int a, int b, int c) : v{a, b, c} {}
    int& operator()(int i) { return v[i]; }
    int operator()(int i) const { return v[i]; }
};

class Structure {
public:
    // Structure.cpp:7-10: the sites of geometry_ that lie inside the scope
    // of total_space, in their given order.
    Structure(VectorXi* total_space, VectorXi* geometry_, bool para_cond_ = false);
    VectorXi* get_geometry() { return &geometry; }
    int get_geometry_size() const { return (int)(geometry.size() / 3); }
    bool para_or_not() const { return para_cond; }

private:
    void cut(VectorXi* big, VectorXi* smalll);
    VectorXi geometry;
    bool para_cond;
};

class Space {
public:
    Space(VectorXi* total_space_, int Nx_, int Ny_, int Nz_, int N_, std::vector<Structure>* ln_);
    VectorXi* get_total_space() { return total_space; }
    int get_ln_size() const { return (int)ln->size(); }
    std::tuple<int, int, int, int> get_Ns() const { return std::make_tuple(Nx, Ny, Nz, N); }
    std::vector<Structure>* get_ln() { return ln; }
    // Space.cpp:55-59: appends s2 to the shared list and counts its sites
    friend Space operator+(const Space& s1, Structure& s2);

private:
    VectorXi* total_space;
    int Nx, Ny, Nz;
    int N;
    std::vector<Structure>* ln;
};

}  // namespace topodda

// The reference's code names these classes without a namespace; define
// TOPODDA_NO_GLOBAL_NAMES to keep them in namespace topodda only.
#ifndef TOPODDA_NO_GLOBAL_NAMES
using topodda::MatrixXi;
using topodda::Vector3i;
using topodda::Structure;
using topodda::Space;
using topodda::find_scope_3_dim;
#endif
