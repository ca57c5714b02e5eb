// Tools.h -- the reference's geometry helpers that the optimisation tasks
// call (tools.cpp:55-98, 402-433, 497-510): the full-box site list in the
// reference's order, the scope of a site list, and the readers of its input
// structure files (a "commondata" file: Nx Ny Nz N, then N sites; a para
// file: 3 values per site).
#pragma once
#include <string>
#include <tuple>
#include <utility>

#include "Structure.h"

namespace topodda {

// tools.cpp:497-510: x slowest, z fastest
VectorXi build_a_bulk(int Nx, int Ny, int Nz);
// tools.cpp:55-83: (sites, design values); a missing file throws
std::pair<VectorXi, VectorXd> getInputStr(std::string pathCommonData, std::string pathPara);
// tools.cpp:85-98
std::tuple<int, int, int> getInputNs(std::string pathCommonData);

}  // namespace topodda

// The reference's code names these functions without a namespace; define
// TOPODDA_NO_GLOBAL_NAMES to keep them in namespace topodda only.
#ifndef TOPODDA_NO_GLOBAL_NAMES
using topodda::build_a_bulk;
using topodda::getInputNs;
using topodda::getInputStr;
#endif
