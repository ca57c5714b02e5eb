// filterOption.h -- the reference's design filter options (filterOption.h,
// FilterOption.cpp:1-86): the projection steepness beta (scheduled from
// beta_min to beta_max: "exp", "piecewise" or "linear", tools.cpp:776-798),
// the threshold ita (eta) of SmoothDensity, and the cone-filter radius,
// which may change at given iterations.
#pragma once
#include <string>
#include <vector>

namespace topodda {

struct filterinfo {
    int iteration;
    double rfilter;
};

class FilterOption {
public:
    FilterOption(double beta_min_, double beta_max_, double ita_, std::string beta_type_,
                 std::vector<filterinfo> rfilterlist_, bool fixit_ = false, int MAX_ITERATION_FIXED_ = 100);
    double get_beta() const { return beta; }
    double get_ita() const { return ita; }
    double get_rfilter() const { return rfilter; }
    // FilterOption.cpp:24-46; an unknown beta_type throws
    void update_beta(const int iteration, const int Max_iteration);
    // FilterOption.cpp:48-62: Heaviside-like projection of a filtered value
    // in [0, 1]; throws outside that range
    double SmoothDensity(double input) const;
    // FilterOption.cpp:72-85: true (and the new radius taken) when the
    // iteration is one of the list's
    bool filterchange(int iteration);

private:
    double beta, ita;
    std::string beta_type;
    double beta_min, beta_max;
    double rfilter = 0.0;
    std::vector<filterinfo> rfilterlist;
    bool fixit;
    int MAX_ITERATION_FIXED;
};

// d SmoothDensity / d input (EvoDDAModel.cpp:1163-1175)
double PtoFderivative(const double input, const double beta, const double ita);

}  // namespace topodda

// The reference's code names these classes without a namespace; define
// TOPODDA_NO_GLOBAL_NAMES to keep them in namespace topodda only.
#ifndef TOPODDA_NO_GLOBAL_NAMES
using topodda::filterinfo;
using topodda::FilterOption;
using topodda::PtoFderivative;
#endif
