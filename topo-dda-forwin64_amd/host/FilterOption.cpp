#include "filterOption.h"

#include <cmath>
#include <stdexcept>

#include "topodda.h"

namespace topodda {

FilterOption::FilterOption(double beta_min_, double beta_max_, double ita_, std::string beta_type_,
                           std::vector<filterinfo> rfilterlist_, bool fixit_, int MAX_ITERATION_FIXED_)
    : beta(beta_min_), ita(ita_), beta_type(beta_type_), beta_min(beta_min_), beta_max(beta_max_),
      rfilterlist(rfilterlist_), fixit(fixit_), MAX_ITERATION_FIXED(MAX_ITERATION_FIXED_)
{
    if (!rfilterlist.empty()) rfilter = rfilterlist[0].rfilter;
}

void FilterOption::update_beta(const int iteration, const int Max_iteration)
{
    const int max_real = fixit ? MAX_ITERATION_FIXED : Max_iteration;
    int kind;
    if (beta_type == "exp") kind = TDDA_BETA_EXP;
    else if (beta_type == "piecewise") kind = TDDA_BETA_PIECEWISE;
    else if (beta_type == "linear") kind = TDDA_BETA_LINEAR;
    else throw std::invalid_argument("FilterOption::update_beta: beta_type not defined");
    beta = tdda_beta_update(kind, iteration, max_real, beta_min, beta_max);
}

double FilterOption::SmoothDensity(double input) const
{
    if (input <= ita && input >= 0.0) return ita * (std::exp(-beta * (1 - input / ita)) - (1 - input / ita) * std::exp(-beta));
    if (input > ita && input <= 1.0)
        return (1 - ita) * (1 - std::exp(-beta * (input - ita) / (1 - ita)) + (input - ita) / (1 - ita) * std::exp(-beta)) +
               ita;
    throw std::out_of_range("FilterOption::SmoothDensity: input out of range");
}

bool FilterOption::filterchange(int iteration)
{
    for (const filterinfo& f : rfilterlist)
        if (iteration == f.iteration) {
            rfilter = f.rfilter;
            return true;
        }
    return false;
}

double PtoFderivative(const double input, const double beta, const double ita)
{
    if (input <= ita && input >= 0.0) return beta * std::exp(-beta * (1 - input / ita)) + std::exp(-beta);
    if (input > ita && input <= 1.0) return beta * std::exp(-beta * (input - ita) / (1 - ita)) + std::exp(-beta);
    throw std::out_of_range("PtoFderivative: input out of range");
}

}  // namespace topodda
