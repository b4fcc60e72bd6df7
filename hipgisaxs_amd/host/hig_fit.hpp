// hig_fit.hpp -- the fit drivers of a .hig model (particle swarm, brute force, LMVM, plain
// objective) on HipGISAXS::fit_objective; the parameter-path updates live in HigInput::update_param.
#pragma once

#include <cstdint>
#include <string>
#include <vector>

#include "hig_sim.hpp"

namespace hig {

// HGX_METRIC_* of a distancemetric name: the reference's sqrt_unit_l2, sqrt_c_l2, cbrt_unit_l2,
// cbrt_c_l2 (src/analyzer/hipgisaxs_fit_main.cpp:79-101), and chi2 and l2; -1 for any other name.
int metric_of(const std::string& name);

// Runs fitting algorithm `algo_index` of the input (in file order) from its fitparam ranges and
// initial values; `seed` seeds the particle swarm's generator (the reference seeds it with the
// clock). best: the best parameter vector found, fbest its objective, evals the objective calls.
bool fit_run(HipGISAXS& sim, int algo_index, uint64_t seed, std::vector<double>& best, double& fbest,
             int64_t& evals, std::string& err);

}  // namespace hig
