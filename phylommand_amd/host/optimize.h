// optimize.h -- model-parameter fitting for the likelihood engine.
//
// The reference maximises the likelihood with NLopt's LN_NELDERMEAD over the
// free rates, the free state frequencies and the rate-model parameters,
// bounds [0, DBL_MAX] for rates and [0, 1] for frequencies, no stopping
// criteria (treeator.cpp:738-911, objective :1122-1182).  NLopt is not part
// of this build; nelder_mead() is an own bounded Nelder-Mead with explicit
// stopping rules, and invalid parameter sets score -infinity (the
// reference's objective returns 0.0, which beats every valid lnL,
// treeator.cpp:1123).
#ifndef PHYLO_AMD_OPTIMIZE_H
#define PHYLO_AMD_OPTIMIZE_H

#include <functional>
#include <vector>

struct phylo_ctx;

namespace phylo_amd {

struct nm_options {
    int max_evals = 2000;
    double ftol_abs = 1e-6;      // stop when the simplex's lnL spread falls below this
    double xtol_rel = 1e-7;      // ... and its extent relative to the best point below sqrt(xtol_rel)
    double initial_step = 0.3;   // simplex edge (per coordinate, absolute)
    std::vector<double> steps;   // per-coordinate edges; empty: initial_step for all
    // Dimension-dependent coefficients (Gao & Han 2012: expansion 1 + 2/n,
    // contraction 0.75 - 1/(2n), shrink 1 - 1/n) instead of 2 / 0.5 / 0.5.
    bool adaptive = false;
    int max_restarts = 4;        // fresh simplices around the optimum (model fits only)
    double restart_step = 0.1;   // their largest edge
};

struct nm_result {
    std::vector<double> x;
    double f = 0.0;              // best (maximised) value
    int evals = 0;
    bool converged = false;
};

// Maximises f over the box [lo, hi] (points outside are clamped).
nm_result nelder_mead(const std::function<double(const std::vector<double>&)>& f, std::vector<double> x0,
                      const std::vector<double>& lo, const std::vector<double>& hi, const nm_options& opt);

// Mean rates of C equal-probability discrete-gamma categories with shape
// alpha and mean 1 (Yang 1994).
std::vector<double> gamma_category_rates(double alpha, int C);
// Regularised lower incomplete gamma P(a, x) and its inverse in x.
double gamma_p(double a, double x);
double gamma_p_inv(double a, double p);

// GTR+Gamma fit on a context that already holds tips and topology.
// x = [log r_AG, log r_AC, log r_AT, log r_GC, log r_GT (r_CT = 1), three
// frequency logits (the fourth is 0), log alpha, log branch-length scale];
// Q is normalised to mean rate 1 and the edge lengths are rel_len * scale.
struct gtr_fit {
    std::vector<double> x;
    double lnl = 0.0;
    int evals = 0;
    double seconds = 0.0;
};
double gtr_gamma_lnl(phylo_ctx* ctx, int n_nodes, const double* rel_len, int n_cat, const std::vector<double>& x);
gtr_fit fit_gtr_gamma(phylo_ctx* ctx, int n_nodes, const double* rel_len, int n_cat, const std::vector<double>& x0,
                      const nm_options& opt);

class simpleML;
class sub_model;
class rate_model;

// treeator's parameter vector for the model fit (treeator.cpp:747-785) and
// its objective (opt_function, opt_rate_in_time, opt_rate_for_clades_function
// through change_non_fixed, treeator.cpp:1122-1182).  Built from the model's
// current values, which are the starting point x0.
struct model_objective {
    model_objective(simpleML& t, sub_model& model, rate_model& rate_mod, const std::vector<unsigned int>& fixed,
                    bool fixed_freq, const std::vector<unsigned int>& fixed_extra, char method);
    // change_non_fixed: writes x into the model and rate model; false when
    // the model rejects a value (a negative rate, frequencies past 1)
    bool apply(const std::vector<double>& x);
    // lnL at x, or -infinity when x is invalid or the likelihood is not
    // finite (the reference's objective returns 0.0 there)
    double operator()(const std::vector<double>& x);

    simpleML& t;
    sub_model& model;
    rate_model& rate_mod;
    std::vector<unsigned int> fixed, fixed_extra;
    bool free_freq;
    char method;
    std::vector<double> x0, lo, hi;
    size_t start_freq = 0, start_extra = 0;
};

// treeator's model-parameter optimisation (treeator.cpp:738-911): the rates
// not in `fixed`, then (reversible models, frequencies not fixed) the first
// n-1 frequencies in [0, 1], then the rate-model parameters not in
// `fixed_extra` (cut-offs bounded by the longest root-to-tip distance for
// methods 't'/'S'); rates and rate-model rates in [0, inf).  method is the
// reference's 'o' (plain), 't'/'S' (rates through time) or 'A' (clade rates).
// The model and rate model are left at the optimum; returns its lnL.
double optimize_parameters(simpleML& t, sub_model& model, rate_model& rate_mod, const std::vector<unsigned int>& fixed,
                           bool fixed_freq, const std::vector<unsigned int>& fixed_extra, char method,
                           const nm_options& opt, int* evals = nullptr);

}  // namespace phylo_amd

#endif
