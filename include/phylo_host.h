/*
 * phylo_host.h -- C entry points of libphylommand_host (the C++ host layer
 * above libphylo_hip): discrete-gamma rates and GTR+Gamma model fitting with
 * the library's own Nelder-Mead.  They take over the model-parameter
 * optimisation treeator runs through NLopt (treeator.cpp:738-911, objective
 * opt_function / change_non_fixed at :1122-1182), which is not available in
 * this build.
 */
#ifndef PHYLO_HOST_H
#define PHYLO_HOST_H

#include "phylo_hip.h"

#ifdef __cplusplus
extern "C" {
#endif

/* Mean rates of n_cat equal-probability discrete-gamma categories with shape
 * alpha and mean 1 (Yang 1994).  New: the reference has no rate variation. */
int phylo_gamma_rates(double alpha, int n_cat, double *rates);

/* GTR+Gamma on a context that holds tips and topology.  x (10 values):
 * log exchangeabilities A-G, A-C, A-T, G-C, G-T (C-T = 1), frequency logits
 * of A, G, C (T = 0), log gamma shape, log branch-length scale; Q is
 * normalised to mean rate 1 and edge lengths are rel_len[node] * scale.
 * phylo_gtr_gamma_lnl evaluates one point (-inf when invalid);
 * phylo_fit_gtr_gamma maximises from x0 under the Nelder-Mead options nm
 * (NULL: phylo_nm_defaults).  After the first run it probes each coordinate
 * reset to its x0 value and reruns the full simplex from the best probe that
 * gains more than ftol_abs (up to three rounds), then restarts from the
 * optimum with small simplices until a restart gains less than ftol_abs, and
 * returns the optimum, its lnL,
 * the number of likelihood evaluations and the wall time. */
typedef struct phylo_nm_options {
    int max_evals;        /* likelihood evaluations, restarts included */
    double ftol_abs;      /* converged when the simplex's lnL spread is below this ... */
    double xtol_rel;      /* ... and its relative extent below sqrt(xtol_rel) */
    double initial_step;  /* simplex edge per coordinate */
    int adaptive;         /* 1: dimension-dependent coefficients (Gao & Han 2012) */
    int max_restarts;
    const double *steps;  /* per-coordinate simplex edges (one per parameter), NULL: initial_step */
    double restart_step;  /* largest simplex edge of a restart */
} phylo_nm_options;
int phylo_nm_defaults(phylo_nm_options *opt);
double phylo_gtr_gamma_lnl(phylo_ctx *ctx, int n_nodes, const double *rel_len, int n_cat, const double *x);
int phylo_fit_gtr_gamma(phylo_ctx *ctx, int n_nodes, const double *rel_len, int n_cat, const double *x0,
                        const phylo_nm_options *nm, double *x_out, double *lnl, int *evals, double *seconds);

#ifdef __cplusplus
}
#endif

#endif /* PHYLO_HOST_H */
