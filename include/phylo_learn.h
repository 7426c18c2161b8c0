/*
 * phylo_learn.h -- learned starting values for model fitting: a GPU forward
 * simulator fused with alignment feature extraction, and a small FP32 MLP
 * (n_in -> n_hidden -> n_hidden -> n_out, ReLU) trained on the GPU with
 * hand-written forward / backward / Adam kernels.  New: the reference starts
 * every fit from fixed values (treeator.cpp:738-911) and has no learned
 * proposal; the simulator extends simpleML::simulate_chars
 * (simpleML.cpp:233-292) with gamma rate categories, gaps and a seeded
 * counter-based random stream.
 */
#ifndef PHYLO_LEARN_H
#define PHYLO_LEARN_H

#include <stdint.h>

#ifdef __cplusplus
extern "C" {
#endif

#define PHYLO_N_FEATURES 256

/* Simulates n_align DNA alignments (n_taxa <= 32, rooted binary trees of
 * 2 n_taxa - 1 nodes, tips 0..n_taxa-1, children -1 for tips) and writes
 * PHYLO_N_FEATURES features per alignment.  Per alignment: left/right
 * [n_nodes], roots, blen [n_nodes] (substitutions per site; the root entry is
 * ignored), eig [36] = V (4x4 row-major), V^-1 (4x4), eigenvalues (4) with
 * Q = V diag(lam) V^-1, pi [4], cat_rates [n_cat] (equal-probability
 * categories, n_cat <= 8).  gap_frac of tip states become gaps.  masks, when
 * not NULL, receives the tip bitmasks [align][taxon][site] (gap = 15). */
int phylo_sim_features(int device, int n_align, int n_taxa, int n_sites, int n_cat, const int32_t *left,
                       const int32_t *right, const int32_t *roots, const double *blen, const double *eig,
                       const double *pi, const double *cat_rates, double gap_frac, uint64_t seed, float *features,
                       uint32_t *masks);

typedef struct phylo_mlp phylo_mlp;

/* Parameters are one flat FP32 vector [W1 (n_in x n_hidden) | b1 | W2 | b2 |
 * W3 (n_hidden x n_out) | b3], row-major with inputs as rows. */
int phylo_mlp_create(int device, int n_in, int n_hidden, int n_out, int max_batch, uint64_t seed, phylo_mlp **out);
void phylo_mlp_destroy(phylo_mlp *mlp);
const char *phylo_mlp_last_error(const phylo_mlp *mlp);
long phylo_mlp_n_params(const phylo_mlp *mlp);
/* Copies the parameters out (get) and/or in (set); either may be NULL. */
int phylo_mlp_params(phylo_mlp *mlp, float *get, const float *set);
int phylo_mlp_forward(phylo_mlp *mlp, int batch, const float *x, float *y);
/* Mean squared error over batch x n_out and its gradient (flat layout). */
int phylo_mlp_grad(phylo_mlp *mlp, int batch, const float *x, const float *t, float *loss, float *grad);
/* Stores a training set on the device, then runs epochs of shuffled
 * minibatch Adam over it; loss_per_epoch (epochs values) may be NULL. */
int phylo_mlp_set_data(phylo_mlp *mlp, int n, const float *x, const float *t);
int phylo_mlp_train(phylo_mlp *mlp, int epochs, int batch, float lr, uint64_t seed, float *loss_per_epoch);

#ifdef __cplusplus
}
#endif

#endif /* PHYLO_LEARN_H */
