/*
 * phylo_hip.h -- C ABI of libphylo_hip, the MI355X (gfx950) Felsenstein-pruning
 * likelihood engine that replaces the compute behind phylommand's simpleML.
 *
 * The reference path is a statically linked C++ class interface
 * (simpleML.h:30-142, sub_model.h:28-128, rate_model.h:29-84); its
 * simpleML-compatible C++ host classes (phylommand_amd/host/) sit above this
 * ABI and keep those signatures.  Every entry point below names the
 * reference interface whose work it takes over.
 *
 * Conventions
 *  - Plain pointers and sizes only; every host buffer is caller-owned and
 *    copied before the call returns.  The library owns its device memory and
 *    its HIP stream (one stream per context).
 *  - int status: 0 = ok, negative = error; phylo_last_error() returns text.
 *  - One context per thread; calls on one context are not reentrant.
 *  - Nodes: 0..n_taxa-1 are the tips (node i carries taxon row i of the tip
 *    masks); nodes n_taxa..n_nodes-1 are internal with left >= 0 and
 *    right >= 0, or right == -1 for a single-child node (the reference keeps
 *    such nodes, simpleML.cpp:170-182).
 *  - Tip states are bit masks, bit i = state i (the reference's
 *    bitset<SIZE> per site, character_vector.h:88); ambiguity sets several
 *    bits, gaps/missing set all bits (treeator.cpp:463-466).
 */
#ifndef PHYLO_HIP_H
#define PHYLO_HIP_H

#include <stdint.h>

#ifdef __cplusplus
extern "C" {
#endif

typedef struct phylo_ctx phylo_ctx;

enum {
    PHYLO_OK = 0,
    PHYLO_ERR_ARG = -1,         /* bad argument (size, index, value) */
    PHYLO_ERR_STATE = -2,       /* a required set_* call is missing */
    PHYLO_ERR_HIP = -3,         /* HIP runtime error or no GPU */
    PHYLO_ERR_DATA = -4,        /* a tip has no state at some site (check_nodes, simpleML.cpp:52-71) */
    PHYLO_ERR_UNSUPPORTED = -5  /* combination not implemented */
};

enum { PHYLO_PREC_FP32 = 0, PHYLO_PREC_FP64 = 1 };

/* P(t) builders.  REFCOMPAT is the reference's (I + Q t / 2^20)^(2^20)
 * (marth.cpp:129-140 via sub_model.cpp:115-127); EIGEN is V exp(L t) V^-1
 * for reversible models; PADE is scaling-and-squaring Pade for any Q. */
enum { PHYLO_P_REFCOMPAT = 0, PHYLO_P_EIGEN = 1, PHYLO_P_PADE = 2 };

/* Library / context life cycle.  Replaces: construction of simpleML and its
 * CLV store (simpleML.h:33-40, init_nodes simpleML.cpp:24-31). */
int phylo_create(int device, phylo_ctx **out);
void phylo_destroy(phylo_ctx *ctx);
const char *phylo_last_error(const phylo_ctx *ctx);
int phylo_version(void);

/* PHYLO_PREC_FP32 (default): partials stored in FP32 with per-site
 * power-of-two scaling, FP64 arithmetic; PHYLO_PREC_FP64: partials stored in FP64. */
int phylo_set_precision(phylo_ctx *ctx, int precision);

/* Tip states for all taxa and sites.  masks: [n_taxa][n_sites] row-major.
 * Replaces: simpleML::set_char (simpleML.h:41-57), one call per (taxon, site). */
int phylo_set_tips(phylo_ctx *ctx, int n_taxa, int n_sites, int n_states, const uint32_t *masks);

/* Optional per-site weights (site-pattern counts); NULL restores weight 1.
 * The reference's MLtree.h:44-47 sketches the same pattern compression. */
int phylo_set_site_weights(phylo_ctx *ctx, const double *weights);

/* Rooted topology.  Replaces: the tree that simpleML inherits
 * (tree_file_parser, tree.cpp:69-237) as the pruning order. */
int phylo_set_topology(phylo_ctx *ctx, int n_nodes, int root, const int32_t *left, const int32_t *right);

/* Effective length of the edge above each node (n_nodes values; the root's
 * is ignored), already transformed by the host rate model.  Replaces:
 * rate_model::get_time_mod_branch_length x get_rate_clade
 * (rate_model.cpp:103-127, rate_model.h:64-67, applied at simpleML.cpp:173,178). */
int phylo_set_edge_lengths(phylo_ctx *ctx, const double *eff_len);

/* Substitution model.  Q: n_states x n_states row-major rate matrix as
 * sub_model::calc_Q_matrix builds it (sub_model.cpp:99-114).  root_weights:
 * per-state weights of the root sum -- pi for time-reversible models, all ones
 * for the reference's non-reversible models (simpleML.h:132-133).  pi: the
 * stationary frequencies, required by PHYLO_P_EIGEN (may be NULL otherwise).
 * Replaces: sub_model::set_P_matrix / get_P_value (sub_model.cpp:115-127,
 * sub_model.h:46-50) and the root weighting (simpleML.h:123-141). */
int phylo_set_model(phylo_ctx *ctx, int n_states, const double *Q, const double *root_weights,
                    const double *pi, int p_builder);

/* Rate categories: branch-length multipliers and mixture weights (1 category
 * = the reference's single-rate model).  New in this library: the reference
 * has no rate heterogeneity. */
int phylo_set_categories(phylo_ctx *ctx, int n_cat, const double *rates, const double *weights);

/* Whether phylo_loglik stores every internal node's partial (keep != 0) or
 * only those its own walk reads back (keep == 0, the default: 11 of the 127
 * of a 128-taxon Yule tree).  Results are identical either way;
 * phylo_node_partials and phylo_nni_score run one storing evaluation first
 * when the last one kept only part.  The reference always keeps every
 * partial in simpleML's likelihoods map (simpleML.h:105). */
int phylo_set_keep_partials(phylo_ctx *ctx, int keep);

/* Whether a whole-tree DNA evaluation whose 64-site blocks do not divide
 * evenly among the GPU's CUs walks the blocks past the last full round (the
 * tail) as subtree pieces beside the other blocks plus a short second launch
 * for the nodes above the cut (enable != 0, the default), or as whole-tree
 * workgroups.  Values differ only by the FP32 rounding of the extra stored
 * partials of the tail sites (FP64 storage: bit-identical).  No counterpart
 * in the reference (a launch-shape choice); kept switchable for tests. */
int phylo_set_tail_split(phylo_ctx *ctx, int enable);

/* One full postorder evaluation.  site_lnl (nullable): n_sites values.
 * total: weighted sum of the site log-likelihoods (compensated FP64 sum).
 * Replaces: simpleML::calculate_log_likelihood(sub_model&[, rate_model&])
 * (simpleML.h:58-90), calculate_likelihood (simpleML.cpp:127-184),
 * branch_likelihood (:186-197), get_log_likelihood_from_root (simpleML.h:123-141). */
int phylo_loglik(phylo_ctx *ctx, double *site_lnl, double *total);

/* Partials of one internal node for the current inputs (those of the last
 * phylo_loglik when nothing changed since; evaluated first if needed): clv
 * [n_sites][n_cat][n_states] (scaled), scale [n_sites] base-2 exponents so
 * that the true partial is clv * 2^scale.  Replaces: reading simpleML's
 * likelihoods map in draw_normalized_likelihood_on_nodes (simpleML.cpp:199-231). */
int phylo_node_partials(phylo_ctx *ctx, int node, double *clv, int32_t *scale);

/* P matrices the library built for the edge above `node`: [n_cat][S][S].
 * Replaces: sub_model::get_P_value after set_P_matrix (sub_model.h:46-50). */
int phylo_get_pmatrix(phylo_ctx *ctx, int node, double *P);

/* Scores the NNI neighbourhood of the current topology.  mode 0: the trees
 * treebender --nni all emits after the original (internal branches 2..n-3,
 * left then right swap, 2*(n-4) variants; treebender.cpp:897-915 with
 * tree::nni / tree::interchange, tree.cpp:2480-2548); mode 1: the full
 * unrooted neighbourhood, 2*(n-3) variants.  Each variant recomputes only the
 * two re-attached nodes and the path to the root against the stored partials
 * of the current tree (evaluated first if stale), all variants in one launch.
 * Branch lengths move with their subtrees, as in interchange.  lnl: one total
 * per variant (capacity max_variants); node/side (nullable): the interchanged
 * node and side (0 left, 1 right); *n_variants: the count; base_lnl
 * (nullable): the unchanged tree's total.  Replaces: scoring each
 * treebender --nni tree from scratch through treeator (treeator.cpp:705-990). */
int phylo_nni_score(phylo_ctx *ctx, int mode, int max_variants, double *lnl, int32_t *node, int32_t *side,
                    int *n_variants, double *base_lnl);

/* Forward simulation on the GPU, replacing simpleML::simulate_chars
 * (simpleML.cpp:233-292): one thread per site draws a rate category from the
 * category weights, the root state from root_freq (n_states values) and then
 * every node's state from row P[parent state] of its edge's P(t) for that
 * category, with the reference's subtract-until-non-positive inverse CDF.
 * Uses the context's topology, edge lengths, model / P builder and
 * categories (phylo_set_tips fixes the numbering; its tip data is unused).
 * states: n_nodes x n_sites bytes, [node][site].  Draws come from a
 * counter-based generator keyed by (seed, site, node), so results do not
 * depend on the launch shape; the reference draws from rand() seeded by the
 * clock and cannot be reproduced. */
int phylo_simulate(phylo_ctx *ctx, int n_sites, const double *root_freq, uint64_t seed, uint8_t *states);

/* Number of partial-recompute ops the last phylo_nni_score ran (updates
 * per site and category), for throughput accounting. */
int phylo_nni_ops(phylo_ctx *ctx, long *n_ops);

/* Waits for all work on the context's device (hipDeviceSynchronize). */
int phylo_synchronize(phylo_ctx *ctx);

/* Device time (ms, HIP events on the library's stream) of the pruning
 * kernel in the last phylo_loglik / phylo_nni_score call. */
int phylo_last_kernel_ms(phylo_ctx *ctx, double *ms);

/* One whole-tree evaluation that allocates nothing: no hipMalloc, hipFree,
 * pinned-host or other device allocation on any call.  No counterpart in the
 * reference: it is for callers that cannot allocate device memory (code under
 * HIP graph capture, tight serving loops, harnesses that forbid it).  The
 * inputs are those of phylo_set_tips, phylo_set_topology,
 * phylo_set_edge_lengths, phylo_set_model (PHYLO_P_EIGEN, so Q must be
 * reversible with respect to pi) and phylo_set_categories, for the DNA walker
 * (1 to 4 states; 1, 2, 4 or 8 categories) with FP32 partials; the outputs are
 * those of phylo_loglik.  The result equals phylo_loglik's on the same inputs
 * with phylo_set_tail_split(ctx, 0) bit for bit.  The context supplies only
 * the device, its stream and the error text (phylo_last_error); its own
 * buffers are not touched.  All device memory comes from one 64 MiB array per
 * GPU, held from the first call on that GPU until the process ends; a problem
 * that needs more (tips, the partials the walk stores, P, sums: about 1 KiB
 * per site at 128 taxa and 4 categories) is rejected with PHYLO_ERR_ARG.
 * Every call on a GPU uses the same array, so calls that overlap on one GPU
 * (from several threads or contexts) race; calls from one thread do not,
 * since each returns only after its work is done. */
int phylo_loglik_noalloc(phylo_ctx *ctx, int n_taxa, int n_sites, int n_states, const uint32_t *masks, int n_nodes,
                         int root, const int32_t *left, const int32_t *right, const double *eff_len, const double *Q,
                         const double *root_weights, const double *pi, int n_cat, const double *cat_rates,
                         const double *cat_weights, double *site_lnl, double *total);

/* ---- Batched scoring ------------------------------------------------------
 * Many independent DNA problems (S <= 4) of uneven taxon and site counts,
 * each optionally with its whole NNI neighbourhood, scored by two persistent
 * work-queue launches (full evaluations, then partial-recompute variants)
 * captured once into a HIP graph and replayed by phylo_batch_run.  All
 * problems share the number of rate categories and the precision.
 * Replaces: running treeator once per tree file / per treebender --nni tree
 * (treeator.cpp:705-990), which has no batching. */
typedef struct phylo_batch phylo_batch;

int phylo_batch_create(int device, int n_cat, int precision, phylo_batch **out);
const char *phylo_batch_last_error(const phylo_batch *b);

/* Adds one problem (copied): tip masks [n_taxa][n_sites], topology as in
 * phylo_set_topology, effective edge lengths, a reversible Q with its pi
 * (P(t) from the eigen-system), root weights, category rates and weights,
 * nni_mode -1 (none), 0 or 1 (as phylo_nni_score).  *index: its position. */
int phylo_batch_add(phylo_batch *b, int n_taxa, int n_sites, int n_states, const uint32_t *masks, int n_nodes,
                    int root, const int32_t *left, const int32_t *right, const double *eff_len, const double *Q,
                    const double *root_weights, const double *pi, const double *cat_rates,
                    const double *cat_weights, int nni_mode, int *index);

/* The tiling finalize uses without autotuning: sites per work item (a
 * multiple of 64, at most 4096; default 128) and rate-category waves per
 * workgroup (n_cat or 1; default n_cat).  Call before phylo_batch_finalize. */
int phylo_batch_set_tiling(phylo_batch *b, int sites_per_item, int category_waves);

/* Uploads everything, builds the work lists and captures the graph.  With
 * autotune != 0 it first times 64, 128, 256 and 512 sites per work item with
 * one rate category per wave, then one wave walking every category at the
 * best size and twice it, then up to three fewer resident workgroups per CU,
 * and keeps the fastest (phylo_batch_tiling). */
int phylo_batch_finalize(phylo_batch *b, int autotune);

/* Replays the graph.  totals: one lnL per problem; nni_lnl (nullable): every
 * problem's variants, concatenated in phylo_batch_nni_layout order. */
int phylo_batch_run(phylo_batch *b, double *totals, double *nni_lnl);

/* offsets [n_problems + 1] into the variant list; node/side per variant. */
int phylo_batch_nni_layout(phylo_batch *b, int32_t *offsets, int32_t *node, int32_t *side);

/* Updates (site x node x category) per run in each phase, device time of the
 * last run (ms), the chosen sites per work item and the variant count. */
int phylo_batch_stats(phylo_batch *b, double *updates_full, double *updates_nni, double *last_ms,
                      int *sites_per_item, int *n_variants);

/* The tiling the autotuner chose: sites per work item, rate-category waves
 * per workgroup (n_cat: one category per wave; 1: one wave walks every
 * category in turn) and resident workgroups per CU of the persistent grid. */
int phylo_batch_tiling(phylo_batch *b, int *sites_per_item, int *category_waves, int *workgroups_per_cu);

void phylo_batch_destroy(phylo_batch *b);

#ifdef __cplusplus
}
#endif

#endif /* PHYLO_HIP_H */
