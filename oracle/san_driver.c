/*
 * san_driver.c -- TEST INFRASTRUCTURE ONLY.
 *
 * Runs every entry point of the oracle (phylo_oracle.c) once or more on
 * small inputs so that a build with AddressSanitizer and
 * UndefinedBehaviorSanitizer (oracle/Makefile, target "sanitize") checks
 * them for memory and undefined-behaviour errors.  Usage:
 *   san_driver TREE.nwk ALIGNMENT.fst
 * Prints a few values; exits non-zero on a failed call.
 */
#include <math.h>
#include <stdint.h>
#include <stdio.h>
#include <stdlib.h>
#include <string.h>

typedef struct orc_tree orc_tree;
orc_tree *orc_tree_parse(const char *s);
void orc_tree_free(orc_tree *t);
int orc_tree_n_nodes(const orc_tree *t);
int orc_tree_root(const orc_tree *t);
void orc_tree_arrays(const orc_tree *t, int *left, int *right, int *parent, double *blen);
const char *orc_tree_label(const orc_tree *t, int i);
int orc_tree_newick(const orc_tree *t, char *out, int cap);
int orc_parse_fasta(const char *text, int alphabet, int set_empty, char *names_out, int names_cap,
                    uint32_t *masks_out, int max_sites, int *row_len_out, int *n_sites_out);
int orc_n_states(const uint32_t *masks, long n);
void orc_build_Q(int S, int tr, const int *spec, const double *rates, const double *freqs_in, double *Q);
int orc_pmatrix(int S, const double *Q, const double *pi, double t, int builder, double *P);
double orc_time_mod_length(double branch_length, float distance_to_root, int n_cut, const float *cut,
                           const float *cut_rate);
void orc_effective_lengths(int n_nodes, int root, const int *left, const int *right, const double *blen, int n_cut,
                           const float *cut, const float *cut_rate, const int *clade_of_node, int n_clade,
                           const float *clade_rate, int root_quirk, double *eff);
void orc_gamma_rates(double alpha, int C, double *rates);
int orc_loglik(int n_nodes, int root, const int *left, const int *right, const double *eff, int m, int S,
               const uint32_t *masks, int C, const double *cat_rate, const double *cat_w, const double *Q,
               const double *pi, int builder, int per_site_p, const double *root_w, double *site_lnl,
               double *total);
int orc_node_partials(int n_nodes, int root, const int *left, const int *right, const double *eff, int m, int S,
                      const uint32_t *masks, int C, const double *cat_rate, const double *cat_w, const double *Q,
                      const double *pi, int builder, const double *root_w, int reset_sum_per_site, double *out);
int orc_nni_reference(const char *newick, char *out, int cap);
int orc_nni_list(int n_nodes, int root, const int *left, const int *right, int mode, int *out_node, int *out_side);
void orc_nni_apply(int n_nodes, int root, const int *left, const int *right, int node, int side, int *out_left,
                   int *out_right);
int orc_simulate(int n_nodes, int root, const int *left, const int *right, const double *blen, int m, int S,
                 const double *Q, const double *pi, const double *freqs, int C, const double *cat_rate,
                 uint64_t seed, int *states);

static char *slurp(const char *path) {
    FILE *f = fopen(path, "rb");
    if (!f) return NULL;
    fseek(f, 0, SEEK_END);
    long n = ftell(f);
    fseek(f, 0, SEEK_SET);
    char *s = malloc((size_t)n + 1);
    if (fread(s, 1, (size_t)n, f) != (size_t)n) { fclose(f); free(s); return NULL; }
    s[n] = 0;
    fclose(f);
    return s;
}

#define CHECK(x) do { if (!(x)) { fprintf(stderr, "failed: %s (line %d)\n", #x, __LINE__); return 1; } } while (0)

int main(int argc, char **argv) {
    if (argc < 3) return 2;
    char *nwk = slurp(argv[1]), *fst = slurp(argv[2]);
    CHECK(nwk && fst);

    /* parsing and printing */
    orc_tree *t = orc_tree_parse(nwk);
    CHECK(t);
    const int n = orc_tree_n_nodes(t), root = orc_tree_root(t);
    int *left = malloc(sizeof(int) * n), *right = malloc(sizeof(int) * n), *parent = malloc(sizeof(int) * n);
    double *blen = malloc(sizeof(double) * n);
    orc_tree_arrays(t, left, right, parent, blen);
    char *buf = malloc(1 << 20);
    CHECK(orc_tree_newick(t, buf, 1 << 20) > 0);
    CHECK(strlen(orc_tree_label(t, 0)) < 1000);
    const int n_nni = orc_nni_reference(nwk, buf, 1 << 20);
    printf("nodes %d root %d nni_trees %d\n", n, root, n_nni);

    /* fasta */
    const int max_sites = 5000, names_cap = 1 << 16;
    char *names = malloc(names_cap);
    uint32_t *fmasks = calloc((size_t)512 * max_sites, sizeof(uint32_t));
    int *row_len = calloc(512, sizeof(int)), n_sites = 0;
    const int n_rows = orc_parse_fasta(fst, 0, 1, names, names_cap, fmasks, max_sites, row_len, &n_sites);
    CHECK(n_rows > 0 && n_sites > 0);
    printf("fasta rows %d sites %d states %d\n", n_rows, n_sites, orc_n_states(fmasks, (long)n_rows * max_sites));

    /* a balanced tree over tips 0..7 and random data, 4 categories */
    enum { NT = 8, NN = 15, M = 200, S = 4, C = 4 };
    int L[NN], R[NN];
    double bl[NN];
    for (int v = 0; v < NN; ++v) { L[v] = R[v] = -1; bl[v] = 0.02 + 0.03 * (v % 5); }
    for (int k = 0; k < NT - 1; ++k) { L[NT + k] = 2 * k < NT ? 2 * k : NT + 2 * k - NT; R[NT + k] = L[NT + k] + 1; }
    const int rt = NN - 1;
    bl[rt] = 0.0;
    const int spec[6] = {0, 1, 2, 3, 4, 5};
    const double rates[6] = {1.2, 0.4, 0.7, 0.9, 2.1, 1.0}, freqs[4] = {0.3, 0.2, 0.25, 0.25};
    double Q[16], P[16], cr[C], cw[C] = {0.25, 0.25, 0.25, 0.25};
    orc_build_Q(S, 1, spec, rates, freqs, Q);
    const double pi[4] = {0.3, 0.2, 0.25, 0.25};
    for (int b = 0; b < 3; ++b) CHECK(orc_pmatrix(S, Q, pi, 0.37, b, P) == 0);
    orc_gamma_rates(0.6, C, cr);
    uint32_t *masks = calloc((size_t)NN * M, sizeof(uint32_t));
    uint64_t x = 12345;
    for (int v = 0; v < NT; ++v)
        for (int s = 0; s < M; ++s) {
            x = x * 6364136223846793005ull + 1442695040888963407ull;
            const int r = (int)(x >> 60);
            masks[(size_t)v * M + s] = r < 14 ? (1u << (r & 3)) : (r == 14 ? 0xFu : 0x5u);
        }
    double site[M], total = 0.0;
    for (int b = 0; b < 3; ++b) {
        CHECK(orc_loglik(NN, rt, L, R, bl, M, S, masks, C, cr, cw, Q, pi, b, 1, pi, site, &total) == 0);
        CHECK(isfinite(total));
    }
    printf("loglik %.6f\n", total);
    double *part = malloc(sizeof(double) * (size_t)NN * M * S);
    CHECK(orc_node_partials(NN, rt, L, R, bl, M, S, masks, C, cr, cw, Q, pi, 1, pi, 0, part) == 0);

    /* rate models */
    const float cut[2] = {0.03f, 0.08f}, cut_rate[2] = {2.0f, 0.5f}, clade_rate[2] = {1.0f, 3.0f};
    int clade_of_node[NN];
    for (int v = 0; v < NN; ++v) clade_of_node[v] = v == NT + 2 ? 1 : -1;
    double eff[NN];
    orc_effective_lengths(NN, rt, L, R, bl, 2, cut, cut_rate, clade_of_node, 2, clade_rate, 1, eff);
    CHECK(isfinite(orc_time_mod_length(0.05, 0.01f, 2, cut, cut_rate)));

    /* NNI and simulation */
    int nodes[64], sides[64], l2[NN], r2[NN];
    for (int mode = 0; mode < 2; ++mode) {
        const int k = orc_nni_list(NN, rt, L, R, mode, nodes, sides);
        for (int i = 0; i < k; ++i) orc_nni_apply(NN, rt, L, R, nodes[i], sides[i], l2, r2);
        printf("nni mode %d variants %d\n", mode, k);
    }
    int *states = malloc(sizeof(int) * (size_t)NN * M);
    CHECK(orc_simulate(NN, rt, L, R, bl, M, S, Q, pi, pi, C, cr, 7, states) == 0);

    free(states); free(part); free(masks); free(names); free(fmasks); free(row_len); free(buf);
    free(left); free(right); free(parent); free(blen);
    orc_tree_free(t);
    free(nwk); free(fst);
    printf("ok\n");
    return 0;
}
